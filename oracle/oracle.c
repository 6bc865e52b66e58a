/*
 * oracle.c -- FP64 CPU restatement of cpolyg's implicit DG path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the reference's
 * formulas, not its code: quadrature is factored (one state evaluation per
 * quadrature point instead of one per test/trial pair), but every sum, sign,
 * index and ordering rule of the reference is kept, including its quirks:
 *   - periodic traces use the truncated, mis-indexed Legendre series of
 *     src/BoundaryConditions.cpp:18 (periodic_trace = 0),
 *   - off-diagonal Euler Jacobian blocks use the row cell's state and the
 *     row cell's wave-speed derivative (src/Equation.cpp:84,109; jacobian = 0),
 *   - "BlockILU0" is block DILU (src/Preconditioners.cpp:49-123),
 *   - GMRES is left-preconditioned classical Gram-Schmidt with Givens
 *     rotations and a tolerance relative to ||P^-1 b|| (src/BlockMatrix.cpp:79-163).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXB 15   /* basis size up to p = 4 */
#define MAXNC 4

/* Host threads of the per-cell loops, the BSR matvec, the level-scheduled block DILU and the
 * GMRES reductions.  1 (the default) runs every loop in its serial order, which is what the
 * tests pin; more threads are for timing the CPU baseline (same arithmetic per cell, row and
 * block; the DILU factor's Schur sums and the dot products then add in another order). */
static int g_threads = 1;

void o_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* ---------------------------------------------------------------------------
 * Quadrature tables, verbatim values of src/Quadrature.cpp:103-171 (the 14-digit
 * triangle rules bias integrals at the 1e-14 level; parity needs the same digits).
 * Triangle rows: r, s, weight (weights sum to 1; multiplied by the area).
 * Line rows: node on [-1,1], weight (sum 2; multiplied by length/2).
 */
static const double LIN2[] = {-0.5773502691896257, 1.0, 0.5773502691896257, 1.0};
static const double LIN4[] = {-0.3399810435848563, 0.6521451548625461,
                              0.3399810435848563, 0.6521451548625461,
                              -0.8611363115940526, 0.3478548451374538,
                              0.8611363115940526, 0.3478548451374538};
static const double LIN6[] = {0.6612093864662645, 0.3607615730481386,
                              -0.6612093864662645, 0.3607615730481386,
                              -0.2386191860831969, 0.4679139345726910,
                              0.2386191860831969, 0.4679139345726910,
                              -0.9324695142031521, 0.1713244923791704,
                              0.9324695142031521, 0.1713244923791704};
static const double LIN10[] = {0.1488743389816312, 0.2955242247147529,
                               -0.1488743389816312, 0.2955242247147529,
                               -0.4333953941292472, 0.2692667193099963,
                               0.4333953941292472, 0.2692667193099963,
                               -0.6794095682990244, 0.2190863625159820,
                               0.6794095682990244, 0.2190863625159820,
                               -0.8650633666889845, 0.1494513491505806,
                               0.8650633666889845, 0.1494513491505806,
                               -0.9739065285171717, 0.0666713443086881,
                               0.9739065285171717, 0.0666713443086881};
static const double TRI2[] = {0.16666666666667, 0.16666666666667, 0.33333333333333,
                              0.16666666666667, 0.66666666666667, 0.33333333333333,
                              0.66666666666667, 0.16666666666667, 0.33333333333333};
static const double TRI4[] = {0.44594849091597, 0.44594849091597, 0.22338158967801,
                              0.44594849091597, 0.10810301816807, 0.22338158967801,
                              0.10810301816807, 0.44594849091597, 0.22338158967801,
                              0.09157621350977, 0.09157621350977, 0.10995174365532,
                              0.09157621350977, 0.81684757298046, 0.10995174365532,
                              0.81684757298046, 0.09157621350977, 0.10995174365532};
static const double TRI6[] = {0.24928674517091, 0.24928674517091, 0.11678627572638,
                              0.24928674517091, 0.50142650965818, 0.11678627572638,
                              0.50142650965818, 0.24928674517091, 0.11678627572638,
                              0.06308901449150, 0.06308901449150, 0.05084490637021,
                              0.06308901449150, 0.87382197101700, 0.05084490637021,
                              0.87382197101700, 0.06308901449150, 0.05084490637021,
                              0.31035245103378, 0.63650249912140, 0.08285107561837,
                              0.63650249912140, 0.05314504984482, 0.08285107561837,
                              0.05314504984482, 0.31035245103378, 0.08285107561837,
                              0.63650249912140, 0.31035245103378, 0.08285107561837,
                              0.31035245103378, 0.05314504984482, 0.08285107561837,
                              0.05314504984482, 0.63650249912140, 0.08285107561837};
static const double TRI10[] = {0.33333333333333, 0.33333333333333, 0.09081799038275,
                               0.48557763338366, 0.48557763338366, 0.03672595775647,
                               0.48557763338366, 0.02884473323269, 0.03672595775647,
                               0.02884473323269, 0.48557763338366, 0.03672595775647,
                               0.10948157548504, 0.10948157548504, 0.04532105943553,
                               0.10948157548504, 0.78103684902993, 0.04532105943553,
                               0.78103684902993, 0.10948157548504, 0.04532105943553,
                               0.30793983876412, 0.55035294182100, 0.07275791684542,
                               0.55035294182100, 0.14170721941488, 0.07275791684542,
                               0.14170721941488, 0.30793983876412, 0.07275791684542,
                               0.55035294182100, 0.30793983876412, 0.07275791684542,
                               0.30793983876412, 0.14170721941488, 0.07275791684542,
                               0.14170721941488, 0.55035294182100, 0.07275791684542,
                               0.24667256063990, 0.72832390459741, 0.02832724253106,
                               0.72832390459741, 0.02500353476269, 0.02832724253106,
                               0.02500353476269, 0.24667256063990, 0.02832724253106,
                               0.72832390459741, 0.24667256063990, 0.02832724253106,
                               0.24667256063990, 0.02500353476269, 0.02832724253106,
                               0.02500353476269, 0.72832390459741, 0.02832724253106,
                               0.06680325101220, 0.92365593358750, 0.00942166696373,
                               0.92365593358750, 0.00954081540030, 0.00942166696373,
                               0.00954081540030, 0.06680325101220, 0.00942166696373,
                               0.92365593358750, 0.06680325101220, 0.00942166696373,
                               0.06680325101220, 0.00954081540030, 0.00942166696373,
                               0.00954081540030, 0.92365593358750, 0.00942166696373};

/* rule choice of Quadrature::polygonIntegral / lineIntegral (Quadrature.cpp:20-27,74-81) */
static void pick_rules(int deg2, const double **tri, int *nt, const double **lin, int *nl)
{
    if (deg2 <= 2) { *tri = TRI2; *nt = 3; *lin = LIN2; *nl = 2; }
    else if (deg2 <= 4) { *tri = TRI4; *nt = 6; *lin = LIN4; *nl = 4; }
    else if (deg2 <= 6) { *tri = TRI6; *nt = 12; *lin = LIN6; *nl = 6; }
    else { *tri = TRI10; *nt = 25; *lin = LIN10; *nl = 10; }
}

int o_rule_sizes(int deg2, int *ntri, int *nlin)
{
    const double *t, *l;
    pick_rules(deg2, &t, ntri, &l, nlin);
    return 0;
}

int o_rule_tables(int deg2, double *tri, double *lin)
{
    const double *t, *l;
    int nt, nl;
    pick_rules(deg2, &t, &nt, &l, &nl);
    memcpy(tri, t, sizeof(double) * 3 * nt);
    memcpy(lin, l, sizeof(double) * 2 * nl);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Legendre basis.  Leg2D (src/Legendre.h:51-68): phi_k = P_a(xi) P_b(eta) with
 * k running over a = 0..p (outer) and b = 0..p-a (inner).  Derivatives are the
 * ones LegDerX/LegDerY (src/Legendre.cpp:34-124) produce as Legendre series,
 * evaluated here directly from P'_{n+1} = x P'_n + (n+1) P_n.
 */
void o_legendre(int p, double x, double *P, double *dP)
{
    P[0] = 1.0;
    dP[0] = 0.0;
    if (p >= 1) { P[1] = x; dP[1] = 1.0; }
    for (int n = 1; n < p; n++) {
        P[n + 1] = ((2 * n + 1) * x * P[n] - n * P[n - 1]) / (n + 1);
        dP[n + 1] = x * dP[n] + (n + 1) * P[n];
    }
}

void o_basis(int p, double xi, double eta, double *phi, double *dxi, double *deta)
{
    double Px[MAXB], dPx[MAXB], Py[MAXB], dPy[MAXB];
    o_legendre(p, xi, Px, dPx);
    o_legendre(p, eta, Py, dPy);
    int k = 0;
    for (int a = 0; a <= p; a++)
        for (int b = 0; b <= p - a; b++) {
            phi[k] = Px[a] * Py[b];
            if (dxi) dxi[k] = dPx[a] * Py[b];
            if (deta) deta[k] = Px[a] * dPy[b];
            k++;
        }
}

/* Periodic trace basis.  reference (quirk): Leg2D with m = int(0.5*(sqrt(4 bs+1)-1))
 * over the first m(m+1)/2 coefficients (src/BoundaryConditions.cpp:18-24); exact: the
 * full basis. */
static void trace_basis(int p, int exact, double xi, double eta, double *phi)
{
    int bs = (p + 1) * (p + 2) / 2;
    if (exact) { o_basis(p, xi, eta, phi, NULL, NULL); return; }
    int mq = (int)(0.5 * (sqrt(4.0 * bs + 1.0) - 1.0));
    double Px[MAXB], dPx[MAXB], Py[MAXB], dPy[MAXB];
    for (int k = 0; k < bs; k++) phi[k] = 0.0;
    if (mq <= 0) return;
    o_legendre(mq - 1, xi, Px, dPx);
    o_legendre(mq - 1, eta, Py, dPy);
    int k = 0;
    for (int a = 0; a < mq; a++)
        for (int b = 0; b < mq - a; b++) phi[k++] = Px[a] * Py[b];
}

/* ---------------------------------------------------------------------------
 * Geometry (src/PolyMesh.h:44-49, src/PolyMesh.cpp:248-273, 307-325)
 */
static void local_coords(const o_mesh *m, int i, double x, double y, double *xi, double *eta)
{
    const double *bb = m->bb + 4 * i;
    *xi = 2.0 * (x - bb[0]) / bb[2] - 1.0;
    *eta = 2.0 * (y - bb[1]) / bb[3] - 1.0;
}

void o_outward_normal(const o_mesh *m, int i, int a, int b, double *nx, double *ny)
{
    int nv = m->pptr[i + 1] - m->pptr[i];
    double cx = 0, cy = 0;
    for (int j = 0; j < nv; j++) {
        cx += m->v[2 * m->pv[m->pptr[i] + j]];
        cy += m->v[2 * m->pv[m->pptr[i] + j] + 1];
    }
    cx /= nv;
    cy /= nv;
    double dx = m->v[2 * a] - m->v[2 * b];
    double dy = m->v[2 * a + 1] - m->v[2 * b + 1];
    double len = sqrt(dx * dx + dy * dy);
    double sv = -dy * (m->v[2 * a] - cx) + dx * (m->v[2 * a + 1] - cy);
    int s = (0.0 < sv) - (sv < 0.0);
    *nx = -s * dy / len;
    *ny = s * dx / len;
}

static void periodic_coords(const o_mesh *m, int id, double x, double y, double *xi, double *eta)
{
    const int *row = m->bc + 7 * (-1 - id);
    int a1 = row[3], a2 = row[5], p2 = row[2];
    double xn = x + (m->v[2 * a2] - m->v[2 * a1]);
    double yn = y + (m->v[2 * a2 + 1] - m->v[2 * a1 + 1]);
    local_coords(m, p2, xn, yn, xi, eta);
}

static double tri_area(const double *t)
{
    return 0.5 * fabs(t[0] * t[3] + t[2] * t[5] + t[4] * t[1] - t[0] * t[5] - t[4] * t[3] - t[2] * t[1]);
}

double o_polygon_integral_monomial(const o_mesh *m, int i, int deg2, int px, int py)
{
    const double *tr, *ln;
    int nt, nl;
    pick_rules(deg2, &tr, &nt, &ln, &nl);
    double integ = 0;
    for (int t = m->tptr[i]; t < m->tptr[i + 1]; t++) {
        const double *T = m->txy + 6 * t;
        double area = tri_area(T), z = 0;
        for (int q = 0; q < nt; q++) {
            double r = tr[3 * q], s = tr[3 * q + 1];
            double x = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
            double y = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
            z += pow(x, px) * pow(y, py) * tr[3 * q + 2];
        }
        integ += z * area;
    }
    return integ;
}

/* ---------------------------------------------------------------------------
 * Equations (eqns/Advection.cpp:8-66, eqns/Euler.cpp:7-179)
 */
int o_nc(const o_eq *eq) { return eq->kind == O_EULER ? 4 : 1; }

static void adv_beta(const o_eq *eq, double x, double y, double *bx, double *by)
{
    *bx = eq->beta[0] + eq->beta[1] * x + eq->beta[2] * y;
    *by = eq->beta[3] + eq->beta[4] * x + eq->beta[5] * y;
}

static void euler_flux(double g, const double *U, double *Fx, double *Fy, double *c, double *u, double *v)
{
    double rho = U[0], ru = U[1], rv = U[2], rE = U[3];
    *u = ru / rho;
    *v = rv / rho;
    double P = (g - 1) * (rE - 0.5 * (ru * ru + rv * rv) / rho);
    Fx[0] = ru; Fx[1] = ru * (*u) + P; Fx[2] = ru * (*v); Fx[3] = (rE + P) * (*u);
    Fy[0] = rv; Fy[1] = ru * (*v); Fy[2] = rv * (*v) + P; Fy[3] = (rE + P) * (*v);
    *c = sqrt(g * P / rho);
}

/* flux Jacobians, row-major 4x4 (Euler::fluxJacobian, Euler.cpp:134-165).  The
 * reference's x-Jacobian has H - (gamma-1)|u|^2 in the energy row, momentum-x column
 * (Euler.cpp:157); the derivative is H - (gamma-1) u^2.  ref != 0 reproduces it. */
static void euler_fluxjac(double g, const double *U, double *Ax, double *Ay, int ref)
{
    double rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3] / rho;
    double usq = u * u + v * v, e = E - 0.5 * usq, P = (g - 1) * rho * e, H = E + P / rho;
    double ax[16] = {0, 1, 0, 0,
                     -u * u + 0.5 * (g - 1) * usq, (3 - g) * u, -(g - 1) * v, g - 1,
                     -u * v, v, u, 0,
                     u * (0.5 * (g - 1) * usq - H), ref ? H - (g - 1) * usq : H - (g - 1) * u * u,
                     -(g - 1) * u * v, g * u};
    double ay[16] = {0, 0, 1, 0,
                     -u * v, v, u, 0,
                     -v * v + 0.5 * (g - 1) * usq, -(g - 1) * u, (3 - g) * v, g - 1,
                     v * (0.5 * (g - 1) * usq - H), -(g - 1) * u * v, H - (g - 1) * v * v, g * v};
    memcpy(Ax, ax, sizeof ax);
    memcpy(Ay, ay, sizeof ay);
}

/* Euler::alphaDerivative (Euler.cpp:35-98): derivative of the LF speed w.r.t. U1,
 * nonzero only when the first maximal |eigenvalue| belongs to U1. */
static void euler_alpha_deriv(double g, const double *U1, const double *U2, double nx, double ny,
                              double *alpha, double *d)
{
    double u1 = U1[1] / U1[0], v1 = U1[2] / U1[0], u2 = U2[1] / U2[0], v2 = U2[2] / U2[0];
    double P1 = (g - 1) * (U1[3] - 0.5 * (U1[1] * U1[1] + U1[2] * U1[2]) / U1[0]);
    double c1 = sqrt(g * P1 / U1[0]);
    double P2 = (g - 1) * (U2[3] - 0.5 * (U2[1] * U2[1] + U2[2] * U2[2]) / U2[0]);
    double c2 = sqrt(g * P2 / U2[0]);
    double vn1 = u1 * nx + v1 * ny, vn2 = u2 * nx + v2 * ny;
    double eig[6] = {vn1 - c1, vn1, vn1 + c1, vn2 - c2, vn2, vn2 + c2};
    int idx = 0;
    double best = fabs(eig[0]);
    for (int k = 1; k < 6; k++)
        if (fabs(eig[k]) > best) { best = fabs(eig[k]); idx = k; }
    *alpha = best;
    d[0] = d[1] = d[2] = d[3] = 0;
    if (idx <= 2) {
        double rho = U1[0], E = U1[3] / rho;
        d[0] = -u1 / rho * nx - v1 / rho * ny;
        d[1] = 1 / rho * nx;
        d[2] = 1 / rho * ny;
        double usq = u1 * u1 + v1 * v1;
        double a = sqrt(g * g - g), b = rho * sqrt(4 * E - 2 * usq);
        if (idx == 0) {
            d[0] -= a * (usq - E) / b; d[1] -= -a * u1 / b; d[2] -= -a * v1 / b; d[3] -= a / b;
        } else if (idx == 2) {
            d[0] += a * (usq - E) / b; d[1] += -a * u1 / b; d[2] += -a * v1 / b; d[3] += a / b;
        }
        if (eig[idx] < 0) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; d[3] = -d[3]; }
    }
}

/* Derivative of the residual's LF speed w.r.t. the OUTER state U2.  The speed is the
 * first maximal |eigenvalue| of the list (U1's three, then U2's three) -- the order in
 * which Euler::numericalFlux (Euler.cpp:100-130) takes std::max -- so that a NaN speed of
 * an unphysical outer state is skipped here exactly as in the residual.  Nonzero only when
 * the maximum belongs to U2.  (Not in the reference: its off-diagonal Jacobian is the
 * row-cell quirk; this is the derivative used by the exact Jacobian option.) */
static void euler_alpha_deriv_outer(double g, const double *U1, const double *U2, double nx, double ny,
                                    double *alpha, double *d)
{
    double u1 = U1[1] / U1[0], v1 = U1[2] / U1[0], u2 = U2[1] / U2[0], v2 = U2[2] / U2[0];
    double P1 = (g - 1) * (U1[3] - 0.5 * (U1[1] * U1[1] + U1[2] * U1[2]) / U1[0]);
    double c1 = sqrt(g * P1 / U1[0]);
    double P2 = (g - 1) * (U2[3] - 0.5 * (U2[1] * U2[1] + U2[2] * U2[2]) / U2[0]);
    double c2 = sqrt(g * P2 / U2[0]);
    double vn1 = u1 * nx + v1 * ny, vn2 = u2 * nx + v2 * ny;
    double eig[6] = {vn1 - c1, vn1, vn1 + c1, vn2 - c2, vn2, vn2 + c2};
    int idx = 0;
    double best = fabs(eig[0]);
    for (int k = 1; k < 6; k++)
        if (fabs(eig[k]) > best) { best = fabs(eig[k]); idx = k; }
    *alpha = best;
    d[0] = d[1] = d[2] = d[3] = 0;
    if (idx >= 3) {
        double rho = U2[0], E = U2[3] / rho;
        d[0] = -u2 / rho * nx - v2 / rho * ny;
        d[1] = 1 / rho * nx;
        d[2] = 1 / rho * ny;
        double usq = u2 * u2 + v2 * v2;
        double a = sqrt(g * g - g), b = rho * sqrt(4 * E - 2 * usq);
        if (idx == 3) {
            d[0] -= a * (usq - E) / b; d[1] += a * u2 / b; d[2] += a * v2 / b; d[3] -= a / b;
        } else if (idx == 5) {
            d[0] += a * (usq - E) / b; d[1] -= a * u2 / b; d[2] -= a * v2 / b; d[3] += a / b;
        }
        if (eig[idx] < 0) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; d[3] = -d[3]; }
    }
}

static void flux(const o_eq *eq, const double *U, double x, double y, double *Fx, double *Fy)
{
    if (eq->kind == O_EULER) {
        double c, u, v;
        euler_flux(eq->gamma, U, Fx, Fy, &c, &u, &v);
    } else {
        double bx, by;
        adv_beta(eq, x, y, &bx, &by);
        Fx[0] = bx * U[0];
        Fy[0] = by * U[0];
    }
}

static void numflux(const o_eq *eq, const double *Um, const double *Up, double x, double y,
                    double nx, double ny, double *F)
{
    if (eq->kind == O_EULER) {
        double Fxm[4], Fym[4], Fxp[4], Fyp[4], cm, um, vm, cp, up, vp;
        euler_flux(eq->gamma, Um, Fxm, Fym, &cm, &um, &vm);
        euler_flux(eq->gamma, Up, Fxp, Fyp, &cp, &up, &vp);
        double vnm = um * nx + vm * ny, vnp = up * nx + vp * ny;
        double e[6] = {fabs(vnm - cm), fabs(vnm), fabs(vnm + cm), fabs(vnp - cp), fabs(vnp), fabs(vnp + cp)};
        double alpha = e[0];
        for (int k = 1; k < 6; k++) if (e[k] > alpha) alpha = e[k];
        for (int c = 0; c < 4; c++)
            F[c] = 0.5 * ((Fxp[c] + Fxm[c]) * nx + (Fyp[c] + Fym[c]) * ny + alpha * (Um[c] - Up[c]));
    } else {
        double bx, by;
        adv_beta(eq, x, y, &bx, &by);
        double bn = bx * nx + by * ny;
        F[0] = bn > 0 ? Um[0] * bn : Up[0] * bn;
    }
}

static void fluxjac(const o_eq *eq, const double *U, double x, double y, double *Ax, double *Ay)
{
    if (eq->kind == O_EULER) euler_fluxjac(eq->gamma, U, Ax, Ay, eq->jacobian == 0);
    else adv_beta(eq, x, y, Ax, Ay);
}

/* Euler::numericalFluxJacobian (Euler.cpp:168-179): 0.5*(A(U) n + sgn alpha I + (U - U2) alpha'^T),
 * alpha' w.r.t. the first argument. */
static void numfluxjac(const o_eq *eq, const double *U, const double *U2, double x, double y,
                       double nx, double ny, int sgn, double *J)
{
    if (eq->kind == O_EULER) {
        double Ax[16], Ay[16], d[4], alpha;
        euler_fluxjac(eq->gamma, U, Ax, Ay, eq->jacobian == 0);
        euler_alpha_deriv(eq->gamma, U, U2, nx, ny, &alpha, d);
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++)
                J[4 * r + c] = 0.5 * (Ax[4 * r + c] * nx + Ay[4 * r + c] * ny + (r == c ? sgn * alpha : 0.0)
                                      + (U[r] - U2[r]) * d[c]);
    } else {
        double bx, by;
        adv_beta(eq, x, y, &bx, &by);
        double bn = nx * bx + ny * by;
        J[0] = (sgn * bn > 0) ? bn : 0.0;
    }
}

/* Dirichlet functors used by the drivers */
void o_eval_bc(const o_eq *eq, double x, double y, double t, double *out)
{
    int nc = o_nc(eq);
    const double *p = eq->bc_par;
    switch (eq->bc_kind) {
    case O_BC_CONSTANT:
        for (int c = 0; c < nc; c++) out[c] = p[c];
        return;
    case O_BC_ISENTROPIC_VORTEX: {
        /* p: beta, x0, y0, uinf, vinf ; gamma from eq (Yee et al. vortex moving with the flow) */
        double g = eq->gamma, beta = p[0];
        double dx = x - p[1] - p[3] * t, dy = y - p[2] - p[4] * t;
        double r2 = dx * dx + dy * dy;
        double ex = exp(0.5 * (1 - r2));
        double u = p[3] - beta / (2 * M_PI) * dy * ex;
        double v = p[4] + beta / (2 * M_PI) * dx * ex;
        double T = 1 - (g - 1) * beta * beta / (8 * g * M_PI * M_PI) * ex * ex;
        double rho = pow(T, 1 / (g - 1));
        double P = rho * T;
        out[0] = rho; out[1] = rho * u; out[2] = rho * v;
        out[3] = P / (g - 1) + 0.5 * rho * (u * u + v * v);
        return;
    }
    case O_BC_REF_VORTEX: {
        /* VortexSolution (eqns/EulerSetups.cpp:7-67) */
        double g = eq->gamma, x0 = 5, y0 = 5, theta = atan2(1, 2), eps = 0.3, rc = 1.5, Minf = 0.5;
        double uInf = 1.0, rhoInf = 1.0, pInf = 1.0, ub = uInf * cos(theta), vb = uInf * sin(theta);
        double f = (1 - pow((x - x0) - ub * t, 2) - pow((y - y0) - vb * t, 2)) / (rc * rc);
        double u = uInf * (cos(theta) - eps * ((y - y0) - vb * t) / (2 * M_PI * rc) * exp(f * 0.5));
        double v = uInf * (sin(theta) + eps * ((x - x0) - ub * t) / (2 * M_PI * rc) * exp(f * 0.5));
        double base = 1 - eps * eps * (g - 1) * (Minf * Minf / (8 * M_PI * M_PI)) * exp(f);
        double rho = rhoInf * pow(base, 1 / (g - 1));
        double P = pInf * pow(base, 1 / (g - 1));
        out[0] = rho; out[1] = rho * u; out[2] = rho * v;
        out[3] = P / (g - 1) + 0.5 * rho * (u * u + v * v);
        return;
    }
    default:
        for (int c = 0; c < nc; c++) out[c] = 0.0;
    }
}

/* state at a point from coefficients U[c*bs + k] and basis values */
static void eval_state(int bs, int nc, const double *U, const double *phi, double *out)
{
    for (int c = 0; c < nc; c++) {
        double s = 0;
        for (int k = 0; k < bs; k++) s += U[c * bs + k] * phi[k];
        out[c] = s;
    }
}

/* ---------------------------------------------------------------------------
 * Mass matrix (MassMatrix::MassMatrix, src/MassMatrix.cpp:30-94): block-diagonal,
 * M_i[(c,k),(c,l)] = int phi_k phi_l, replicated per component.
 */
int o_mass(const o_mesh *m, int deg, int nc, double *blocks)
{
    int bs = (deg + 1) * (deg + 2) / 2, bl = nc * bs;
    const double *tr, *ln;
    int nt, nl;
    pick_rules(2 * deg, &tr, &nt, &ln, &nl);
    double phi[MAXB], Mk[MAXB * MAXB];
    for (int i = 0; i < m->np; i++) {
        memset(Mk, 0, sizeof Mk);
        for (int t = m->tptr[i]; t < m->tptr[i + 1]; t++) {
            const double *T = m->txy + 6 * t;
            double area = tri_area(T);
            for (int q = 0; q < nt; q++) {
                double r = tr[3 * q], s = tr[3 * q + 1], W = tr[3 * q + 2] * area;
                double x = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
                double y = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
                double xi, eta;
                local_coords(m, i, x, y, &xi, &eta);
                o_basis(deg, xi, eta, phi, NULL, NULL);
                for (int k = 0; k < bs; k++)
                    for (int l = 0; l < bs; l++) Mk[k * bs + l] += W * phi[k] * phi[l];
            }
        }
        double *B = blocks + (size_t)i * bl * bl;
        memset(B, 0, sizeof(double) * bl * bl);
        for (int c = 0; c < nc; c++)
            for (int k = 0; k < bs; k++)
                for (int l = 0; l < bs; l++) B[(c * bs + l) * bl + c * bs + k] = Mk[k * bs + l];
    }
    return 0;
}

/* MeshFn::L2Norm (src/MeshFn.cpp:299-312) */
void o_l2norm(const o_mesh *m, int deg, int nc, const double *u, double *out)
{
    int bs = (deg + 1) * (deg + 2) / 2, bl = nc * bs;
    const double *tr, *ln;
    int nt, nl;
    pick_rules(2 * deg, &tr, &nt, &ln, &nl);
    double acc[MAXNC] = {0, 0, 0, 0}, phi[MAXB], val[MAXNC];
    for (int i = 0; i < m->np; i++) {
        double cell[MAXNC] = {0, 0, 0, 0};
        for (int t = m->tptr[i]; t < m->tptr[i + 1]; t++) {
            const double *T = m->txy + 6 * t;
            double area = tri_area(T), z[MAXNC] = {0, 0, 0, 0};
            for (int q = 0; q < nt; q++) {
                double r = tr[3 * q], s = tr[3 * q + 1];
                double x = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
                double y = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
                double xi, eta;
                local_coords(m, i, x, y, &xi, &eta);
                o_basis(deg, xi, eta, phi, NULL, NULL);
                eval_state(bs, nc, u + (size_t)i * bl, phi, val);
                for (int c = 0; c < nc; c++) z[c] += val[c] * val[c] * tr[3 * q + 2];
            }
            for (int c = 0; c < nc; c++) cell[c] += z[c] * area;
        }
        for (int c = 0; c < nc; c++) acc[c] += cell[c];
    }
    for (int c = 0; c < nc; c++) out[c] = sqrt(acc[c]);
}

/* ---------------------------------------------------------------------------
 * MeshFn::interp (src/MeshFn.cpp:56-145) and MeshFn::L2Error (src/MeshFn.cpp:277-296)
 */

/* the m Gauss-Legendre nodes, ascending (gsl_integration_glfixed_table in the reference):
 * bisection on P_m between the roots of P_{m-1}, which interlace with those of P_m */
void o_gauss_nodes(int m, double *x)
{
    double prev[16], cur[16];
    int np = 0;
    for (int n = 1; n <= m; n++) {
        /* brackets: -1, roots of P_{n-1}, 1 */
        double br[18];
        int nb = 0;
        br[nb++] = -1.0;
        for (int k = 0; k < np; k++) br[nb++] = prev[k];
        br[nb++] = 1.0;
        for (int k = 0; k + 1 < nb; k++) {
            double a = br[k], b = br[k + 1], Pa, dPa[MAXB + 2], Pv[MAXB + 2];
            o_legendre(n, a, Pv, dPa);
            Pa = Pv[n];
            for (int it = 0; it < 200; it++) {
                double c = 0.5 * (a + b), Pc;
                o_legendre(n, c, Pv, dPa);
                Pc = Pv[n];
                if ((Pc < 0) == (Pa < 0)) { a = c; Pa = Pc; } else b = c;
            }
            cur[k] = 0.5 * (a + b);
        }
        np = n;
        for (int k = 0; k < np; k++) prev[k] = cur[k];
    }
    for (int k = 0; k < m; k++) x[k] = prev[k];
}

void o_interp_points(const o_mesh *m, int deg, double *xy)
{
    int M = deg + 1, Q = M * M;
    double nodes[16];
    o_gauss_nodes(M, nodes);
    for (int i = 0; i < m->np; i++) {
        const double *bb = m->bb + 4 * i;
        for (int qi = 0; qi < M; qi++)
            for (int qj = 0; qj < M; qj++) {
                double *o = xy + 2 * ((size_t)i * Q + qi * M + qj);
                o[0] = (nodes[qi] + 1.0) * bb[2] / 2.0 + bb[0];
                o[1] = (nodes[qj] + 1.0) * bb[3] / 2.0 + bb[1];
            }
    }
}

/* least squares of G c = v by Householder QR of G (Q x bs, the basis at the tensor points) */
int o_interp_values(const o_mesh *m, int deg, int nc, const double *vals, double *u)
{
    int M = deg + 1, Q = M * M, bs = (deg + 1) * (deg + 2) / 2;
    double nodes[16], G[25 * MAXB], hv[MAXB][25], beta[MAXB], phi[MAXB];
    o_gauss_nodes(M, nodes);
    for (int qi = 0; qi < M; qi++)
        for (int qj = 0; qj < M; qj++) {
            o_basis(deg, nodes[qi], nodes[qj], phi, NULL, NULL);
            for (int k = 0; k < bs; k++) G[(qi * M + qj) * bs + k] = phi[k];
        }
    /* Householder: column k reflected onto e_k; G becomes R in its upper triangle */
    for (int k = 0; k < bs; k++) {
        double nrm = 0;
        for (int r = k; r < Q; r++) nrm += G[r * bs + k] * G[r * bs + k];
        nrm = sqrt(nrm);
        if (nrm == 0) return 1;
        double alpha = G[k * bs + k] > 0 ? -nrm : nrm, vn = 0;
        for (int r = 0; r < Q; r++) hv[k][r] = r < k ? 0.0 : G[r * bs + k];
        hv[k][k] -= alpha;
        for (int r = k; r < Q; r++) vn += hv[k][r] * hv[k][r];
        beta[k] = vn > 0 ? 2.0 / vn : 0.0;
        for (int c = k; c < bs; c++) {
            double d = 0;
            for (int r = k; r < Q; r++) d += hv[k][r] * G[r * bs + c];
            for (int r = k; r < Q; r++) G[r * bs + c] -= beta[k] * d * hv[k][r];
        }
    }
    int bl = nc * bs;
    for (int i = 0; i < m->np; i++)
        for (int c = 0; c < nc; c++) {
            double v[25];
            for (int q = 0; q < Q; q++) v[q] = vals[((size_t)i * Q + q) * nc + c];
            for (int k = 0; k < bs; k++) {
                double d = 0;
                for (int r = k; r < Q; r++) d += hv[k][r] * v[r];
                for (int r = k; r < Q; r++) v[r] -= beta[k] * d * hv[k][r];
            }
            double *o = u + (size_t)i * bl + c * bs;
            for (int k = bs - 1; k >= 0; k--) {
                double s = v[k];
                for (int l = k + 1; l < bs; l++) s -= G[k * bs + l] * o[l];
                o[k] = s / G[k * bs + k];
            }
        }
    return 0;
}

int o_quad_count(const o_mesh *m, int deg2)
{
    const double *tr, *ln;
    int nt, nl;
    pick_rules(deg2, &tr, &nt, &ln, &nl);
    return m->tptr[m->np] * nt;
}

void o_quad_points(const o_mesh *m, int deg2, double *xyw)
{
    const double *tr, *ln;
    int nt, nl;
    pick_rules(deg2, &tr, &nt, &ln, &nl);
    size_t e = 0;
    for (int t = 0; t < m->tptr[m->np]; t++) {
        const double *T = m->txy + 6 * t;
        double area = tri_area(T);
        for (int q = 0; q < nt; q++, e++) {
            double r = tr[3 * q], s = tr[3 * q + 1];
            xyw[3 * e] = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
            xyw[3 * e + 1] = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
            xyw[3 * e + 2] = tr[3 * q + 2] * area;
        }
    }
}

void o_l2_error_values(const o_mesh *m, int deg, int nc, const double *u, const double *f, double *out)
{
    int bs = (deg + 1) * (deg + 2) / 2, bl = nc * bs;
    const double *tr, *ln;
    int nt, nl;
    pick_rules(2 * deg, &tr, &nt, &ln, &nl);
    double acc[MAXNC] = {0, 0, 0, 0}, phi[MAXB], val[MAXNC];
    size_t e = 0;
    for (int i = 0; i < m->np; i++) {
        double cell[MAXNC] = {0, 0, 0, 0};
        for (int t = m->tptr[i]; t < m->tptr[i + 1]; t++) {
            const double *T = m->txy + 6 * t;
            double area = tri_area(T), z[MAXNC] = {0, 0, 0, 0};
            for (int q = 0; q < nt; q++, e++) {
                double r = tr[3 * q], s = tr[3 * q + 1];
                double x = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
                double y = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
                double xi, eta;
                local_coords(m, i, x, y, &xi, &eta);
                o_basis(deg, xi, eta, phi, NULL, NULL);
                eval_state(bs, nc, u + (size_t)i * bl, phi, val);
                for (int c = 0; c < nc; c++) {
                    double d = val[c] - f[e * nc + c];
                    z[c] += d * d * tr[3 * q + 2];
                }
            }
            for (int c = 0; c < nc; c++) cell[c] += z[c] * area;
        }
        for (int c = 0; c < nc; c++) acc[c] += cell[c];
    }
    for (int c = 0; c < nc; c++) out[c] = sqrt(acc[c]);
}

/* exterior state U+ at a boundary point of cell i across edge with neighbour id nb */
static void outer_state(const o_mesh *m, const o_eq *eq, int deg, const double *u, int nb,
                        double x, double y, double t, double *Up, double *phip)
{
    int bs = (deg + 1) * (deg + 2) / 2, nc = o_nc(eq), bl = nc * bs;
    double xi, eta;
    if (nb >= 0) {
        local_coords(m, nb, x, y, &xi, &eta);
        o_basis(deg, xi, eta, phip, NULL, NULL);
        eval_state(bs, nc, u + (size_t)nb * bl, phip, Up);
    } else if (m->bc[7 * (-1 - nb)]) {
        int p2 = m->bc[7 * (-1 - nb) + 2];
        periodic_coords(m, nb, x, y, &xi, &eta);
        trace_basis(deg, eq->periodic_trace, xi, eta, phip);
        eval_state(bs, nc, u + (size_t)p2 * bl, phip, Up);
    } else {
        o_eval_bc(eq, x, y, t, Up);
        for (int k = 0; k < bs; k++) phip[k] = 0.0;
    }
}

/* ---------------------------------------------------------------------------
 * Residual (Equation::assemble, src/Equation.cpp:152-191 with the integrands of
 * :20-57 and boundaryIntegral :114-145):
 *   b_i[c*bs+k] = sum_vol W (Fx_c dpsi_k/dx + Fy_c dpsi_k/dy) - sum_edges sum_q W psi_k Fhat_c
 */
int o_assemble(const o_mesh *m, const o_eq *eq, int deg, const double *u, double t, double *b)
{
    int bs = (deg + 1) * (deg + 2) / 2, nc = o_nc(eq), bl = nc * bs;
    const double *tr, *ln;
    int nt, nl;
    pick_rules(2 * deg, &tr, &nt, &ln, &nl);
    double phi[MAXB], dxi[MAXB], deta[MAXB], phip[MAXB], Um[MAXNC], Up[MAXNC], Fx[MAXNC], Fy[MAXNC], F[MAXNC];
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) if (g_threads > 1) \
    private(phi, dxi, deta, phip, Um, Up, Fx, Fy, F)
    for (int i = 0; i < m->np; i++) {
        const double *Ui = u + (size_t)i * bl;
        double *bi = b + (size_t)i * bl;
        double w = m->bb[4 * i + 2], h = m->bb[4 * i + 3];
        for (int r = 0; r < bl; r++) bi[r] = 0.0;
        for (int tt = m->tptr[i]; tt < m->tptr[i + 1]; tt++) {
            const double *T = m->txy + 6 * tt;
            double area = tri_area(T);
            for (int q = 0; q < nt; q++) {
                double r = tr[3 * q], s = tr[3 * q + 1], W = tr[3 * q + 2] * area;
                double x = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
                double y = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
                double xi, eta;
                local_coords(m, i, x, y, &xi, &eta);
                o_basis(deg, xi, eta, phi, dxi, deta);
                eval_state(bs, nc, Ui, phi, Um);
                flux(eq, Um, x, y, Fx, Fy);
                for (int c = 0; c < nc; c++)
                    for (int k = 0; k < bs; k++)
                        bi[c * bs + k] += W * (Fx[c] * dxi[k] * 2.0 / w + Fy[c] * deta[k] * 2.0 / h);
            }
        }
        int nv = m->pptr[i + 1] - m->pptr[i];
        for (int e = 0; e < nv; e++) {
            int a1 = m->pv[m->pptr[i] + e], b1 = m->pv[m->pptr[i] + (e + 1) % nv];
            int nb = m->p2p[m->pptr[i] + e];
            double nx, ny;
            o_outward_normal(m, i, a1, b1, &nx, &ny);
            double x1 = m->v[2 * a1], y1 = m->v[2 * a1 + 1], x2 = m->v[2 * b1], y2 = m->v[2 * b1 + 1];
            double len = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
            for (int q = 0; q < nl; q++) {
                double x = x1 + 0.5 * (ln[2 * q] + 1) * (x2 - x1);
                double y = y1 + 0.5 * (ln[2 * q] + 1) * (y2 - y1);
                double W = 0.5 * len * ln[2 * q + 1];
                double xi, eta;
                local_coords(m, i, x, y, &xi, &eta);
                o_basis(deg, xi, eta, phi, NULL, NULL);
                eval_state(bs, nc, Ui, phi, Um);
                outer_state(m, eq, deg, u, nb, x, y, t, Up, phip);
                numflux(eq, Um, Up, x, y, nx, ny, F);
                for (int c = 0; c < nc; c++)
                    for (int k = 0; k < bs; k++) bi[c * bs + k] -= W * phi[k] * F[c];
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Jacobian pattern (Jacobian::Jacobian, src/Jacobian.cpp:7-47) and values
 * (Equation::jacobian, src/Equation.cpp:193-279).
 */
int o_jacobian_pattern(const o_mesh *m, int *rowptr, int *colind)
{
    int nb = 0;
    for (int i = 0; i < m->np; i++) {
        rowptr[i] = nb;
        if (colind) colind[nb] = i;
        nb++;
        for (int e = m->pptr[i]; e < m->pptr[i + 1]; e++) {
            int k = m->p2p[e];
            if (k >= 0) { if (colind) colind[nb] = k; nb++; }
            else if (m->bc[7 * (-1 - k)]) { if (colind) colind[nb] = m->bc[7 * (-1 - k) + 2]; nb++; }
        }
    }
    rowptr[m->np] = nb;
    return nb;
}

int o_jacobian(const o_mesh *m, const o_eq *eq, int deg, const double *u, double t,
               const int *rowptr, double *blocks)
{
    int bs = (deg + 1) * (deg + 2) / 2, nc = o_nc(eq), bl = nc * bs;
    const double *tr, *ln;
    int nt, nl;
    pick_rules(2 * deg, &tr, &nt, &ln, &nl);
    double phi[MAXB], dxi[MAXB], deta[MAXB], phip[MAXB], Um[MAXNC], Up[MAXNC];
    double Ax[16], Ay[16], Jd[16], Jo[16];
    size_t bb2 = (size_t)bl * bl;
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) if (g_threads > 1) \
    private(phi, dxi, deta, phip, Um, Up, Ax, Ay, Jd, Jo)
    for (int i = 0; i < m->np; i++) {
        const double *Ui = u + (size_t)i * bl;
        double w = m->bb[4 * i + 2], h = m->bb[4 * i + 3];
        int d = rowptr[i];
        memset(blocks + (size_t)d * bb2, 0, sizeof(double) * bb2 * (rowptr[i + 1] - rowptr[i]));
        double *D = blocks + (size_t)d * bb2;
        for (int tt = m->tptr[i]; tt < m->tptr[i + 1]; tt++) {
            const double *T = m->txy + 6 * tt;
            double area = tri_area(T);
            for (int q = 0; q < nt; q++) {
                double r = tr[3 * q], s = tr[3 * q + 1], W = tr[3 * q + 2] * area;
                double x = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
                double y = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
                double xi, eta;
                local_coords(m, i, x, y, &xi, &eta);
                o_basis(deg, xi, eta, phi, dxi, deta);
                eval_state(bs, nc, Ui, phi, Um);
                fluxjac(eq, Um, x, y, Ax, Ay);
                for (int c1 = 0; c1 < nc; c1++)
                    for (int c2 = 0; c2 < nc; c2++) {
                        double ax = Ax[c1 * nc + c2], ay = Ay[c1 * nc + c2];
                        for (int l = 0; l < bs; l++) {
                            double *col = D + (size_t)(c2 * bs + l) * bl + c1 * bs;
                            for (int k = 0; k < bs; k++)
                                col[k] += W * phi[l] * (ax * dxi[k] * 2.0 / w + ay * deta[k] * 2.0 / h);
                        }
                    }
            }
        }
        int nv = m->pptr[i + 1] - m->pptr[i];
        int slot = d + 1;
        for (int e = 0; e < nv; e++) {
            int a1 = m->pv[m->pptr[i] + e], b1 = m->pv[m->pptr[i] + (e + 1) % nv];
            int nb = m->p2p[m->pptr[i] + e];
            int coupled = nb >= 0 || m->bc[7 * (-1 - nb)];
            double nx, ny;
            o_outward_normal(m, i, a1, b1, &nx, &ny);
            double x1 = m->v[2 * a1], y1 = m->v[2 * a1 + 1], x2 = m->v[2 * b1], y2 = m->v[2 * b1 + 1];
            double len = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
            double *O = coupled ? blocks + (size_t)slot * bb2 : NULL;
            for (int q = 0; q < nl; q++) {
                double x = x1 + 0.5 * (ln[2 * q] + 1) * (x2 - x1);
                double y = y1 + 0.5 * (ln[2 * q] + 1) * (y2 - y1);
                double W = 0.5 * len * ln[2 * q + 1];
                double xi, eta;
                local_coords(m, i, x, y, &xi, &eta);
                o_basis(deg, xi, eta, phi, NULL, NULL);
                eval_state(bs, nc, Ui, phi, Um);
                outer_state(m, eq, deg, u, nb, x, y, t, Up, phip);
                numfluxjac(eq, Um, Up, x, y, nx, ny, +1, Jd);
                if (coupled) {
                    if (eq->jacobian && eq->kind == O_EULER) {
                        /* exact: dFhat/dU+ = 0.5 (A(U+) n - alpha I + (U- - U+) dalpha/dU+^T) */
                        double Bx[16], By[16], dd[4], alpha;
                        euler_fluxjac(eq->gamma, Up, Bx, By, 0);
                        euler_alpha_deriv_outer(eq->gamma, Um, Up, nx, ny, &alpha, dd);
                        for (int r = 0; r < 4; r++)
                            for (int c = 0; c < 4; c++)
                                Jo[4 * r + c] = 0.5 * (Bx[4 * r + c] * nx + By[4 * r + c] * ny
                                                       - (r == c ? alpha : 0.0) + (Um[r] - Up[r]) * dd[c]);
                    } else {
                        numfluxjac(eq, Um, Up, x, y, nx, ny, -1, Jo);
                    }
                    for (int c1 = 0; c1 < nc; c1++)
                        for (int c2 = 0; c2 < nc; c2++) {
                            double jo = Jo[c1 * nc + c2];
                            for (int l = 0; l < bs; l++) {
                                double *col = O + (size_t)(c2 * bs + l) * bl + c1 * bs;
                                for (int k = 0; k < bs; k++) col[k] -= W * phi[k] * phip[l] * jo;
                            }
                        }
                }
                for (int c1 = 0; c1 < nc; c1++)
                    for (int c2 = 0; c2 < nc; c2++) {
                        double jd = Jd[c1 * nc + c2];
                        for (int l = 0; l < bs; l++) {
                            double *col = D + (size_t)(c2 * bs + l) * bl + c1 * bs;
                            for (int k = 0; k < bs; k++) col[k] -= W * phi[k] * phi[l] * jd;
                        }
                    }
            }
            if (coupled) slot++;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Block sparse algebra
 */
void o_bsr_matvec(int n, int bl, const int *rowptr, const int *colind, const double *A,
                  const double *x, double *y)
{
    size_t bb2 = (size_t)bl * bl;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int i = 0; i < n; i++) {
        double *yi = y + (size_t)i * bl;
        for (int r = 0; r < bl; r++) yi[r] = 0.0;
        for (int s = rowptr[i]; s < rowptr[i + 1]; s++) {
            const double *B = A + s * bb2, *xj = x + (size_t)colind[s] * bl;
            for (int c = 0; c < bl; c++) {
                double xc = xj[c];
                for (int r = 0; r < bl; r++) yi[r] += B[c * bl + r] * xc;
            }
        }
    }
}

/* LU with partial pivoting (first maximal |pivot|), column-major, LAPACK getrf semantics */
static int lu_factor(int n, double *A, int *piv)
{
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int r = k + 1; r < n; r++)
            if (fabs(A[k * n + r]) > best) { best = fabs(A[k * n + r]); p = r; }
        piv[k] = p;
        if (A[k * n + p] == 0.0) { if (!info) info = k + 1; continue; }
        if (p != k)
            for (int c = 0; c < n; c++) { double tmp = A[c * n + k]; A[c * n + k] = A[c * n + p]; A[c * n + p] = tmp; }
        double inv = 1.0 / A[k * n + k];
        for (int r = k + 1; r < n; r++) A[k * n + r] *= inv;
        for (int c = k + 1; c < n; c++) {
            double akc = A[c * n + k];
            if (akc != 0.0)
                for (int r = k + 1; r < n; r++) A[c * n + r] -= A[k * n + r] * akc;
        }
    }
    return info;
}

static void lu_solve(int n, const double *LU, const int *piv, double *x)
{
    for (int k = 0; k < n; k++)
        if (piv[k] != k) { double tmp = x[k]; x[k] = x[piv[k]]; x[piv[k]] = tmp; }
    for (int c = 0; c < n; c++) {
        double xc = x[c];
        for (int r = c + 1; r < n; r++) x[r] -= LU[c * n + r] * xc;
    }
    for (int c = n - 1; c >= 0; c--) {
        x[c] /= LU[c * n + c];
        double xc = x[c];
        for (int r = 0; r < c; r++) x[r] -= LU[c * n + r] * xc;
    }
}

int o_inv(int bl, const double *A, double *Ainv)
{
    double *LU = malloc(sizeof(double) * bl * bl);
    int *piv = malloc(sizeof(int) * bl);
    memcpy(LU, A, sizeof(double) * bl * bl);
    int info = lu_factor(bl, LU, piv);
    if (!info)
        for (int c = 0; c < bl; c++) {
            double *col = Ainv + (size_t)c * bl;
            for (int r = 0; r < bl; r++) col[r] = (r == c);
            lu_solve(bl, LU, piv, col);
        }
    free(LU);
    free(piv);
    return info;
}

/* C = A*B (all bl x bl column-major) */
static void gemm(int bl, const double *A, const double *B, double *C)
{
    for (int c = 0; c < bl; c++) {
        double *Cc = C + (size_t)c * bl;
        for (int r = 0; r < bl; r++) Cc[r] = 0.0;
        for (int k = 0; k < bl; k++) {
            double bkc = B[(size_t)c * bl + k];
            const double *Ak = A + (size_t)k * bl;
            for (int r = 0; r < bl; r++) Cc[r] += Ak[r] * bkc;
        }
    }
}

static int find_slot(const int *rowptr, const int *colind, int row, int col, int skip_diag)
{
    for (int s = rowptr[row]; s < rowptr[row + 1]; s++)
        if (colind[s] == col && !(skip_diag && col == row)) return s;
    return -1;
}

/* BlockILU0::BlockILU0 (src/Preconditioners.cpp:49-123): block DILU in natural row order.
 * F gets A's slot layout: F[diag] = D_i^-1, F[(i,j<i)] = L_ij = A_ij D_j^-1,
 * F[(i,j>i)] = DU_ij = D_i^-1 A_ij. */
/* Dependency levels of the lower (fwd) or upper (!fwd) triangle of the block pattern: rows of one
 * level have no lower (upper) neighbour in the same or a later level, so they are independent.
 * Returns the level count; order[] lists the rows level by level, lptr[] the level starts. */
static int dilu_levels(int n, const int *rowptr, const int *colind, int fwd, int *order, int *lptr)
{
    int *lev = malloc(sizeof(int) * (n > 0 ? n : 1)), nl = 0;
    for (int t = 0; t < n; t++) {
        int i = fwd ? t : n - 1 - t, l = 0;
        for (int s = rowptr[i]; s < rowptr[i + 1]; s++) {
            int j = colind[s];
            if ((fwd && j < i) || (!fwd && j > i)) l = lev[j] + 1 > l ? lev[j] + 1 : l;
        }
        lev[i] = l;
        if (l + 1 > nl) nl = l + 1;
    }
    for (int l = 0; l <= nl; l++) lptr[l] = 0;
    for (int i = 0; i < n; i++) lptr[lev[i] + 1]++;
    for (int l = 0; l < nl; l++) lptr[l + 1] += lptr[l];
    int *fill = malloc(sizeof(int) * (nl + 1));
    memcpy(fill, lptr, sizeof(int) * (nl + 1));
    for (int t = 0; t < n; t++) {
        int i = fwd ? t : n - 1 - t;
        order[fill[lev[i]]++] = i;
    }
    free(fill);
    free(lev);
    return nl;
}

/* The same factor with host threads: left-looking per row, the rows of one level in parallel
 * (S_j = A_jj - sum_i L_ji A_ij with L_ji = A_ji D_i^-1 over the lower neighbours i, D_j^-1 =
 * S_j^-1), then DU_ij = D_i^-1 A_ij for every upper block. */
static int dilu_factor_threaded(int n, int bl, const int *rowptr, const int *colind, double *F)
{
    size_t bb2 = (size_t)bl * bl;
    int *order = malloc(sizeof(int) * (n > 0 ? n : 1)), *lptr = malloc(sizeof(int) * (n + 2));
    int nl = dilu_levels(n, rowptr, colind, 1, order, lptr), info = 0;
    for (int l = 0; l < nl; l++) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
        for (int t = lptr[l]; t < lptr[l + 1]; t++) {
            int j = order[t], dj = find_slot(rowptr, colind, j, j, 0);
            double *T = malloc(sizeof(double) * bb2), *Dn = malloc(sizeof(double) * bb2);
            for (int s = rowptr[j]; s < rowptr[j + 1]; s++) {
                int i = colind[s];
                if (i >= j) continue;
                int di = find_slot(rowptr, colind, i, i, 0), sij = find_slot(rowptr, colind, i, j, 1);
                gemm(bl, F + s * bb2, F + di * bb2, T);          /* L_ji = A_ji D_i^-1 */
                memcpy(F + s * bb2, T, sizeof(double) * bb2);
                if (sij < 0) continue;
                gemm(bl, F + s * bb2, F + sij * bb2, T);         /* S_j -= L_ji A_ij */
                for (size_t e = 0; e < bb2; e++) F[dj * bb2 + e] -= T[e];
            }
            int st = o_inv(bl, F + dj * bb2, Dn);
            memcpy(F + dj * bb2, Dn, sizeof(double) * bb2);
            if (st) {
#pragma omp critical
                if (!info || j + 1 < info) info = j + 1;
            }
            free(T);
            free(Dn);
        }
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
    for (int i = 0; i < n; i++) {
        double *T = malloc(sizeof(double) * bb2);
        int di = find_slot(rowptr, colind, i, i, 0);
        for (int s = rowptr[i]; s < rowptr[i + 1]; s++) {
            if (colind[s] <= i) continue;
            gemm(bl, F + di * bb2, F + s * bb2, T);              /* DU_ij = D_i^-1 A_ij */
            memcpy(F + s * bb2, T, sizeof(double) * bb2);
        }
        free(T);
    }
    free(order);
    free(lptr);
    return info;
}

int o_dilu_factor(int n, int bl, const int *rowptr, const int *colind, const double *A, double *F)
{
    size_t bb2 = (size_t)bl * bl;
    size_t nnzb = rowptr[n];
    memcpy(F, A, sizeof(double) * bb2 * nnzb);
    if (g_threads > 1) return dilu_factor_threaded(n, bl, rowptr, colind, F);
    double *D = malloc(sizeof(double) * bb2), *T = malloc(sizeof(double) * bb2);
    int info = 0;
    for (int i = 0; i < n; i++) {
        int di = find_slot(rowptr, colind, i, i, 0);
        if (di < 0) { info = -1; break; }
        int st = o_inv(bl, F + di * bb2, D);
        if (st && !info) info = i + 1;
        memcpy(F + di * bb2, D, sizeof(double) * bb2);
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            int j = colind[k];
            if (j == i || j < i) continue;
            int k2 = find_slot(rowptr, colind, j, i, 1);
            if (k2 < 0) continue;
            int dj = find_slot(rowptr, colind, j, j, 0);
            gemm(bl, F + k2 * bb2, D, T);             /* L_ji = A_ji D_i */
            memcpy(F + k2 * bb2, T, sizeof(double) * bb2);
            gemm(bl, F + k2 * bb2, F + k * bb2, T);   /* D_j -= L_ji A_ij */
            for (size_t e = 0; e < bb2; e++) F[dj * bb2 + e] -= T[e];
            gemm(bl, D, F + k * bb2, T);              /* DU_ij = D_i A_ij */
            memcpy(F + k * bb2, T, sizeof(double) * bb2);
        }
    }
    free(D);
    free(T);
    return info;
}

/* the apply with host threads: each sweep level by level, the rows of a level in parallel, each
 * row's arithmetic in the serial order (the same result as the serial apply) */
static void dilu_apply_threaded(int n, int bl, const int *rowptr, const int *colind, const double *F, double *x)
{
    size_t bb2 = (size_t)bl * bl;
    int *order = malloc(sizeof(int) * (n > 0 ? n : 1)), *lptr = malloc(sizeof(int) * (n + 2));
    for (int fwd = 1; fwd >= 0; fwd--) {
        int nl = dilu_levels(n, rowptr, colind, fwd, order, lptr);
        for (int l = 0; l < nl; l++) {
#pragma omp parallel for schedule(dynamic, 32) num_threads(g_threads)
            for (int t = lptr[l]; t < lptr[l + 1]; t++) {
                int i = order[t];
                for (int s = rowptr[i]; s < rowptr[i + 1]; s++) {
                    int j = colind[s];
                    if (fwd ? j >= i : j <= i) continue;
                    const double *B = F + s * bb2;
                    for (int c = 0; c < bl; c++)
                        for (int r = 0; r < bl; r++) x[(size_t)i * bl + r] -= B[c * bl + r] * x[(size_t)j * bl + c];
                }
            }
        }
        if (fwd) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
            for (int i = 0; i < n; i++) {
                double tmp[64 * MAXNC];
                int d = find_slot(rowptr, colind, i, i, 0);
                const double *B = F + d * bb2;
                for (int r = 0; r < bl; r++) tmp[r] = 0.0;
                for (int c = 0; c < bl; c++)
                    for (int r = 0; r < bl; r++) tmp[r] += B[c * bl + r] * x[(size_t)i * bl + c];
                memcpy(x + (size_t)i * bl, tmp, sizeof(double) * bl);
            }
        }
    }
    free(order);
    free(lptr);
}

/* BlockILU0::solve (src/Preconditioners.cpp:125-158) */
void o_dilu_apply(int n, int bl, const int *rowptr, const int *colind, const double *F, double *x)
{
    size_t bb2 = (size_t)bl * bl;
    if (g_threads > 1 && bl <= 64 * MAXNC) {
        dilu_apply_threaded(n, bl, rowptr, colind, F, x);
        return;
    }
    double *tmp = malloc(sizeof(double) * bl);
    for (int i = 0; i < n; i++)
        for (int s = rowptr[i]; s < rowptr[i + 1]; s++) {
            int j = colind[s];
            if (j >= i) continue;
            const double *B = F + s * bb2;
            for (int c = 0; c < bl; c++)
                for (int r = 0; r < bl; r++) x[(size_t)i * bl + r] -= B[c * bl + r] * x[(size_t)j * bl + c];
        }
    for (int i = 0; i < n; i++) {
        int d = find_slot(rowptr, colind, i, i, 0);
        const double *B = F + d * bb2;
        for (int r = 0; r < bl; r++) tmp[r] = 0.0;
        for (int c = 0; c < bl; c++)
            for (int r = 0; r < bl; r++) tmp[r] += B[c * bl + r] * x[(size_t)i * bl + c];
        memcpy(x + (size_t)i * bl, tmp, sizeof(double) * bl);
    }
    for (int i = n - 1; i >= 0; i--)
        for (int s = rowptr[i]; s < rowptr[i + 1]; s++) {
            int j = colind[s];
            if (j <= i) continue;
            const double *B = F + s * bb2;
            for (int c = 0; c < bl; c++)
                for (int r = 0; r < bl; r++) x[(size_t)i * bl + r] -= B[c * bl + r] * x[(size_t)j * bl + c];
        }
    free(tmp);
}

/* BlockJacobi (src/Preconditioners.cpp:8-47) */
int o_bjacobi_factor(int n, int bl, const int *rowptr, const int *colind, const double *A,
                     double *LU, int *piv)
{
    size_t bb2 = (size_t)bl * bl;
    int info = 0;
    for (int i = 0; i < n; i++) {
        int d = find_slot(rowptr, colind, i, i, 0);
        memcpy(LU + i * bb2, A + d * bb2, sizeof(double) * bb2);
        int st = lu_factor(bl, LU + i * bb2, piv + (size_t)i * bl);
        if (st && !info) info = i + 1;
    }
    return info;
}

void o_bjacobi_apply(int n, int bl, const double *LU, const int *piv, double *x)
{
    size_t bb2 = (size_t)bl * bl;
    for (int i = 0; i < n; i++) lu_solve(bl, LU + i * bb2, piv + (size_t)i * bl, x + (size_t)i * bl);
}

typedef struct {
    int n, bl, kind;
    const int *rowptr, *colind;
    const double *data;
    const int *piv;
} pc_t;

static void pc_solve(const pc_t *pc, double *x)
{
    if (pc->kind == O_PC_DILU) o_dilu_apply(pc->n, pc->bl, pc->rowptr, pc->colind, pc->data, x);
    else if (pc->kind == O_PC_BJACOBI) o_bjacobi_apply(pc->n, pc->bl, pc->data, pc->piv, x);
}

static double dot(size_t n, const double *x, const double *y)
{
    double s = 0;
    if (g_threads > 1) {
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : s)
        for (size_t i = 0; i < n; i++) s += x[i] * y[i];
        return s;
    }
    for (size_t i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

static double nrm2(size_t n, const double *x)
{
    return sqrt(dot(n, x, x));
}

/* BlockMatrix::gmres (src/BlockMatrix.cpp:79-163) -- mode O_GMRES_LEFT_REF; and a
 * right-preconditioned flexible GMRES (mode O_GMRES_FLEXIBLE, tolerance relative to ||b||). */
int o_gmres(int n, int bl, const int *rowptr, const int *colind, const double *A,
            int pc_kind, const double *pc_data, const int *piv,
            const double *b, double *x, int m, double *tol, int *maxit, int mode)
{
    size_t N = (size_t)n * bl;
    pc_t pc = {n, bl, pc_kind, rowptr, colind, pc_data, piv};
    double *v = malloc(sizeof(double) * N * (m + 1));
    double *z = mode == O_GMRES_FLEXIBLE ? malloc(sizeof(double) * N * m) : NULL;
    double *h = malloc(sizeof(double) * (m * (m + 1) / 2 + 1));
    double *r = malloc(sizeof(double) * N);
    double *y = malloc(sizeof(double) * (m + 1));
    double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
    double nrm2b, beta, hnew, rd, dd;
    int i = 0, j, uki, u0i;
    double tl = *tol;
    int mx = *maxit;

    memcpy(r, b, sizeof(double) * N);
    if (mode == O_GMRES_LEFT_REF) pc_solve(&pc, r);
    nrm2b = nrm2(N, r);
    if (mx % m > 0) mx = (mx / m + 1) * m;
    if (nrm2b == 0.0) {
        memset(x, 0, sizeof(double) * N);
        *tol = 0.0;
        *maxit = 0;
        free(v); free(z); free(h); free(r); free(y); free(cs); free(sn);
        return 0;
    }
    y[0] = 0;
    for (j = 0; j < mx / m; j++) {
        if (mode == O_GMRES_LEFT_REF) {
            /* r = P^-1 (A x - b) */
            o_bsr_matvec(n, bl, rowptr, colind, A, x, r);
            for (size_t e = 0; e < N; e++) r[e] -= b[e];
            pc_solve(&pc, r);
        } else {
            /* r = A x - b (sign kept as in the reference so that x -= Z y) */
            o_bsr_matvec(n, bl, rowptr, colind, A, x, r);
            for (size_t e = 0; e < N; e++) r[e] -= b[e];
        }
        beta = nrm2(N, r);
        for (size_t e = 0; e < N; e++) v[e] = r[e] / beta;
        y[0] = beta;
        uki = 0;
        for (i = 0; i < m; i++) {
            if (fabs(y[i]) < tl * nrm2b) break;
            double *vi = v + N * i, *vi1 = v + N * (i + 1);
            u0i = uki;
            if (mode == O_GMRES_LEFT_REF) {
                o_bsr_matvec(n, bl, rowptr, colind, A, vi, vi1);
                pc_solve(&pc, vi1);
            } else {
                double *zi = z + N * i;
                memcpy(zi, vi, sizeof(double) * N);
                pc_solve(&pc, zi);
                o_bsr_matvec(n, bl, rowptr, colind, A, zi, vi1);
            }
            for (int k = 0; k <= i; k++) h[u0i + k] = dot(N, v + N * k, vi1);
            for (int k = 0; k <= i; k++) {
                double hk = h[u0i + k];
                const double *vk = v + N * k;
                for (size_t e = 0; e < N; e++) vi1[e] -= hk * vk[e];
            }
            hnew = nrm2(N, vi1);
            for (size_t e = 0; e < N; e++) vi1[e] *= 1.0 / hnew;
            for (int k = 0; k < i; k++) {
                double tmp = cs[k] * h[uki] - sn[k] * h[uki + 1];
                h[uki + 1] = sn[k] * h[uki] + cs[k] * h[uki + 1];
                h[uki] = tmp;
                ++uki;
            }
            rd = h[uki];
            dd = sqrt(rd * rd + hnew * hnew);
            cs[i] = rd / dd;
            sn[i] = -hnew / dd;
            h[uki] = dd;
            ++uki;
            y[i + 1] = sn[i] * y[i];
            y[i] = cs[i] * y[i];
        }
        /* packed upper-triangular solve (dtpsv 'U','N','N'): column k of H starts at k(k+1)/2 */
        for (int c = i - 1; c >= 0; c--) {
            y[c] /= h[c * (c + 1) / 2 + c];
            double yc = y[c];
            for (int rr = 0; rr < c; rr++) y[rr] -= h[c * (c + 1) / 2 + rr] * yc;
        }
        const double *basis = mode == O_GMRES_LEFT_REF ? v : z;
        for (int k = 0; k < i; k++) {
            const double *bk = basis + N * k;
            double yk = y[k];
            for (size_t e = 0; e < N; e++) x[e] -= bk[e] * yk;
        }
        if (fabs(y[i]) < tl * nrm2b) break;
    }
    *maxit = m * j + i;
    *tol = fabs(y[i]) / nrm2b;
    free(v); free(z); free(h); free(r); free(y); free(cs); free(sn);
    return 0;
}

/* BlockMatrix::jacobi (src/BlockMatrix.cpp:51-77), omega = 1 */
int o_jacobi_solve(int n, int bl, const int *rowptr, const int *colind, const double *A,
                   const double *LU, const int *piv, const double *b, double *x,
                   double *tol, int *maxit)
{
    size_t N = (size_t)n * bl;
    double *Dinvb = malloc(sizeof(double) * N), *t = malloc(sizeof(double) * N), *r = malloc(sizeof(double) * N);
    memcpy(Dinvb, b, sizeof(double) * N);
    o_bjacobi_apply(n, bl, LU, piv, Dinvb);
    int it;
    double rn = 0;
    for (it = 0; it < *maxit; it++) {
        o_bsr_matvec(n, bl, rowptr, colind, A, x, t);
        for (size_t e = 0; e < N; e++) r[e] = t[e] - b[e];
        rn = nrm2(N, r);
        if (rn < *tol) break;
        o_bjacobi_apply(n, bl, LU, piv, t);
        for (size_t e = 0; e < N; e++) x[e] = Dinvb[e] + x[e] - t[e];
    }
    *tol = rn;
    *maxit = it;
    free(Dinvb); free(t); free(r);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Time stepping (src/TimeIntegration.cpp)
 */
/* Alexander's L-stable 3-stage SDIRK (R. Alexander, SIAM J. Numer. Anal. 14 (1977) 1006-1021):
 * gamma is the root in (1/6, 1/2) of x^3 - 3x^2 + 3x/2 - 1/6, c = (gamma, (1+gamma)/2, 1), and the
 * last row of A is b (stiffly accurate), so R(-inf) = 0.  Not in the reference: the reference's
 * 3-stage tableau (src/TimeIntegration.cpp:71-77) is not A-stable (|R(-inf)| = 1.78). */
static void alexander_sdirk3(double *A, double *b, double *c)
{
    const double g = 0.43586652150845899942;
    const double b1 = -(6 * g * g - 16 * g + 1) / 4.0, b2 = (6 * g * g - 20 * g + 5) / 4.0;
    double a[9] = {g, 0, 0, (1 - g) / 2.0, g, 0, b1, b2, g};
    memcpy(A, a, sizeof a);
    b[0] = b1; b[1] = b2; b[2] = g;
    c[0] = g; c[1] = (1 + g) / 2.0; c[2] = 1;
}

void o_dirk_tableau_v(int nstages, int variant, double *A, double *b, double *c)
{
    if (nstages == 3 && variant == O_DIRK3_L_STABLE) { alexander_sdirk3(A, b, c); return; }
    o_dirk_tableau(nstages, A, b, c);
}

void o_dirk_tableau(int nstages, double *A, double *b, double *c)
{
    if (nstages == 1) { A[0] = 1; b[0] = 1; c[0] = 1; return; }   /* backward Euler */
    if (nstages == 3) {
        const double al = 0.435866521508459;
        double a[9] = {(1 + al) / 2.0, 0, 0, -al / 2.0, (1 + al) / 2.0, 0, 1 + al, -(1 + 2 * al), (1 + al) / 2.0};
        memcpy(A, a, sizeof a);
        b[0] = 1.0 / 6.0 / (al * al); b[1] = 1 - 1.0 / 3.0 / (al * al); b[2] = 1.0 / 6.0 / (al * al);
        c[0] = (1 + al) / 2.0; c[1] = 0.5; c[2] = (1 - al) / 2.0;
        return;
    }
    if (nstages == 5) {
        double s6 = sqrt(6.0), g = (6 - s6) / 10.0;
        double a[25] = {g, 0, 0, 0, 0,
                        (-6 + 5 * s6) / 14.0, g, 0, 0, 0,
                        (888 + 607 * s6) / 2850.0, (126 - 161 * s6) / 1425.0, g, 0, 0,
                        (3153 - 3082 * s6) / 14250.0, (3213 + 1148 * s6) / 28500.0, (-267 + 88 * s6) / 500.0, g, 0,
                        (-32583 + 14638 * s6) / 71250.0, (-17199 + 364 * s6) / 142500.0, (1329 - 544 * s6) / 2500.0,
                        (-96 + 131 * s6) / 625.0, g};
        memcpy(A, a, sizeof a);
        b[0] = 0; b[1] = 0; b[2] = 1 / 9.0; b[3] = (16 - s6) / 36.0; b[4] = (16 + s6) / 36.0;
        c[0] = g; c[1] = (6 + 9 * s6) / 35.0; c[2] = 1; c[3] = (4 - s6) / 10.0; c[4] = (4 + s6) / 10.0;
        return;
    }
    if (nstages == 2) {
        /* two-stage, L-stable, second-order SDIRK (gamma = 1 - 1/sqrt(2)); not in the reference */
        double gm = 1.0 - 1.0 / sqrt(2.0);
        A[0] = gm; A[1] = 0; A[2] = 1 - gm; A[3] = gm;
        b[0] = 1 - gm; b[1] = gm;
        c[0] = gm; c[1] = 1;
    }
}

void o_irk_tableau(int nstages, double *A, double *b, double *c)
{
    if (nstages == 2) {
        double a[4] = {5 / 12.0, -1 / 12.0, 3 / 4.0, 1 / 4.0};
        memcpy(A, a, sizeof a);
        b[0] = 3 / 4.0; b[1] = 1 / 4.0; c[0] = 1 / 3.0; c[1] = 1;
    } else {
        double s6 = sqrt(6.0);
        double a[9] = {(88 - 7 * s6) / 360.0, (296 - 169 * s6) / 1800.0, (-2 + 3 * s6) / 225.0,
                       (296 + 169 * s6) / 1800.0, (88 + 7 * s6) / 360.0, (-2 - 3 * s6) / 225.0,
                       (16 - s6) / 36.0, (16 + s6) / 36.0, 1 / 9.0};
        memcpy(A, a, sizeof a);
        b[0] = (16 - s6) / 36.0; b[1] = (16 + s6) / 36.0; b[2] = 1 / 9.0;
        c[0] = (4 - s6) / 10.0; c[1] = (4 + s6) / 10.0; c[2] = 1;
    }
}

static double max_l2(const o_mesh *m, int deg, int nc, const double *r)
{
    double l2[MAXNC];
    o_l2norm(m, deg, nc, r, l2);
    double mx = l2[0];
    for (int c = 1; c < nc; c++) if (l2[c] > mx) mx = l2[c];
    return mx;
}

/* solve B x = r with the configured preconditioner and GMRES; x starts from the
 * caller's content (zero in Jacobian::solve, src/Jacobian.cpp:72) */
static int linear_solve(int n, int bl, const int *rowptr, const int *colind, const double *B,
                        const o_solver *s, const double *r, double *x, o_stats *st)
{
    size_t bb2 = (size_t)bl * bl, nnzb = rowptr[n];
    double *pcd = NULL;
    int *piv = NULL;
    if (s->pc_kind == O_PC_DILU) {
        pcd = malloc(sizeof(double) * bb2 * nnzb);
        o_dilu_factor(n, bl, rowptr, colind, B, pcd);
    } else if (s->pc_kind == O_PC_BJACOBI) {
        pcd = malloc(sizeof(double) * bb2 * n);
        piv = malloc(sizeof(int) * (size_t)n * bl);
        o_bjacobi_factor(n, bl, rowptr, colind, B, pcd, piv);
    }
    double tol = s->rtol;
    int it = s->maxit;
    o_gmres(n, bl, rowptr, colind, B, s->pc_kind, pcd, piv, r, x, s->restart, &tol, &it, s->gmres_mode);
    if (st) { st->gmres_iters += it; st->linear_solves++; }
    free(pcd);
    free(piv);
    return 0;
}

/* DIRK::advance (src/TimeIntegration.cpp:93-151); nstages == 1 is BackwardEuler::advance
 * (:31-66) written in increment form k = u_{n+1} - u_n. */
int o_dirk_step(const o_mesh *m, const o_eq *eq, int deg, int nstages, const double *u,
                double dt, double t, const o_solver *s, double *unp1, o_stats *st)
{
    int n = m->np, bs = (deg + 1) * (deg + 2) / 2, nc = o_nc(eq), bl = nc * bs;
    size_t N = (size_t)n * bl, bb2 = (size_t)bl * bl;
    double A[25], b[5], c[5];
    o_dirk_tableau_v(nstages, s->dirk_tableau, A, b, c);
    int *rowptr = malloc(sizeof(int) * (n + 1));
    int nnzb = o_jacobian_pattern(m, rowptr, NULL);
    int *colind = malloc(sizeof(int) * nnzb);
    o_jacobian_pattern(m, rowptr, colind);
    double *M = malloc(sizeof(double) * bb2 * n);
    o_mass(m, deg, nc, M);
    double *k = calloc(N * nstages, sizeof(double));
    double *cu = malloc(sizeof(double) * N), *rhs = malloc(sizeof(double) * N), *r = malloc(sizeof(double) * N);
    double *x = malloc(sizeof(double) * N);
    double *B = malloc(sizeof(double) * bb2 * nnzb);
    if (st) memset(st, 0, sizeof *st);
    for (int stg = 0; stg < nstages; stg++) {
        for (int it = 0; it < s->newton_maxit; it++) {
            memcpy(cu, u, sizeof(double) * N);
            for (int j = 0; j <= stg; j++)
                for (size_t e = 0; e < N; e++) cu[e] += A[stg * nstages + j] * k[j * N + e];
            double ts = t + dt * c[stg];
            o_assemble(m, eq, deg, cu, ts, rhs);
            /* r = M k_s - dt b(cu); M is block diagonal */
            for (int i = 0; i < n; i++) {
                const double *Mi = M + i * bb2;
                const double *ki = k + stg * N + (size_t)i * bl;
                double *ri = r + (size_t)i * bl;
                for (int rr = 0; rr < bl; rr++) ri[rr] = 0.0;
                for (int cc = 0; cc < bl; cc++)
                    for (int rr = 0; rr < bl; rr++) ri[rr] += Mi[cc * bl + rr] * ki[cc];
            }
            for (size_t e = 0; e < N; e++) r[e] -= dt * rhs[e];
            double rn = max_l2(m, deg, nc, r);
            if (st) st->last_resid = rn;
            if (rn < s->newton_tol) break;
            if (st) st->newton_iters++;
            o_jacobian(m, eq, deg, cu, ts, rowptr, B);
            double sc = -dt * A[stg * nstages + stg];
            for (size_t e = 0; e < bb2 * nnzb; e++) B[e] *= sc;
            for (int i = 0; i < n; i++)
                for (size_t e = 0; e < bb2; e++) B[(size_t)rowptr[i] * bb2 + e] += M[i * bb2 + e];
            memset(x, 0, sizeof(double) * N);
            linear_solve(n, bl, rowptr, colind, B, s, r, x, st);
            for (size_t e = 0; e < N; e++) k[stg * N + e] -= x[e];
        }
    }
    memcpy(unp1, u, sizeof(double) * N);
    for (int i = 0; i < nstages; i++)
        for (size_t e = 0; e < N; e++) unp1[e] += b[i] * k[i * N + e];
    free(rowptr); free(colind); free(M); free(k); free(cu); free(rhs); free(r); free(x); free(B);
    return 0;
}

/* IRK::advance (src/TimeIntegration.cpp:188-281): stage-major coupled system built by
 * BlockMatrix::blockBlockMatrix (src/BlockMatrix.cpp:228-258); the GMRES initial guess is
 * the previous Newton correction (bigK is not reset between Newton iterations). */
int o_irk_step(const o_mesh *m, const o_eq *eq, int deg, int ns, const double *u,
               double dt, double t, const o_solver *s, double *unp1, o_stats *st)
{
    int n = m->np, bs = (deg + 1) * (deg + 2) / 2, nc = o_nc(eq), bl = nc * bs;
    size_t N = (size_t)n * bl, bb2 = (size_t)bl * bl;
    double A[9], b[3], c[3];
    o_irk_tableau(ns, A, b, c);
    int *rowptr = malloc(sizeof(int) * (n + 1));
    int nnzb = o_jacobian_pattern(m, rowptr, NULL);
    int *colind = malloc(sizeof(int) * nnzb);
    o_jacobian_pattern(m, rowptr, colind);
    double *M = malloc(sizeof(double) * bb2 * n);
    o_mass(m, deg, nc, M);
    double *k = calloc(N * ns, sizeof(double)), *w = malloc(sizeof(double) * N * ns);
    double *rhs = malloc(sizeof(double) * N), *bigR = malloc(sizeof(double) * N * ns);
    double *bigK = calloc(N * ns, sizeof(double));
    double *Js = malloc(sizeof(double) * bb2 * nnzb * ns);
    /* stage-major pattern */
    int bn = n * ns;
    int *brow = malloc(sizeof(int) * (bn + 1));
    int *bcol = malloc(sizeof(int) * (size_t)nnzb * ns * ns);
    double *bB = malloc(sizeof(double) * bb2 * nnzb * ns * ns);
    if (st) memset(st, 0, sizeof *st);
    for (int it = 0; it < s->newton_maxit; it++) {
        double rn = 0;
        for (int stg = 0; stg < ns; stg++) {
            double *ws = w + stg * N;
            memcpy(ws, u, sizeof(double) * N);
            for (int j = 0; j < ns; j++)
                for (size_t e = 0; e < N; e++) ws[e] += A[stg * ns + j] * k[j * N + e];
            o_assemble(m, eq, deg, ws, t + dt * c[stg], rhs);
            double *r = bigR + stg * N;
            for (int i = 0; i < n; i++) {
                const double *Mi = M + i * bb2;
                const double *ki = k + stg * N + (size_t)i * bl;
                double *ri = r + (size_t)i * bl;
                for (int rr = 0; rr < bl; rr++) ri[rr] = 0.0;
                for (int cc = 0; cc < bl; cc++)
                    for (int rr = 0; rr < bl; rr++) ri[rr] += Mi[cc * bl + rr] * ki[cc];
            }
            for (size_t e = 0; e < N; e++) r[e] -= dt * rhs[e];
            double l = max_l2(m, deg, nc, r);
            rn += l * l;
        }
        rn = sqrt(rn);
        if (st) st->last_resid = rn;
        if (rn < s->newton_tol) break;
        if (st) st->newton_iters++;
        for (int stg = 0; stg < ns; stg++)
            o_jacobian(m, eq, deg, w + stg * N, t + dt * c[stg], rowptr, Js + (size_t)stg * bb2 * nnzb);
        int nbk = 0;
        for (int si = 0; si < ns; si++)
            for (int i = 0; i < n; i++) {
                brow[si * n + i] = nbk;
                for (int sj = 0; sj < ns; sj++)
                    for (int sl = rowptr[i]; sl < rowptr[i + 1]; sl++) {
                        double sc = -dt * A[si * ns + sj];
                        const double *src = Js + ((size_t)si * nnzb + sl) * bb2;
                        double *dst = bB + (size_t)nbk * bb2;
                        for (size_t e = 0; e < bb2; e++) dst[e] = sc * src[e];
                        if (si == sj && sl == rowptr[i])
                            for (size_t e = 0; e < bb2; e++) dst[e] += M[i * bb2 + e];
                        bcol[nbk] = sj * n + colind[sl];
                        nbk++;
                    }
            }
        brow[bn] = nbk;
        linear_solve(bn, bl, brow, bcol, bB, s, bigR, bigK, st);
        for (size_t e = 0; e < N * ns; e++) k[e] -= bigK[e];
    }
    memcpy(unp1, u, sizeof(double) * N);
    for (int i = 0; i < ns; i++)
        for (size_t e = 0; e < N; e++) unp1[e] += b[i] * k[i * N + e];
    free(rowptr); free(colind); free(M); free(k); free(w); free(rhs); free(bigR); free(bigK);
    free(Js); free(brow); free(bcol); free(bB);
    return 0;
}

/* dt M^-1 b(v, tv) into k, M's per-cell LU factors in Mf/piv */
static void explicit_stage(const o_mesh *m, const o_eq *eq, int deg, const double *Mf, const int *piv,
                           const double *v, double tv, double dt, double *k)
{
    int n = m->np, bs = (deg + 1) * (deg + 2) / 2, bl = o_nc(eq) * bs;
    size_t N = (size_t)n * bl, bb2 = (size_t)bl * bl;
    o_assemble(m, eq, deg, v, tv, k);
    for (int i = 0; i < n; i++) lu_solve(bl, Mf + i * bb2, piv + (size_t)i * bl, k + (size_t)i * bl);
    for (size_t e = 0; e < N; e++) k[e] *= dt;
}

/* ForwardEuler::advance (src/TimeIntegration.cpp:10-12): u + dt M^-1 b(u, t), with a
 * per-component block mass solve (the reference's MassMatrix::solve indexes past its output for
 * nc > 1, src/MassMatrix.cpp:108-109) */
int o_forward_euler_step(const o_mesh *m, const o_eq *eq, int deg, const double *u, double dt, double t,
                         double *unp1)
{
    int n = m->np, bs = (deg + 1) * (deg + 2) / 2, bl = o_nc(eq) * bs;
    size_t N = (size_t)n * bl, bb2 = (size_t)bl * bl;
    double *M = malloc(sizeof(double) * bb2 * n), *k = malloc(sizeof(double) * N);
    int *piv = malloc(sizeof(int) * N);
    o_mass(m, deg, o_nc(eq), M);
    for (int i = 0; i < n; i++) lu_factor(bl, M + i * bb2, piv + (size_t)i * bl);
    explicit_stage(m, eq, deg, M, piv, u, t, dt, k);
    for (size_t e = 0; e < N; e++) unp1[e] = u[e] + k[e];
    free(M); free(k); free(piv);
    return 0;
}

/* RK2::advance (src/TimeIntegration.cpp:24-29): k1 = dt M^-1 b(u, t),
 * k2 = dt M^-1 b(u + k1/2, t + dt/2), u + (k1 + k2)/2 -- the midpoint stage with the trapezoidal
 * weights, as the reference has it (first order) */
int o_rk2_step(const o_mesh *m, const o_eq *eq, int deg, const double *u, double dt, double t, double *unp1)
{
    int n = m->np, bs = (deg + 1) * (deg + 2) / 2, bl = o_nc(eq) * bs;
    size_t N = (size_t)n * bl, bb2 = (size_t)bl * bl;
    double *M = malloc(sizeof(double) * bb2 * n), *k1 = malloc(sizeof(double) * N);
    double *k2 = malloc(sizeof(double) * N), *v = malloc(sizeof(double) * N);
    int *piv = malloc(sizeof(int) * N);
    o_mass(m, deg, o_nc(eq), M);
    for (int i = 0; i < n; i++) lu_factor(bl, M + i * bb2, piv + (size_t)i * bl);
    explicit_stage(m, eq, deg, M, piv, u, t, dt, k1);
    for (size_t e = 0; e < N; e++) v[e] = u[e] + 0.5 * k1[e];
    explicit_stage(m, eq, deg, M, piv, v, t + 0.5 * dt, dt, k2);
    for (size_t e = 0; e < N; e++) unp1[e] = u[e] + 0.5 * (k1[e] + k2[e]);
    free(M); free(k1); free(k2); free(v); free(piv);
    return 0;
}

/* RK4::advance (src/TimeIntegration.cpp:14-22) with a per-component block mass solve */
int o_rk4_step(const o_mesh *m, const o_eq *eq, int deg, const double *u, double dt, double t,
               double *unp1)
{
    int n = m->np, bs = (deg + 1) * (deg + 2) / 2, nc = o_nc(eq), bl = nc * bs;
    size_t N = (size_t)n * bl, bb2 = (size_t)bl * bl;
    double *M = malloc(sizeof(double) * bb2 * n);
    int *piv = malloc(sizeof(int) * (size_t)n * bl);
    o_mass(m, deg, nc, M);
    for (int i = 0; i < n; i++) lu_factor(bl, M + i * bb2, piv + (size_t)i * bl);
    double *kk = malloc(sizeof(double) * N * 4), *tmp = malloc(sizeof(double) * N);
    double cs[4] = {0, 0.5, 0.5, 1.0};
    for (int s = 0; s < 4; s++) {
        memcpy(tmp, u, sizeof(double) * N);
        if (s > 0)
            for (size_t e = 0; e < N; e++) tmp[e] += cs[s] * kk[(s - 1) * N + e];
        o_assemble(m, eq, deg, tmp, t + cs[s] * dt, kk + s * N);
        for (int i = 0; i < n; i++) lu_solve(bl, M + i * bb2, piv + (size_t)i * bl, kk + s * N + (size_t)i * bl);
        for (size_t e = 0; e < N; e++) kk[s * N + e] *= dt;
    }
    for (size_t e = 0; e < N; e++)
        unp1[e] = u[e] + (kk[e] + 2 * kk[N + e] + 2 * kk[2 * N + e] + kk[3 * N + e]) / 6.0;
    free(M); free(piv); free(kk); free(tmp);
    return 0;
}
