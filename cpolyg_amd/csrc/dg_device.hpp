// Device-side DG building blocks: Legendre basis, quadrature rules, Euler and
// advection fluxes with their Jacobians, Dirichlet data.  Formulas follow
// src/Legendre.h:51-68, src/Quadrature.cpp:103-171, eqns/Euler.cpp:7-179,
// eqns/Advection.cpp:8-66 and src/BoundaryConditions.cpp:7-37.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace cpg {

// Equation parameters as they travel to the kernels
struct DevEq {
    int kind;
    double beta[6];
    double gamma;
    int bc_kind;
    double bc_par[8];
    int periodic_trace;
    int jacobian;
};

template <int DEG>
struct Deg {
    static constexpr int BS = (DEG + 1) * (DEG + 2) / 2;
    // rule choice for degree 2p (Quadrature.cpp:20-27,74-81)
    static constexpr int NQL = (2 * DEG <= 2) ? 2 : (2 * DEG <= 4) ? 4 : (2 * DEG <= 6) ? 6 : 10;
    // periodic trace of the reference: Leg2D with m = int(0.5*(sqrt(4 bs + 1) - 1))
    static constexpr int MQ = DEG == 1 ? 1 : DEG == 2 ? 2 : DEG == 3 ? 2 : DEG == 4 ? 3 : 0;
};

// Gauss-Legendre line nodes/weights, values of src/Quadrature.cpp:103-125
__constant__ double c_lin2[4] = {-0.5773502691896257, 1.0, 0.5773502691896257, 1.0};
__constant__ double c_lin4[8] = {-0.3399810435848563, 0.6521451548625461, 0.3399810435848563, 0.6521451548625461,
                                 -0.8611363115940526, 0.3478548451374538, 0.8611363115940526, 0.3478548451374538};
__constant__ double c_lin6[12] = {0.6612093864662645,  0.3607615730481386, -0.6612093864662645, 0.3607615730481386,
                                  -0.2386191860831969, 0.4679139345726910, 0.2386191860831969,  0.4679139345726910,
                                  -0.9324695142031521, 0.1713244923791704, 0.9324695142031521,  0.1713244923791704};
__constant__ double c_lin10[20] = {0.1488743389816312,  0.2955242247147529, -0.1488743389816312, 0.2955242247147529,
                                   -0.4333953941292472, 0.2692667193099963, 0.4333953941292472,  0.2692667193099963,
                                   -0.6794095682990244, 0.2190863625159820, 0.6794095682990244,  0.2190863625159820,
                                   -0.8650633666889845, 0.1494513491505806, 0.8650633666889845,  0.1494513491505806,
                                   -0.9739065285171717, 0.0666713443086881, 0.9739065285171717,  0.0666713443086881};

template <int NQL>
__device__ inline const double *line_rule() {
    if constexpr (NQL == 2) return c_lin2;
    else if constexpr (NQL == 4) return c_lin4;
    else if constexpr (NQL == 6) return c_lin6;
    else return c_lin10;
}

template <int P>
__device__ inline void legendre(double x, double *Pv, double *dP) {
    Pv[0] = 1.0;
    dP[0] = 0.0;
    if constexpr (P >= 1) { Pv[1] = x; dP[1] = 1.0; }
#pragma unroll
    for (int n = 1; n < P; n++) {
        Pv[n + 1] = ((2 * n + 1) * x * Pv[n] - n * Pv[n - 1]) / (n + 1);
        dP[n + 1] = x * dP[n] + (n + 1) * Pv[n];
    }
}

// phi_k = P_a(xi) P_b(eta), k over a = 0..p (outer), b = 0..p-a (inner)
template <int DEG>
__device__ inline void basis(double xi, double eta, double *phi, double *dxi, double *deta) {
    double Px[DEG + 1], dPx[DEG + 1], Py[DEG + 1], dPy[DEG + 1];
    legendre<DEG>(xi, Px, dPx);
    legendre<DEG>(eta, Py, dPy);
    int k = 0;
#pragma unroll
    for (int a = 0; a <= DEG; a++)
#pragma unroll
        for (int b = 0; b <= DEG - a; b++) {
            phi[k] = Px[a] * Py[b];
            if (dxi) dxi[k] = dPx[a] * Py[b];
            if (deta) deta[k] = Px[a] * dPy[b];
            k++;
        }
}

template <int DEG>
__device__ inline void trace_basis(int exact, double xi, double eta, double *phi) {
    constexpr int BS = Deg<DEG>::BS;
    if (exact) {
        basis<DEG>(xi, eta, phi, nullptr, nullptr);
        return;
    }
    constexpr int MQ = Deg<DEG>::MQ;
#pragma unroll
    for (int k = 0; k < BS; k++) phi[k] = 0.0;
    double Px[MQ + 1], dPx[MQ + 1], Py[MQ + 1], dPy[MQ + 1];
    legendre<MQ - 1>(xi, Px, dPx);
    legendre<MQ - 1>(eta, Py, dPy);
    int k = 0;
#pragma unroll
    for (int a = 0; a < MQ; a++)
#pragma unroll
        for (int b = 0; b < MQ - a; b++) phi[k++] = Px[a] * Py[b];
}

template <int BS, int NC>
__device__ inline void eval_state(const double *U, const double *phi, double *out) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < BS; k++) s += U[c * BS + k] * phi[k];
        out[c] = s;
    }
}

// traces of two coefficient vectors stored interleaved, UV[2r] = u_r, UV[2r+1] = v_r
// (one 16-byte read per coefficient pair)
template <int BS, int NC>
__device__ inline void eval_state2(const double *UV, const double *phi, double *outU, double *outV) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 *p = reinterpret_cast<const d2 *>(UV);
#pragma unroll
    for (int c = 0; c < NC; c++) {
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int k = 0; k < BS; k++) {
            const d2 w = p[c * BS + k];
            su += w.x * phi[k];
            sv += w.y * phi[k];
        }
        outU[c] = su;
        outV[c] = sv;
    }
}

// ---- Euler ---------------------------------------------------------------------------
__device__ inline void euler_flux(double g, const double *U, double *Fx, double *Fy, double &c, double &u, double &v) {
    const double rho = U[0], ru = U[1], rv = U[2], rE = U[3];
    u = ru / rho;
    v = rv / rho;
    const double P = (g - 1) * (rE - 0.5 * (ru * ru + rv * rv) / rho);
    Fx[0] = ru; Fx[1] = ru * u + P; Fx[2] = ru * v; Fx[3] = (rE + P) * u;
    Fy[0] = rv; Fy[1] = ru * v; Fy[2] = rv * v + P; Fy[3] = (rE + P) * v;
    c = sqrt(g * P / rho);
}

// row-major 4x4 flux Jacobians.  ref != 0 keeps the reference's energy-row entry of
// A_x, H - (gamma-1)|u|^2 (eqns/Euler.cpp:157), where the derivative is H - (gamma-1) u^2.
__device__ inline void euler_fluxjac(double g, const double *U, double *Ax, double *Ay, int ref) {
    const double rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3] / rho;
    const double usq = u * u + v * v, e = E - 0.5 * usq, P = (g - 1) * rho * e, H = E + P / rho;
    Ax[0] = 0; Ax[1] = 1; Ax[2] = 0; Ax[3] = 0;
    Ax[4] = -u * u + 0.5 * (g - 1) * usq; Ax[5] = (3 - g) * u; Ax[6] = -(g - 1) * v; Ax[7] = g - 1;
    Ax[8] = -u * v; Ax[9] = v; Ax[10] = u; Ax[11] = 0;
    Ax[12] = u * (0.5 * (g - 1) * usq - H); Ax[13] = ref ? H - (g - 1) * usq : H - (g - 1) * u * u;
    Ax[14] = -(g - 1) * u * v; Ax[15] = g * u;
    Ay[0] = 0; Ay[1] = 0; Ay[2] = 1; Ay[3] = 0;
    Ay[4] = -u * v; Ay[5] = v; Ay[6] = u; Ay[7] = 0;
    Ay[8] = -v * v + 0.5 * (g - 1) * usq; Ay[9] = -(g - 1) * u; Ay[10] = (3 - g) * v; Ay[11] = g - 1;
    Ay[12] = v * (0.5 * (g - 1) * usq - H); Ay[13] = -(g - 1) * u * v; Ay[14] = H - (g - 1) * v * v; Ay[15] = g * v;
}

// d alpha / d U1 when the first maximal |eigenvalue| belongs to U1 (Euler::alphaDerivative)
__device__ inline void euler_alpha_deriv(double g, const double *U1, const double *U2, double nx, double ny,
                                         double &alpha, double *d) {
    const double u1 = U1[1] / U1[0], v1 = U1[2] / U1[0], u2 = U2[1] / U2[0], v2 = U2[2] / U2[0];
    const double P1 = (g - 1) * (U1[3] - 0.5 * (U1[1] * U1[1] + U1[2] * U1[2]) / U1[0]);
    const double c1 = sqrt(g * P1 / U1[0]);
    const double P2 = (g - 1) * (U2[3] - 0.5 * (U2[1] * U2[1] + U2[2] * U2[2]) / U2[0]);
    const double c2 = sqrt(g * P2 / U2[0]);
    const double vn1 = u1 * nx + v1 * ny, vn2 = u2 * nx + v2 * ny;
    const double eig[6] = {vn1 - c1, vn1, vn1 + c1, vn2 - c2, vn2, vn2 + c2};
    int idx = 0;
    double best = fabs(eig[0]);
#pragma unroll
    for (int k = 1; k < 6; k++)
        if (fabs(eig[k]) > best) { best = fabs(eig[k]); idx = k; }
    alpha = best;
    d[0] = d[1] = d[2] = d[3] = 0.0;
    if (idx <= 2) {
        const double rho = U1[0], E = U1[3] / rho;
        d[0] = -u1 / rho * nx - v1 / rho * ny;
        d[1] = 1 / rho * nx;
        d[2] = 1 / rho * ny;
        const double usq = u1 * u1 + v1 * v1;
        const double a = sqrt(g * g - g), b = rho * sqrt(4 * E - 2 * usq);
        if (idx == 0) {
            d[0] -= a * (usq - E) / b; d[1] -= -a * u1 / b; d[2] -= -a * v1 / b; d[3] -= a / b;
        } else if (idx == 2) {
            d[0] += a * (usq - E) / b; d[1] += -a * u1 / b; d[2] += -a * v1 / b; d[3] += a / b;
        }
        if (eig[idx] < 0) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; d[3] = -d[3]; }
    }
}

__device__ inline void adv_beta(const DevEq &eq, double x, double y, double &bx, double &by) {
    bx = eq.beta[0] + eq.beta[1] * x + eq.beta[2] * y;
    by = eq.beta[3] + eq.beta[4] * x + eq.beta[5] * y;
}

template <int NC>
__device__ inline void flux(const DevEq &eq, const double *U, double x, double y, double *Fx, double *Fy) {
    if constexpr (NC == 4) {
        double c, u, v;
        euler_flux(eq.gamma, U, Fx, Fy, c, u, v);
    } else {
        double bx, by;
        adv_beta(eq, x, y, bx, by);
        Fx[0] = bx * U[0];
        Fy[0] = by * U[0];
    }
}

template <int NC>
__device__ inline void numflux(const DevEq &eq, const double *Um, const double *Up, double x, double y, double nx,
                               double ny, double *F) {
    if constexpr (NC == 4) {
        double Fxm[4], Fym[4], Fxp[4], Fyp[4], cm, um, vm, cp, up, vp;
        euler_flux(eq.gamma, Um, Fxm, Fym, cm, um, vm);
        euler_flux(eq.gamma, Up, Fxp, Fyp, cp, up, vp);
        const double vnm = um * nx + vm * ny, vnp = up * nx + vp * ny;
        // std::max over the list (eqns/Euler.cpp:127-131): a NaN first element stays NaN,
        // so a state with negative pressure yields a non-finite residual as in the reference
        const double e[6] = {fabs(vnm - cm), fabs(vnm), fabs(vnm + cm), fabs(vnp - cp), fabs(vnp), fabs(vnp + cp)};
        double alpha = e[0];
#pragma unroll
        for (int k = 1; k < 6; k++)
            if (alpha < e[k]) alpha = e[k];
#pragma unroll
        for (int c = 0; c < 4; c++)
            F[c] = 0.5 * ((Fxp[c] + Fxm[c]) * nx + (Fyp[c] + Fym[c]) * ny + alpha * (Um[c] - Up[c]));
    } else {
        double bx, by;
        adv_beta(eq, x, y, bx, by);
        const double bn = bx * nx + by * ny;
        F[0] = bn > 0 ? Um[0] * bn : Up[0] * bn;
    }
}

template <int NC>
__device__ inline void fluxjac(const DevEq &eq, const double *U, double x, double y, double *Ax, double *Ay) {
    if constexpr (NC == 4) euler_fluxjac(eq.gamma, U, Ax, Ay, eq.jacobian == 0);
    else adv_beta(eq, x, y, Ax[0], Ay[0]);
}

// 0.5 (A(U) n + sgn alpha I + (U - U2) alpha'^T)  (Euler::numericalFluxJacobian)
template <int NC>
__device__ inline void numfluxjac(const DevEq &eq, const double *U, const double *U2, double x, double y, double nx,
                                  double ny, int sgn, double *J) {
    if constexpr (NC == 4) {
        double Ax[16], Ay[16], d[4], alpha;
        euler_fluxjac(eq.gamma, U, Ax, Ay, eq.jacobian == 0);
        euler_alpha_deriv(eq.gamma, U, U2, nx, ny, alpha, d);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++)
                J[4 * r + c] = 0.5 * (Ax[4 * r + c] * nx + Ay[4 * r + c] * ny + (r == c ? sgn * alpha : 0.0) +
                                      (U[r] - U2[r]) * d[c]);
    } else {
        double bx, by;
        adv_beta(eq, x, y, bx, by);
        const double bn = nx * bx + ny * by;
        J[0] = (sgn * bn > 0) ? bn : 0.0;
    }
}

// Derivative of the residual's LF speed with respect to the outer state U+: the speed is
// the ordered maximum of (|vn-c|, |vn|, |vn+c|) of U- then of U+ (NaN-skipping exactly as
// the residual's), and only when the maximum is one of U+'s three does it depend on U+.
__device__ inline void euler_alpha_deriv_plus(double g, const double *Um, const double *Up, double nx, double ny,
                                              double &alpha, double *d) {
    const double u1 = Um[1] / Um[0], v1 = Um[2] / Um[0], u2 = Up[1] / Up[0], v2 = Up[2] / Up[0];
    const double P1 = (g - 1) * (Um[3] - 0.5 * (Um[1] * Um[1] + Um[2] * Um[2]) / Um[0]);
    const double P2 = (g - 1) * (Up[3] - 0.5 * (Up[1] * Up[1] + Up[2] * Up[2]) / Up[0]);
    const double c1 = sqrt(g * P1 / Um[0]), c2 = sqrt(g * P2 / Up[0]);
    const double vn1 = u1 * nx + v1 * ny, vn2 = u2 * nx + v2 * ny;
    const double eig[6] = {vn1 - c1, vn1, vn1 + c1, vn2 - c2, vn2, vn2 + c2};
    int idx = 0;
    alpha = fabs(eig[0]);
#pragma unroll
    for (int k = 1; k < 6; k++)
        if (alpha < fabs(eig[k])) { alpha = fabs(eig[k]); idx = k; }
    d[0] = d[1] = d[2] = d[3] = 0.0;
    if (idx >= 3) {
        const double rho = Up[0], E = Up[3] / rho;
        d[0] = -u2 / rho * nx - v2 / rho * ny;
        d[1] = 1 / rho * nx;
        d[2] = 1 / rho * ny;
        if (idx != 4) {
            const double usq = u2 * u2 + v2 * v2;
            const double a = sqrt(g * g - g), b = rho * sqrt(4 * E - 2 * usq);
            const double s = idx == 3 ? -1.0 : 1.0;
            d[0] += s * a * (usq - E) / b; d[1] -= s * a * u2 / b; d[2] -= s * a * v2 / b; d[3] += s * a / b;
        }
        if (eig[idx] < 0) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; d[3] = -d[3]; }
    }
}

// exact derivative of the LF flux with respect to the outer state
__device__ inline void euler_offdiag_exact(double g, const double *Um, const double *Up, double nx, double ny,
                                           double *J) {
    double Bx[16], By[16], d[4], alpha;
    euler_fluxjac(g, Up, Bx, By, 0);
    euler_alpha_deriv_plus(g, Um, Up, nx, ny, alpha, d);
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
            J[4 * r + c] = 0.5 * (Bx[4 * r + c] * nx + By[4 * r + c] * ny - (r == c ? alpha : 0.0) + (Um[r] - Up[r]) * d[c]);
}

// Euler state primitives for the flux derivatives: one reciprocal of the density
struct EulerPrim {
    double ir, u, v, E, k, H;   // 1/rho, velocity, specific energy, (g-1)|u|^2/2, enthalpy
};
__device__ inline EulerPrim euler_prim(double g, const double *U) {
    EulerPrim p;
    p.ir = 1.0 / U[0];
    p.u = U[1] * p.ir;
    p.v = U[2] * p.ir;
    p.E = U[3] * p.ir;
    const double usq = p.u * p.u + p.v * p.v;
    p.k = 0.5 * (g - 1) * usq;
    p.H = p.E + (g - 1) * (p.E - 0.5 * usq);
    return p;
}

// ---- state-dependent data of the Euler Newton operator at the quadrature points ---------
// The Newton matrix M - dt a J(u) is applied many times for one u (every GMRES iteration of
// a Newton step), so what depends on u alone is evaluated once per quadrature point
// (k_linprep) and the operator kernel combines it with the traces of v:
//   volume point: the primitives p = (u, v, k, H) of U, with k = (gamma-1)|u|^2/2 and H the
//                 enthalpy (the entries of euler_fluxjac are polynomials in them);
//   edge point:   p(U-), p(U+), U- - U+, the LF speed alpha, and alpha's derivative with respect
//                 to U- and to U+ (one of the two is zero: alpha belongs to one state).
constexpr int LPV = 4;    // doubles per volume point
constexpr int LPE = 17;   // doubles per edge point: pm[4] pp[4] dU[4] alpha d[4] (euler_edge_prep)

__device__ inline void euler_prim4(double g, const double *U, double *p) {
    const EulerPrim q = euler_prim(g, U);
    p[0] = q.u;
    p[1] = q.v;
    p[2] = q.k;
    p[3] = q.H;
}

// A_x(U) V and A_y(U) V from p = (u, v, k, H): the products with the matrices of
// euler_fluxjac (ref: the reference's energy-row entry H - (gamma-1)|u|^2 of A_x)
__device__ inline void euler_jvp_p(double g, const double *p, const double *V, int ref, double *fx, double *fy) {
    const double gm = g - 1, u = p[0], v = p[1], k = p[2], H = p[3], uv = u * v;
    const double r2 = -uv * V[0] + v * V[1] + u * V[2];
    fx[0] = V[1];
    fx[1] = (k - u * u) * V[0] + (3 - g) * u * V[1] - gm * v * V[2] + gm * V[3];
    fx[2] = r2;
    fx[3] = u * (k - H) * V[0] + (ref ? H - 2 * k : H - gm * u * u) * V[1] - gm * uv * V[2] + g * u * V[3];
    fy[0] = V[2];
    fy[1] = r2;
    fy[2] = (k - v * v) * V[0] - gm * u * V[1] + (3 - g) * v * V[2] + gm * V[3];
    fy[3] = v * (k - H) * V[0] - gm * uv * V[1] + (H - gm * v * v) * V[2] + g * v * V[3];
}

// (nx A_x + ny A_y) V from p
__device__ inline void euler_an_v_p(double g, const double *p, const double *V, int ref, double nx, double ny,
                                    double *f) {
    double fx[4], fy[4];
    euler_jvp_p(g, p, V, ref, fx, fy);
#pragma unroll
    for (int c = 0; c < 4; c++) f[c] = nx * fx[c] + ny * fy[c];
}

// The u-dependent edge-point data: LF speed alpha = first maximum of |vn-c|, |vn|, |vn+c| of
// Um, then of Up (the tie rule of euler_alpha_deriv / euler_alpha_deriv_plus), and d alpha
// with respect to the state that owns it (Euler::alphaDerivative), written to rec[LPE]: the
// derivative once, its side in the sign of the stored alpha (alpha >= 0; -alpha, -0.0 included,
// when Up owns it)
__device__ inline void euler_edge_prep(double g, const double *Um, const double *Up, double nx, double ny,
                                       double *rec) {
    const EulerPrim pm = euler_prim(g, Um), pp = euler_prim(g, Up);
    const double c1 = sqrt(g * (g - 1) * (pm.E - 0.5 * (pm.u * pm.u + pm.v * pm.v)));
    const double c2 = sqrt(g * (g - 1) * (pp.E - 0.5 * (pp.u * pp.u + pp.v * pp.v)));
    const double vn1 = pm.u * nx + pm.v * ny, vn2 = pp.u * nx + pp.v * ny;
    const double eig[6] = {vn1 - c1, vn1, vn1 + c1, vn2 - c2, vn2, vn2 + c2};
    int idx = 0;
    double e = eig[0], alpha = fabs(e);
#pragma unroll
    for (int k = 1; k < 6; k++)
        if (alpha < fabs(eig[k])) { alpha = fabs(eig[k]); idx = k; e = eig[k]; }
    const int side = idx >= 3, j = idx - 3 * side;
    const double ir = side ? pp.ir : pm.ir, uu = side ? pp.u : pm.u, vv = side ? pp.v : pm.v;
    const double E = side ? pp.E : pm.E;
    double d[4] = {-(uu * nx + vv * ny) * ir, nx * ir, ny * ir, 0.0};
    if (j != 1) {
        const double usq = uu * uu + vv * vv;
        const double sc = (j == 0 ? -1.0 : 1.0) * sqrt(g * g - g) * ir / sqrt(4 * E - 2 * usq);
        d[0] += sc * (usq - E); d[1] -= sc * uu; d[2] -= sc * vv; d[3] += sc;
    }
    if (e < 0) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; d[3] = -d[3]; }
    rec[0] = pm.u; rec[1] = pm.v; rec[2] = pm.k; rec[3] = pm.H;
    rec[4] = pp.u; rec[5] = pp.v; rec[6] = pp.k; rec[7] = pp.H;
#pragma unroll
    for (int c = 0; c < 4; c++) rec[8 + c] = Um[c] - Up[c];
    rec[12] = side ? -alpha : alpha;
#pragma unroll
    for (int c = 0; c < 4; c++) rec[13 + c] = d[c];
}

// Derivative of the Euler numerical flux along (Vm, Vp) from the edge-point data: Jd Vm + Jo Vp
// with the matrices k_jacobian uses (Jd = numfluxjac(+1); Jo = euler_offdiag_exact for the exact
// Jacobian form, numfluxjac(-1), i.e. A(U-) and alpha' with respect to U-, for the reference
// form).  Vp = 0 on Dirichlet edges.
__device__ inline void euler_numflux_jvp_p(const DevEq &eq, const double *rec, const double *Vm, const double *Vp,
                                           double nx, double ny, double *f) {
    const double g = eq.gamma, alpha = fabs(rec[12]);
    const bool plus = signbit(rec[12]);   // alpha's derivative is with respect to Up
    const double *dU = rec + 8, *d = rec + 13;
    double fm[4], fp[4], dv = 0.0;
    if (eq.jacobian) {
        euler_an_v_p(g, rec, Vm, 0, nx, ny, fm);
        euler_an_v_p(g, rec + 4, Vp, 0, nx, ny, fp);
#pragma unroll
        for (int c = 0; c < 4; c++) dv = fma(d[c], plus ? Vp[c] : Vm[c], dv);
    } else {
        double Vs[4];
#pragma unroll
        for (int c = 0; c < 4; c++) Vs[c] = Vm[c] + Vp[c];
        euler_an_v_p(g, rec, Vs, 1, nx, ny, fm);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            fp[c] = 0.0;
            dv = plus ? dv : fma(d[c], Vs[c], dv);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) f[c] = 0.5 * (fm[c] + fp[c] + alpha * (Vm[c] - Vp[c]) + dU[c] * dv);
}

// advection: upwind flux derivative (numfluxjac for nc = 1)
__device__ inline double adv_numflux_jvp(const DevEq &eq, double Vm, double Vp, double x, double y, double nx,
                                         double ny) {
    double bx, by;
    adv_beta(eq, x, y, bx, by);
    const double bn = bx * nx + by * ny;
    return bn > 0 ? bn * Vm : bn * Vp;
}

// Dirichlet data
template <int NC>
__device__ inline void eval_bc(const DevEq &eq, double x, double y, double t, double *out) {
    const double *p = eq.bc_par;
    if (eq.bc_kind == 1) {
#pragma unroll
        for (int c = 0; c < NC; c++) out[c] = p[c];
        return;
    }
    if constexpr (NC == 4) {
        if (eq.bc_kind == 2) {
            const double g = eq.gamma, beta = p[0];
            const double dx = x - p[1] - p[3] * t, dy = y - p[2] - p[4] * t;
            const double ex = exp(0.5 * (1 - (dx * dx + dy * dy)));
            const double u = p[3] - beta / (2 * M_PI) * dy * ex;
            const double v = p[4] + beta / (2 * M_PI) * dx * ex;
            const double T = 1 - (g - 1) * beta * beta / (8 * g * M_PI * M_PI) * ex * ex;
            const double rho = pow(T, 1 / (g - 1));
            const double P = rho * T;
            out[0] = rho; out[1] = rho * u; out[2] = rho * v; out[3] = P / (g - 1) + 0.5 * rho * (u * u + v * v);
            return;
        }
        if (eq.bc_kind == 3) {
            const double g = eq.gamma, x0 = 5, y0 = 5, theta = atan2(1.0, 2.0), eps = 0.3, rc = 1.5, Minf = 0.5;
            const double ub = cos(theta), vb = sin(theta);
            const double f = (1 - pow((x - x0) - ub * t, 2) - pow((y - y0) - vb * t, 2)) / (rc * rc);
            const double u = cos(theta) - eps * ((y - y0) - vb * t) / (2 * M_PI * rc) * exp(f * 0.5);
            const double v = sin(theta) + eps * ((x - x0) - ub * t) / (2 * M_PI * rc) * exp(f * 0.5);
            const double base = 1 - eps * eps * (g - 1) * (Minf * Minf / (8 * M_PI * M_PI)) * exp(f);
            const double rho = pow(base, 1 / (g - 1));
            const double P = pow(base, 1 / (g - 1));
            out[0] = rho; out[1] = rho * u; out[2] = rho * v; out[3] = P / (g - 1) + 0.5 * rho * (u * u + v * v);
            return;
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) out[c] = 0.0;
}

}  // namespace cpg
