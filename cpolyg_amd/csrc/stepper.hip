// Time steppers with the Newton-Krylov loop on the device.
//
// DIRK::advance (src/TimeIntegration.cpp:93-151), BackwardEuler::advance (:31-66)
// in increment form, an SDIRK2 tableau (not in the reference), and RK4::advance
// (:14-22).  Per Newton iteration of a stage:
//   cu = u + sum_{j<=s} A_sj k_j                 (one fused vector kernel)
//   r  = M k_s - dt b(cu, t + c_s dt)            (residual kernel + mass kernel)
//   stop when max_c ||r_c||_L2 < tol             (||r_c||^2 = sum_i r_ic^T M_i r_ic; one D2H read)
//   B  = M - dt A_ss J(cu)                       (Jacobian kernel, written into BSR slots)
//   factor the preconditioner of B, GMRES B x = r from x = 0, k_s -= x.
#include "cpg_internal.hpp"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

struct cpg_equation_s;

namespace cpg {
void eq_residual(const cpg_equation_s &E, const double *u, double t, double *b, hipStream_t s);
void eq_jacobian(const cpg_equation_s &E, const double *u, double t, double scale, int add_mass, double *J,
                 hipStream_t s);
void eq_mass_matvec(const cpg_equation_s &E, const double *x, double *y, hipStream_t s);
double eq_newton_resid(const cpg_equation_s &E, const double *kv, const double *b, double dt, double *r,
                       double *out_nc, hipStream_t s);
void krylov_timing(cpg_krylov K, bool on);
void krylov_timing_read(cpg_krylov K, int *nm, double *msm, int *na, double *msa);
void eq_newton_prepare(cpg_equation_s &E, const double *u, double t, hipStream_t s, double *vol = nullptr,
                       double *edg = nullptr);
void eq_newton_data_size(const cpg_equation_s &E, size_t *nvol, size_t *nedg);
void eq_newton_apply_dev(const cpg_equation_s &E, const double *scale_dev, const double *v, double *y,
                         const int *active, hipStream_t s, double scale = 0.0, const double *xm = nullptr,
                         const double *vol = nullptr, const double *edg = nullptr);

// the Newton matrix M - dt a J(cu) applied at the quadrature points, from the point data of
// the last eq_newton_prepare; the scale stays in device memory, so captured Krylov cycles are
// replayed for every stage and Newton iteration
struct NewtonOp {
    const cpg_equation_s *E;
    const double *scale;
};
void newton_linop(void *ctx, const double *x, double *y, const int *active, hipStream_t s) {
    const auto *o = static_cast<const NewtonOp *>(ctx);
    eq_newton_apply_dev(*o->E, o->scale, x, y, active, s);
}

__global__ void k_set_scalar(double *p, double v) { *p = v; }

// ---- implicit Runge-Kutta with the matrix-free coupled operator ------------------------
// The coupled Newton system of IRK::advance (src/TimeIntegration.cpp:188-281) in its
// stage-major layout: y_s = M x_s - dt sum_t A[s][t] J(cu_s) x_t.  Row stage s is the Newton
// operator of stage s's state applied to w_s = sum_t A[s][t] x_t, with its mass term on x_s.
// Preconditioner: block Jacobi on the cell-interleaved coupled diagonal blocks (one
// (ns bl) x (ns bl) block per cell, D_i[(s,.),(t,.)] = delta_st M_i - dt A[s][t] J_ii(cu_s)),
// applied on an interleaved copy of the vector.

struct IrkCoef3 {
    double a[3];
};

// w = sum_t a[t] x_t  (stage-major x, ns stages of N)
__global__ void k_irk_combine(IrkCoef3 c, int ns, const double *__restrict__ x, size_t N, double *__restrict__ w,
                              const int *__restrict__ active) {
    if (active && !*active) return;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    double v = 0.0;
    for (int t = 0; t < ns; t++) v = fma(c.a[t], x[(size_t)t * N + e], v);
    w[e] = v;
}

// stage-major [s][i][r] <-> cell-interleaved [i][s][r] (forward: to interleaved)
__global__ void k_irk_interleave(int ns, int n, int bl, const double *__restrict__ in, double *__restrict__ out,
                                 int forward, const int *__restrict__ active) {
    if (active && !*active) return;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // index in the interleaved vector
    const size_t tot = (size_t)ns * n * bl;
    if (e >= tot) return;
    const size_t i = e / ((size_t)ns * bl);
    const int rem = (int)(e - i * ns * bl), st = rem / bl, r = rem - st * bl;
    const size_t sm = ((size_t)st * n + i) * bl + r;
    if (forward) out[e] = in[sm];
    else out[sm] = in[e];
}

// row stripe s of every coupled diagonal block (column-major, side ns bl) from the diagonal
// blocks of J(cu_s) (pattern: diagonal first in each row) and the mass blocks
__global__ void k_irk_diag_scatter(int ns, int s, int n, int bl, int bs, const int *__restrict__ rowptr,
                                   const double *__restrict__ J, const double *__restrict__ mass, IrkCoef3 adt,
                                   double *__restrict__ D) {
    const int L = ns * bl;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * bl * L) return;
    const size_t i = e / ((size_t)bl * L);
    const int rem = (int)(e - i * bl * L);
    const int col = rem / bl, r = rem - col * bl;   // column of the coupled block, row within stripe s
    const int t = col / bl, c = col - t * bl;
    const double *Jii = J + (size_t)rowptr[i] * bl * bl;
    double v = adt.a[t] * Jii[(size_t)c * bl + r];
    if (t == s && r / bs == c / bs) v += mass[(size_t)i * bs * bs + (c % bs) * bs + (r % bs)];
    D[(size_t)i * L * L + (size_t)col * L + s * bl + r] = v;
}

struct IrkOp {
    const cpg_equation_s *E;
    int ns, n, bl;
    size_t N;
    double A[9];
    const double *negdt;          // -dt in device memory
    double *w;                    // N scratch
    double *vol[3], *edg[3];      // point data of each stage's state
};
void irk_linop(void *ctx, const double *x, double *y, const int *active, hipStream_t s) {
    const auto *o = static_cast<const IrkOp *>(ctx);
    for (int st = 0; st < o->ns; st++) {
        IrkCoef3 c{};
        for (int t = 0; t < o->ns; t++) c.a[t] = o->A[st * o->ns + t];
        hipLaunchKernelGGL(k_irk_combine, dim3(ceil_div(o->N, 256)), dim3(256), 0, s, c, o->ns, x, o->N, o->w, active);
        CPG_LAUNCH_CHECK();
        eq_newton_apply_dev(*o->E, o->negdt, o->w, y + o->N * st, active, s, 0.0, x + o->N * st, o->vol[st],
                            o->edg[st]);
    }
}

struct IrkPc {
    cpg_prec P;
    int ns, n, bl;
    double *pin, *pout;
};
void irk_pcop(void *ctx, const double *in, double *out, const int *active, hipStream_t s) {
    const auto *o = static_cast<const IrkPc *>(ctx);
    const size_t tot = (size_t)o->ns * o->n * o->bl;
    hipLaunchKernelGGL(k_irk_interleave, dim3(ceil_div(tot, 256)), dim3(256), 0, s, o->ns, o->n, o->bl, in, o->pin, 1,
                       active);
    CPG_LAUNCH_CHECK();
    prec_apply_to(o->P, o->pin, o->pout, s, active);
    hipLaunchKernelGGL(k_irk_interleave, dim3(ceil_div(tot, 256)), dim3(256), 0, s, o->ns, o->n, o->bl, o->pout, out,
                       0, active);
    CPG_LAUNCH_CHECK();
}

// ---- Radau IIA in the W form (IRK::newAdvance, src/TimeIntegration.cpp:283-395) -----------
// Unknowns w_s (stage states u + w_s); row s of the Newton matrix applied to x is
// M (sum_j Ainv[s][j] x_j) - dt J(u + w_s) x_s: the Newton operator of stage s's state on x_s
// with its mass term on the combination.  Preconditioner: the factor of each stage's diagonal
// block Ainv[s][s] M - dt J_s on its own part of the vector.
struct IrkWOp {
    const cpg_equation_s *E;
    int ns;
    size_t N;
    double Ainv[9];
    const double *negdt;          // -dt in device memory
    double *z;                    // N scratch: the mass term's vector
    double *vol[3], *edg[3];      // point data of each stage's state
};
void irkw_linop(void *ctx, const double *x, double *y, const int *active, hipStream_t s) {
    const auto *o = static_cast<const IrkWOp *>(ctx);
    for (int st = 0; st < o->ns; st++) {
        IrkCoef3 c{};
        for (int j = 0; j < o->ns; j++) c.a[j] = o->Ainv[st * o->ns + j];
        hipLaunchKernelGGL(k_irk_combine, dim3(ceil_div(o->N, 256)), dim3(256), 0, s, c, o->ns, x, o->N, o->z, active);
        CPG_LAUNCH_CHECK();
        eq_newton_apply_dev(*o->E, o->negdt, x + o->N * st, y + o->N * st, active, s, 0.0, o->z, o->vol[st],
                            o->edg[st]);
    }
}

struct IrkWPc {
    cpg_prec P[3];
    int ns;
    size_t N;
};
void irkw_pcop(void *ctx, const double *in, double *out, const int *active, hipStream_t s) {
    const auto *o = static_cast<const IrkWPc *>(ctx);
    for (int st = 0; st < o->ns; st++) prec_apply_to(o->P[st], in + o->N * st, out + o->N * st, s, active);
}

struct LinComb {
    const double *v[6];
    double a[6];
    int nv;
};

// out = sum_j a_j v_j
__global__ void k_lincomb(LinComb L, double *__restrict__ out, size_t N) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    double s = 0.0;
    for (int j = 0; j < L.nv; j++) s = fma(L.a[j], L.v[j][e], s);
    out[e] = s;
}

// x += a * y
__global__ void k_axpy(double *__restrict__ x, double a, const double *__restrict__ y, size_t N) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) x[e] = fma(a, y[e], x[e]);
}

// y_i = Minv_i x_i blockwise per component (bs x bs blocks)
__global__ void k_block_apply(int n, int bs, int nc, const double *__restrict__ Mi, const double *__restrict__ x,
                              double *__restrict__ y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t bl = (size_t)bs * nc;
    if (e >= (size_t)n * bl) return;
    const size_t i = e / bl;
    const int r = (int)(e - i * bl);
    const int c = r / bs, k = r - c * bs;
    const double *M = Mi + i * bs * bs;
    const double *xi = x + i * bl + c * bs;
    double s = 0.0;
    for (int l = 0; l < bs; l++) s = fma(M[l * bs + k], xi[l], s);
    y[e] = s;
}

// Stage-coupled Newton matrix of IRK::advance (src/TimeIntegration.cpp:233-245) in the
// stage-major layout of BlockMatrix::blockBlockMatrix (src/BlockMatrix.cpp:228-258):
// big row (si, i) holds, for sj = 0..ns-1, the blocks of Jacobian row i scaled by
// -dt A[si][sj], plus the mass block on (si, i, si, i).  This kernel writes the row-stage
// si from the stage-si Jacobian J (Jacobian pattern, diagonal block first in each row).
struct IrkCoef {
    double c[3];
};

__global__ void k_irk_scatter(int ns, int si, int nnzb, int bl, int bs, const int *__restrict__ rowptr,
                              const int *__restrict__ slot_row, const double *__restrict__ J, IrkCoef coef,
                              const double *__restrict__ mass, double *__restrict__ big) {
    const size_t bb = (size_t)bl * bl;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)ns * nnzb * bb) return;
    const int sj = (int)(e / ((size_t)nnzb * bb));
    const size_t rem = e - (size_t)sj * nnzb * bb;
    const int s = (int)(rem / bb);
    const int off = (int)(rem - (size_t)s * bb);
    const int i = slot_row[s];
    const int q = s - rowptr[i], len = rowptr[i + 1] - rowptr[i];
    const size_t dst = ((size_t)si * ns * nnzb + (size_t)ns * rowptr[i] + (size_t)sj * len + q) * bb + off;
    double v = coef.c[sj] * J[(size_t)s * bb + off];
    if (sj == si && q == 0) {
        const int r = off % bl, c = off / bl;
        if (r / bs == c / bs) v += mass[(size_t)i * bs * bs + (c % bs) * bs + (r % bs)];
    }
    big[dst] = v;
}

void tableau(int method, int ns, double *A, double *b, double *c, int dirk3) {
    std::memset(A, 0, sizeof(double) * 25);
    if (method == CPG_STEP_IRK) {   // Radau IIA (src/TimeIntegration.cpp:153-186)
        if (ns == 2) {
            const double a[4] = {5 / 12.0, -1 / 12.0, 3 / 4.0, 1 / 4.0};
            std::memcpy(A, a, sizeof a);
            b[0] = 3 / 4.0; b[1] = 1 / 4.0; c[0] = 1 / 3.0; c[1] = 1;
            return;
        }
        if (ns == 3) {
            const double s6 = std::sqrt(6.0);
            const double a[9] = {(88 - 7 * s6) / 360.0, (296 - 169 * s6) / 1800.0, (-2 + 3 * s6) / 225.0,
                                 (296 + 169 * s6) / 1800.0, (88 + 7 * s6) / 360.0, (-2 - 3 * s6) / 225.0,
                                 (16 - s6) / 36.0, (16 + s6) / 36.0, 1 / 9.0};
            std::memcpy(A, a, sizeof a);
            b[0] = (16 - s6) / 36.0; b[1] = (16 + s6) / 36.0; b[2] = 1 / 9.0;
            c[0] = (4 - s6) / 10.0; c[1] = (4 + s6) / 10.0; c[2] = 1;
            return;
        }
    }
    if (method == CPG_STEP_DIRK) {
        if (ns == 1) { A[0] = 1; b[0] = 1; c[0] = 1; return; }
        if (ns == 2) {
            const double g = 1.0 - 1.0 / std::sqrt(2.0);
            A[0] = g; A[1] = 0; A[2] = 1 - g; A[3] = g;
            b[0] = 1 - g; b[1] = g; c[0] = g; c[1] = 1;
            return;
        }
        if (ns == 3 && dirk3 == CPG_DIRK3_L_STABLE) {
            // Alexander's L-stable SDIRK3 (SIAM J. Numer. Anal. 14 (1977) 1006-1021), not in the
            // reference: g is the root in (1/6, 1/2) of x^3 - 3x^2 + 3x/2 - 1/6; the last row of A is b
            // (stiffly accurate), so R(-inf) = 0 and |R(iy)| <= 1
            const double g = 0.43586652150845899942;
            const double b1 = -(6 * g * g - 16 * g + 1) / 4.0, b2 = (6 * g * g - 20 * g + 5) / 4.0;
            const double a[9] = {g, 0, 0, (1 - g) / 2.0, g, 0, b1, b2, g};
            std::memcpy(A, a, sizeof a);
            b[0] = b1; b[1] = b2; b[2] = g;
            c[0] = g; c[1] = (1 + g) / 2.0; c[2] = 1;
            return;
        }
        if (ns == 3) {
            // the reference's tableau (src/TimeIntegration.cpp:71-77): Crouzeix's form with
            // Alexander's gamma as alpha, second order and not A-stable (R(-inf) = -1.78)
            const double al = 0.435866521508459;
            const double a[9] = {(1 + al) / 2.0, 0, 0, -al / 2.0, (1 + al) / 2.0, 0, 1 + al, -(1 + 2 * al), (1 + al) / 2.0};
            std::memcpy(A, a, sizeof a);
            b[0] = 1.0 / 6.0 / (al * al); b[1] = 1 - 1.0 / 3.0 / (al * al); b[2] = 1.0 / 6.0 / (al * al);
            c[0] = (1 + al) / 2.0; c[1] = 0.5; c[2] = (1 - al) / 2.0;
            return;
        }
        if (ns == 5) {  // Cooper and Sayfy (1979)
            const double s6 = std::sqrt(6.0), g = (6 - s6) / 10.0;
            const double a[25] = {g, 0, 0, 0, 0,
                                  (-6 + 5 * s6) / 14.0, g, 0, 0, 0,
                                  (888 + 607 * s6) / 2850.0, (126 - 161 * s6) / 1425.0, g, 0, 0,
                                  (3153 - 3082 * s6) / 14250.0, (3213 + 1148 * s6) / 28500.0, (-267 + 88 * s6) / 500.0, g, 0,
                                  (-32583 + 14638 * s6) / 71250.0, (-17199 + 364 * s6) / 142500.0,
                                  (1329 - 544 * s6) / 2500.0, (-96 + 131 * s6) / 625.0, g};
            std::memcpy(A, a, sizeof a);
            b[0] = 0; b[1] = 0; b[2] = 1 / 9.0; b[3] = (16 - s6) / 36.0; b[4] = (16 + s6) / 36.0;
            c[0] = g; c[1] = (6 + 9 * s6) / 35.0; c[2] = 1; c[3] = (4 - s6) / 10.0; c[4] = (4 + s6) / 10.0;
            return;
        }
    }
    if (method == CPG_STEP_RK4) {
        b[0] = 1.0 / 6; b[1] = 1.0 / 3; b[2] = 1.0 / 3; b[3] = 1.0 / 6;
        c[0] = 0; c[1] = 0.5; c[2] = 0.5; c[3] = 1.0;
        return;
    }
    if (method == CPG_STEP_RK2) {   // RK2::advance (src/TimeIntegration.cpp:24-29)
        b[0] = 0.5; b[1] = 0.5;
        c[0] = 0; c[1] = 0.5;
        return;
    }
    if (method == CPG_STEP_FORWARD_EULER) {   // ForwardEuler::advance (src/TimeIntegration.cpp:10-12)
        b[0] = 1.0;
        c[0] = 0;
        return;
    }
    CPG_CHECK(false, CPG_ERR_UNSUPPORTED, "method %d with %d stages", method, ns);
}

// the explicit methods: stage input u + c_q k_{q-1}, one block mass solve per stage
inline bool explicit_method(int method) {
    return method == CPG_STEP_RK4 || method == CPG_STEP_RK2 || method == CPG_STEP_FORWARD_EULER;
}
inline int explicit_stages(int method) {
    return method == CPG_STEP_RK4 ? 4 : method == CPG_STEP_RK2 ? 2 : 1;
}

using clk = std::chrono::steady_clock;
inline double ms_since(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

}  // namespace cpg

using namespace cpg;

struct cpg_stepper_s {
    cpg_equation E = nullptr;
    int method = 0, ns = 0, n = 0, bl = 0;
    size_t N = 0;
    cpg_solver_params sp{};
    double A[25], b[5], c[5];
    cpg_bsr J = nullptr;
    cpg_bsr big = nullptr;     // IRK stage-coupled matrix
    cpg_prec P = nullptr;
    cpg_krylov K = nullptr;
    DevBuf<double> k, cu, rhs, r, x, minv, bigR, bigK;
    DevBuf<int> slot_row;
    DevBuf<double> opscale;    // -dt a_ss of the matrix-free Newton operator (IRK: -dt)
    // IRK with the matrix-free coupled operator
    cpg_bsr Dc = nullptr;      // coupled diagonal blocks, one per cell (block Jacobi)
    bool irk_mf = false;       // the matrix-free coupled path is set up
    cpg_bsr Jst[2] = {nullptr, nullptr};   // coupled DILU: the Jacobians of stages 1, 2 (J: stage 0)
    cpg_bsr Cpat = nullptr;    // coupled DILU: the (ns bl)-wide pattern (no blocks)
    // IRK in the W form
    cpg_prec Pw[3] = {nullptr, nullptr, nullptr};   // factors of the stage diagonal blocks
    double Ainv[9] = {};       // A^-1 (row-major, ns x ns)
    double btAinv[3] = {};     // b^T A^-1: u_{n+1} = u + sum_s btAinv[s] w_s
    DevBuf<double> wv, zbuf;   // w (ns N), the operator's mass-term vector (N)
    cpg::IrkWOp irkwop{};
    cpg::IrkWPc irkwpc{};
    DevBuf<double> irk_data, pin, pout, wbuf;
    cpg::IrkOp irkop{};
    cpg::IrkPc irkpc{};
    bool factored = false;     // the preconditioner holds a factor of this step (pc_refresh 2, 3)
    // pc_refresh 3: the factor is kept across steps and renewed when a step's GMRES iteration
    // count exceeds 1.5 x (+2) that of the first step after the last renewal, or dt changes
    double dt_factored = 0.0;
    int its_ref = 0, last_its = 0;
    bool refreshed = false;    // a factor was formed during the current step
    cpg::NewtonOp nop{};
    // CPOLYG_VERBOSE=1 in the environment at creation: one stderr line per Newton iteration,
    // the progress the reference prints (TimeIntegration.cpp:209-264)
    bool verbose = false;
    ~cpg_stepper_s() {
        if (P) cpg_prec_destroy(P);
        if (K) cpg_krylov_destroy(K);
        if (J) cpg_bsr_destroy(J);
        if (big) cpg_bsr_destroy(big);
        if (Dc) cpg_bsr_destroy(Dc);
        for (cpg_bsr j : Jst)
            if (j) cpg_bsr_destroy(j);
        for (cpg_prec p : Pw)
            if (p) cpg_prec_destroy(p);
        if (Cpat) cpg_bsr_destroy(Cpat);
    }
};

namespace cpg {
const double *equation_mass(const cpg_equation_s &E);
int equation_bs(const cpg_equation_s &E);
}  // namespace cpg

namespace {

void lincomb(cpg_stepper_s &S, double *out, const double *u, const double *coef, int nk, const double *kb,
             hipStream_t s) {
    LinComb L{};
    L.nv = 0;
    L.v[L.nv] = u;
    L.a[L.nv++] = 1.0;
    for (int j = 0; j < nk; j++) {
        if (coef[j] == 0.0) continue;
        L.v[L.nv] = kb + S.N * j;
        L.a[L.nv++] = coef[j];
    }
    hipLaunchKernelGGL(k_lincomb, dim3(ceil_div(S.N, 256)), dim3(256), 0, s, L, out, S.N);
    CPG_LAUNCH_CHECK();
}

int check(int st) {
    if (st) throw Error{st};
    return st;
}

// The state a step returns is checked before the step reports success.  The residual at u_{n+1}
// evaluates the state at every quadrature point and the Lax-Friedrichs sound speed at every edge
// point, so a NaN or infinity anywhere, or a state whose sound speed is not real, makes its norm
// non-finite, and the step that produced it fails instead of the one after it.
void check_state(cpg_stepper_s &S, const double *unp1, double t_end, hipStream_t s, cpg_step_stats &st) {
    const auto t0 = clk::now();
    eq_residual(*S.E, unp1, t_end, S.rhs.p, s);
    const double rn = eq_newton_resid(*S.E, unp1, S.rhs.p, 1.0, S.r.p, nullptr, s);
    const double ms = ms_since(t0);
    st.ms_residual += ms;
    st.ms_total += ms;
    CPG_CHECK(std::isfinite(rn), CPG_ERR_NONFINITE,
              "the step to t = %.17g returned a non-physical state (the residual at u_{n+1} is not finite)", t_end);
}

// re-raise a failed library call with the stage / Newton iteration it happened in
void check_ctx(int st, int stage, int it, double resid) {
    if (!st) return;
    std::string msg = cpg_last_error();
    set_error("stage %d, Newton iteration %d (residual %.3e): %s", stage, it, resid, msg.c_str());
    throw Error{st};
}

// start of a step: whether the preconditioner's factor is kept (pc_refresh 3) or formed anew
void begin_step(cpg_stepper_s &S, double dt) {
    S.refreshed = false;
    if (S.sp.pc_refresh != 3) {
        S.factored = false;
        return;
    }
    if (!S.factored || dt != S.dt_factored || (S.its_ref > 0 && S.last_its > 1.5 * S.its_ref + 2))
        S.factored = false;
}

// a factor was formed in this step
void mark_factored(cpg_stepper_s &S, double dt) {
    S.factored = true;
    S.refreshed = true;
    S.dt_factored = dt;
}

// end of a converged step: its GMRES iterations, the reference count after a renewal
void end_step(cpg_stepper_s &S, int gmres_its) {
    S.last_its = gmres_its;
    if (S.refreshed) S.its_ref = gmres_its;
}

}  // namespace

extern "C" {

int cpg_stepper_create(cpg_equation E, int method, int nstages, const cpg_solver_params *params, cpg_stepper *out) {
    return guard([&] {
        CPG_CHECK(E && params && out, CPG_ERR_ARG, "null argument");
        CPG_CHECK(method == CPG_STEP_DIRK || explicit_method(method) || method == CPG_STEP_IRK ||
                      method == CPG_STEP_IRK_W, CPG_ERR_UNSUPPORTED, "stepper method %d not available", method);
        auto *S = new cpg_stepper_s;
        if (const char *v = std::getenv("CPOLYG_VERBOSE")) S->verbose = v[0] != '\0' && v[0] != '0';
        try {
            S->E = E;
            S->method = method;
            S->ns = explicit_method(method) ? explicit_stages(method) : nstages;
            S->sp = *params;
            CPG_CHECK(params->dirk_tableau == CPG_DIRK3_REFERENCE || params->dirk_tableau == CPG_DIRK3_L_STABLE,
                      CPG_ERR_ARG, "bad dirk_tableau %d", params->dirk_tableau);
            tableau(method == CPG_STEP_IRK_W ? CPG_STEP_IRK : method, S->ns, S->A, S->b, S->c, params->dirk_tableau);
            int n, bl, nnzb;
            check(cpg_equation_info(E, &n, &bl, &nnzb));
            S->n = n;
            S->bl = bl;
            S->N = (size_t)n * bl;
            S->k.alloc(S->N * S->ns);
            S->cu.alloc(S->N);
            S->rhs.alloc(S->N);
            S->r.alloc(S->N);
            S->x.alloc(S->N);
            if (method == CPG_STEP_DIRK) {
                CPG_CHECK(params->restart > 0 && params->maxit > 0 && params->newton_maxit > 0, CPG_ERR_ARG,
                          "bad solver parameters");
                check(cpg_jacobian_create(E, &S->J));
                check(cpg_prec_create(S->J, params->pc_kind, params->precision, &S->P));
                check(cpg_prec_set_flow(S->P, params->pc_level_launches ? 0 : -1, -1));
                if (params->pc_kind == CPG_PC_INNER_GMRES) {
                    CPG_CHECK(params->gmres_mode == CPG_GMRES_FLEXIBLE, CPG_ERR_ARG,
                              "the inner-GMRES preconditioner varies between iterations: use flexible GMRES");
                    check(cpg_prec_set_inner(S->P, params->inner_iters > 0 ? params->inner_iters : 20));
                }
                check(cpg_krylov_create(n, bl, params->restart, params->gmres_mode, &S->K));
                CPG_CHECK(params->newton_operator == 0 || params->newton_operator == 1, CPG_ERR_ARG,
                          "bad newton_operator %d", params->newton_operator);
                CPG_CHECK(params->pc_refresh >= 0 && params->pc_refresh <= 3, CPG_ERR_ARG, "bad pc_refresh %d",
                          params->pc_refresh);
                S->opscale.alloc(1);
                S->nop = cpg::NewtonOp{E, S->opscale.p};
            } else if (method == CPG_STEP_IRK_W) {
                CPG_CHECK(nstages == 2 || nstages == 3, CPG_ERR_UNSUPPORTED, "Radau IIA with %d stages", nstages);
                CPG_CHECK(params->restart > 0 && params->maxit > 0 && params->newton_maxit > 0, CPG_ERR_ARG,
                          "bad solver parameters");
                CPG_CHECK(params->gmres_mode == CPG_GMRES_FLEXIBLE, CPG_ERR_ARG,
                          "the W-form IRK step applies its matrix matrix-free: use flexible GMRES");
                CPG_CHECK(params->pc_kind == CPG_PC_BJACOBI || params->pc_kind == CPG_PC_DILU, CPG_ERR_UNSUPPORTED,
                          "the W-form IRK step takes block Jacobi or block DILU per stage");
                CPG_CHECK(params->pc_refresh >= 0 && params->pc_refresh <= 3, CPG_ERR_ARG, "bad pc_refresh %d",
                          params->pc_refresh);
                const int ns = nstages;
                // A^-1 by Gauss-Jordan (ns <= 3), and b^T A^-1
                double a[3][6] = {};
                for (int i = 0; i < ns; i++) {
                    for (int j = 0; j < ns; j++) a[i][j] = S->A[i * ns + j];
                    a[i][ns + i] = 1.0;
                }
                for (int c = 0; c < ns; c++) {
                    int pr = c;
                    for (int r = c + 1; r < ns; r++)
                        if (std::fabs(a[r][c]) > std::fabs(a[pr][c])) pr = r;
                    for (int j = 0; j < 2 * ns; j++) std::swap(a[c][j], a[pr][j]);
                    const double d = a[c][c];
                    for (int j = 0; j < 2 * ns; j++) a[c][j] /= d;
                    for (int r = 0; r < ns; r++) {
                        if (r == c) continue;
                        const double f = a[r][c];
                        for (int j = 0; j < 2 * ns; j++) a[r][j] -= f * a[c][j];
                    }
                }
                for (int i = 0; i < ns; i++)
                    for (int j = 0; j < ns; j++) S->Ainv[i * ns + j] = a[i][ns + j];
                for (int j = 0; j < ns; j++) {
                    S->btAinv[j] = 0.0;
                    for (int i = 0; i < ns; i++) S->btAinv[j] += S->b[i] * S->Ainv[i * ns + j];
                }
                check(cpg_jacobian_create(E, &S->J));
                for (int st = 0; st < ns; st++) {
                    check(cpg_prec_create(S->J, params->pc_kind, params->precision, &S->Pw[st]));
                    check(cpg_prec_set_flow(S->Pw[st], params->pc_level_launches ? 0 : -1, -1));
                }
                check(cpg_krylov_create(ns * n, bl, params->restart, CPG_GMRES_FLEXIBLE, &S->K));
                S->bigR.alloc(S->N * ns);
                S->bigK.alloc(S->N * ns);
                S->wv.alloc(S->N * ns);
                S->zbuf.alloc(S->N);
                S->opscale.alloc(1);
                size_t nvol = 0, nedg = 0;
                eq_newton_data_size(*E, &nvol, &nedg);
                S->irk_data.alloc((nvol + nedg) * ns);
                S->irkwop = cpg::IrkWOp{E, ns, S->N, {}, S->opscale.p, S->zbuf.p, {}, {}};
                for (int j = 0; j < ns * ns; j++) S->irkwop.Ainv[j] = S->Ainv[j];
                for (int st = 0; st < ns; st++) {
                    S->irkwop.vol[st] = nvol ? S->irk_data.p + (nvol + nedg) * st : nullptr;
                    S->irkwop.edg[st] = nvol ? S->irk_data.p + (nvol + nedg) * st + nvol : nullptr;
                }
                S->irkwpc = cpg::IrkWPc{{S->Pw[0], S->Pw[1], S->Pw[2]}, ns, S->N};
            } else if (method == CPG_STEP_IRK && params->newton_operator == 1) {
                CPG_CHECK(nstages == 2 || nstages == 3, CPG_ERR_UNSUPPORTED, "Radau IIA with %d stages", nstages);
                CPG_CHECK(params->restart > 0 && params->maxit > 0 && params->newton_maxit > 0, CPG_ERR_ARG,
                          "bad solver parameters");
                CPG_CHECK(params->gmres_mode == CPG_GMRES_FLEXIBLE, CPG_ERR_ARG,
                          "the matrix-free IRK operator uses flexible GMRES");
                CPG_CHECK(params->pc_kind == CPG_PC_BJACOBI || params->pc_kind == CPG_PC_DILU, CPG_ERR_UNSUPPORTED,
                          "the matrix-free IRK step takes the coupled block-Jacobi or block DILU preconditioner");
                CPG_CHECK(params->pc_refresh >= 0 && params->pc_refresh <= 3, CPG_ERR_ARG, "bad pc_refresh %d",
                          params->pc_refresh);
                const int ns = nstages;
                check(cpg_jacobian_create(E, &S->J));
                if (params->pc_kind == CPG_PC_DILU) {
                    // block DILU on the cell-interleaved coupled system: its blocks are formed in the
                    // factor from the three stage Jacobians
                    CPG_CHECK(ns == 3 && bl == 40, CPG_ERR_UNSUPPORTED,
                              "the coupled DILU takes 3 stages of 40 x 40 blocks (got %d x %d)", ns, bl);
                    for (cpg_bsr &j : S->Jst) check(cpg_jacobian_create(E, &j));
                    auto *C = new cpg_bsr_s;
                    S->Cpat = C;
                    C->n = n;
                    C->bl = ns * bl;
                    C->nnzb = S->J->nnzb;
                    C->nbmax = S->J->nbmax;
                    C->h_rowptr = S->J->h_rowptr;
                    C->h_colind = S->J->h_colind;
                    C->rowptr.upload(C->h_rowptr);
                    C->colind.upload(C->h_colind);
                    check(cpg_prec_create(S->Cpat, CPG_PC_DILU, params->precision, &S->P));
                    check(cpg_prec_set_flow(S->P, params->pc_level_launches ? 0 : -1, -1));
                } else {
                    std::vector<int> rp(n + 1), ci(n);
                    for (int i = 0; i <= n; i++) rp[i] = i;
                    for (int i = 0; i < n; i++) ci[i] = i;
                    check(cpg_bsr_create(n, ns * bl, n, rp.data(), ci.data(), &S->Dc));
                    check(cpg_prec_create(S->Dc, CPG_PC_BJACOBI, params->precision, &S->P));
                }
                S->irk_mf = true;
                check(cpg_krylov_create(ns * n, bl, params->restart, CPG_GMRES_FLEXIBLE, &S->K));
                S->bigR.alloc(S->N * ns);
                S->bigK.alloc(S->N * ns);
                S->k.alloc(S->N * ns);
                S->pin.alloc(S->N * ns);
                S->pout.alloc(S->N * ns);
                S->wbuf.alloc(S->N);
                S->opscale.alloc(1);
                size_t nvol = 0, nedg = 0;
                eq_newton_data_size(*E, &nvol, &nedg);
                S->irk_data.alloc((nvol + nedg) * ns);
                S->irkop = cpg::IrkOp{E, ns, n, bl, S->N, {}, S->opscale.p, S->wbuf.p, {}, {}};
                for (int j = 0; j < ns * ns; j++) S->irkop.A[j] = S->A[j];
                for (int st = 0; st < ns; st++) {
                    S->irkop.vol[st] = nvol ? S->irk_data.p + (nvol + nedg) * st : nullptr;
                    S->irkop.edg[st] = nvol ? S->irk_data.p + (nvol + nedg) * st + nvol : nullptr;
                }
                S->irkpc = cpg::IrkPc{S->P, ns, n, bl, S->pin.p, S->pout.p};
            } else if (method == CPG_STEP_IRK) {
                CPG_CHECK(nstages == 2 || nstages == 3, CPG_ERR_UNSUPPORTED, "Radau IIA with %d stages", nstages);
                CPG_CHECK(params->restart > 0 && params->maxit > 0 && params->newton_maxit > 0, CPG_ERR_ARG,
                          "bad solver parameters");
                const int ns = nstages;
                check(cpg_jacobian_create(E, &S->J));
                std::vector<int> rp(n + 1), ci(nnzb), srow(nnzb);
                check(cpg_jacobian_pattern(E, rp.data(), ci.data()));
                for (int i = 0; i < n; i++)
                    for (int q = rp[i]; q < rp[i + 1]; q++) srow[q] = i;
                std::vector<int> brow((size_t)ns * n + 1), bcol;
                bcol.reserve((size_t)ns * ns * nnzb);
                for (int si = 0; si < ns; si++)
                    for (int i = 0; i < n; i++) {
                        brow[(size_t)si * n + i] = (int)bcol.size();
                        for (int sj = 0; sj < ns; sj++)
                            for (int q = rp[i]; q < rp[i + 1]; q++) bcol.push_back(sj * n + ci[q]);
                    }
                brow[(size_t)ns * n] = (int)bcol.size();
                check(cpg_bsr_create(ns * n, bl, (int)bcol.size(), brow.data(), bcol.data(), &S->big));
                S->slot_row.upload(srow);
                check(cpg_prec_create(S->big, params->pc_kind, params->precision, &S->P));
                check(cpg_prec_set_flow(S->P, params->pc_level_launches ? 0 : -1, -1));
                check(cpg_krylov_create(ns * n, bl, params->restart, params->gmres_mode, &S->K));
                S->bigR.alloc(S->N * ns);
                S->bigK.alloc(S->N * ns);
                S->k.alloc(S->N * ns);
            } else {
                // block inverse of the mass matrix (per component, bs x bs)
                const int bs = (int)std::lround((std::sqrt(8.0 * bl + 1) - 1) / 2);  // bl = nc*bs, try nc=1
                int nc = 1, bsx = bl;
                for (int cand : {1, 4}) {
                    if (bl % cand) continue;
                    const int b2 = bl / cand;
                    for (int d = 1; d <= 4; d++)
                        if ((d + 1) * (d + 2) / 2 == b2) { nc = cand; bsx = b2; }
                }
                (void)bs;
                std::vector<double> Mb((size_t)n * bl * bl);
                check(cpg_mass_download(E, Mb.data()));
                std::vector<double> inv((size_t)n * bsx * bsx);
                std::vector<double> a((size_t)bsx * 2 * bsx);
                for (int i = 0; i < n; i++) {
                    // Gauss-Jordan with partial pivoting on the first component block
                    for (int r = 0; r < bsx; r++)
                        for (int cc = 0; cc < bsx; cc++) {
                            a[r * 2 * bsx + cc] = Mb[(size_t)i * bl * bl + (size_t)cc * bl + r];
                            a[r * 2 * bsx + bsx + cc] = (r == cc);
                        }
                    for (int p = 0; p < bsx; p++) {
                        int pr = p;
                        for (int r = p + 1; r < bsx; r++)
                            if (std::fabs(a[r * 2 * bsx + p]) > std::fabs(a[pr * 2 * bsx + p])) pr = r;
                        CPG_CHECK(a[pr * 2 * bsx + p] != 0.0, CPG_ERR_SINGULAR, "singular mass block %d", i);
                        for (int cc = 0; cc < 2 * bsx; cc++) std::swap(a[p * 2 * bsx + cc], a[pr * 2 * bsx + cc]);
                        const double iv = 1.0 / a[p * 2 * bsx + p];
                        for (int cc = 0; cc < 2 * bsx; cc++) a[p * 2 * bsx + cc] *= iv;
                        for (int r = 0; r < bsx; r++) {
                            if (r == p) continue;
                            const double f = a[r * 2 * bsx + p];
                            for (int cc = 0; cc < 2 * bsx; cc++) a[r * 2 * bsx + cc] -= f * a[p * 2 * bsx + cc];
                        }
                    }
                    for (int r = 0; r < bsx; r++)
                        for (int cc = 0; cc < bsx; cc++) inv[(size_t)i * bsx * bsx + cc * bsx + r] = a[r * 2 * bsx + bsx + cc];
                }
                S->minv.upload(inv);
                S->A[0] = nc;  // remember nc and bs in unused tableau slots
                S->A[1] = bsx;
            }
        } catch (...) {
            delete S;
            throw;
        }
        *out = S;
    });
}

int cpg_prec_factor_coupled(cpg_prec P, cpg_equation E, cpg_bsr J0, cpg_bsr J1, cpg_bsr J2, const double *coef,
                            void *stream) {
    return guard([&] {
        CPG_CHECK(P && E && J0 && J1 && J2 && coef, CPG_ERR_ARG, "null argument");
        const cpg_bsr Js[3] = {J0, J1, J2};
        int n = 0, bl = 0, nnzb = 0;
        check(cpg_equation_info(E, &n, &bl, &nnzb));
        for (const cpg_bsr J : Js)
            CPG_CHECK(J->n == n && J->bl == bl && J->nnzb == nnzb, CPG_ERR_ARG, "stage Jacobian does not match E");
        for (const cpg_bsr J : Js) CPG_CHECK(J->h_colind == J0->h_colind, CPG_ERR_ARG, "stage Jacobian patterns differ");
        CoupledBlocks cb{};
        for (int st = 0; st < 3; st++) cb.J[st] = Js[st]->blocks.p;
        cb.mass = equation_mass(*E);
        for (int j = 0; j < 9; j++) cb.coef[j] = coef[j];
        cb.bs = equation_bs(*E);
        cb.n = n;
        cb.nnzb = nnzb;
        prec_factor_coupled(P, cb, S(stream));
    });
}

int cpg_stepper_destroy(cpg_stepper S) {
    return guard([&] { delete S; });
}

int cpg_stepper_enable_kernel_timing(cpg_stepper S, int enable) {
    return guard([&] {
        CPG_CHECK(S, CPG_ERR_ARG, "null stepper");
        if (S->K) krylov_timing(S->K, enable != 0);
    });
}

int cpg_stepper_kernel_timing(cpg_stepper S, int *n_matvec, double *ms_matvec, int *n_apply, double *ms_apply) {
    return guard([&] {
        CPG_CHECK(S && n_matvec && ms_matvec && n_apply && ms_apply, CPG_ERR_ARG, "null argument");
        *n_matvec = *n_apply = 0;
        *ms_matvec = *ms_apply = 0;
        if (S->K) krylov_timing_read(S->K, n_matvec, ms_matvec, n_apply, ms_apply);
    });
}

int cpg_stepper_advance(cpg_stepper S, const double *u, double dt, double t, double *unp1, cpg_step_stats *stats,
                        void *stream) {
    cpg_step_stats st{};
    int ret = guard([&] {
        CPG_CHECK(S && u && unp1, CPG_ERR_ARG, "null argument");
        hipStream_t s = cpg::S(stream);
        const auto t_all = clk::now();
        const size_t N = S->N;
        double *k = S->k.p;
        if (explicit_method(S->method)) {
            const int nc = (int)S->A[0], bs = (int)S->A[1];
            for (int q = 0; q < S->ns; q++) {
                const double coef[1] = {S->c[q]};
                // stage input u + c_q k_{q-1}
                if (q == 0) CPG_HIP(hipMemcpyAsync(S->cu.p, u, N * sizeof(double), hipMemcpyDeviceToDevice, s));
                else lincomb(*S, S->cu.p, u, coef, 1, k + N * (q - 1), s);
                eq_residual(*S->E, S->cu.p, t + S->c[q] * dt, S->rhs.p, s);
                hipLaunchKernelGGL(k_block_apply, dim3(ceil_div(N, 256)), dim3(256), 0, s, S->n, bs, nc, S->minv.p,
                                   S->rhs.p, k + N * q);
                CPG_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_lincomb, dim3(ceil_div(N, 256)), dim3(256), 0, s,
                                   LinComb{{k + N * q}, {dt}, 1}, k + N * q, N);
                CPG_LAUNCH_CHECK();
            }
            lincomb(*S, unp1, u, S->b, S->ns, k, s);
            CPG_HIP(hipStreamSynchronize(s));
            st.ms_total = ms_since(t_all);
            check_state(*S, unp1, t + dt, s, st);
            return;
        }
        if (S->method == CPG_STEP_IRK_W) {
            // IRK::newAdvance (src/TimeIntegration.cpp:283-395): Newton on w with k = A^-1 w, the
            // stage-block-diagonal factors refreshed per pc_refresh, flexible GMRES warm-started
            // from the previous correction (bigU is never reset in the reference)
            const int ns = S->ns;
            const size_t NB = N * ns;
            double *w = S->wv.p;
            CPG_HIP(hipMemsetAsync(w, 0, NB * sizeof(double), s));
            CPG_HIP(hipMemsetAsync(S->bigK.p, 0, NB * sizeof(double), s));
            hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, s, S->opscale.p, -dt);
            CPG_LAUNCH_CHECK();
            const double one[1] = {1.0};
            begin_step(*S, dt);
            bool converged = false;
            bool &factored = S->factored;
            for (int it = 0; it < S->sp.newton_maxit; it++) {
                auto t0 = clk::now();
                double rn2 = 0.0;
                for (int stg = 0; stg < ns; stg++) {
                    IrkCoef3 c{};
                    for (int j = 0; j < ns; j++) c.a[j] = S->Ainv[stg * ns + j];
                    hipLaunchKernelGGL(k_irk_combine, dim3(ceil_div(N, 256)), dim3(256), 0, s, c, ns, w, N,
                                       k + N * stg, nullptr);
                    CPG_LAUNCH_CHECK();
                    lincomb(*S, S->cu.p, u, one, 1, w + N * stg, s);
                    eq_residual(*S->E, S->cu.p, t + dt * S->c[stg], S->rhs.p, s);
                    const double rs = eq_newton_resid(*S->E, k + N * stg, S->rhs.p, dt, S->bigR.p + N * stg, nullptr, s);
                    rn2 += rs * rs;
                }
                const double rn = std::sqrt(rn2);
                st.ms_residual += ms_since(t0);
                st.last_resid = rn;
                CPG_CHECK(std::isfinite(rn), CPG_ERR_NONFINITE, "non-finite IRK Newton residual (it %d)", it);
                if (rn < S->sp.newton_tol) { converged = true; break; }
                st.newton_iters++;
                if (S->sp.pc_refresh == 0 || !factored) {
                    // stage s: Ainv[s][s] M - dt J(u + w_s), formed as a (M - dt / a J)
                    for (int stg = 0; stg < ns; stg++) {
                        t0 = clk::now();
                        const double a = S->Ainv[stg * ns + stg];
                        lincomb(*S, S->cu.p, u, one, 1, w + N * stg, s);
                        eq_jacobian(*S->E, S->cu.p, t + dt * S->c[stg], -dt / a, 1, S->J->blocks.p, s);
                        check(cpg_bsr_scale(S->J, a, s));
                        CPG_HIP(hipStreamSynchronize(s));
                        st.ms_jacobian += ms_since(t0);
                        t0 = clk::now();
                        check(cpg_prec_factor(S->Pw[stg], S->J, s));
                        st.ms_factor += ms_since(t0);
                    }
                    mark_factored(*S, dt);
                }
                t0 = clk::now();
                for (int stg = 0; stg < ns; stg++) {
                    lincomb(*S, S->cu.p, u, one, 1, w + N * stg, s);
                    eq_newton_prepare(*S->E, S->cu.p, t + dt * S->c[stg], s, S->irkwop.vol[stg], S->irkwop.edg[stg]);
                }
                double tol = S->sp.rtol;
                int maxit = S->sp.maxit;
                check(guard([&] {
                    gmres_flexible_op(S->K, LinOp{irkw_linop, &S->irkwop}, PcOp{irkw_pcop, &S->irkwpc}, S->bigR.p,
                                      S->bigK.p, false, &tol, &maxit, s);
                }));
                st.ms_gmres += ms_since(t0);
                st.gmres_iters += maxit;
                st.linear_solves++;
                if (S->verbose)
                    std::fprintf(stderr, "cpolyg: IRK (W form) Newton iteration %d, residual norm = %.6e, GMRES "
                                 "iterations %d\n", it, rn, maxit);
                hipLaunchKernelGGL(k_axpy, dim3(ceil_div(NB, 256)), dim3(256), 0, s, w, -1.0, S->bigK.p, NB);
                CPG_LAUNCH_CHECK();
            }
            lincomb(*S, unp1, u, S->btAinv, ns, w, s);
            CPG_HIP(hipStreamSynchronize(s));
            st.ms_total = ms_since(t_all);
            if (!converged) {
                st.status = CPG_ERR_NOT_CONVERGED;
                set_error("IRK Newton did not reach %g in %d iterations (last %g)", S->sp.newton_tol,
                          S->sp.newton_maxit, st.last_resid);
                throw Error{CPG_ERR_NOT_CONVERGED};
            }
            check_state(*S, unp1, t + dt, s, st);
            end_step(*S, st.gmres_iters);
            return;
        }
        if (S->method == CPG_STEP_IRK && S->irk_mf) {
            // IRK::advance (src/TimeIntegration.cpp:188-281) with the coupled Newton system
            // applied matrix-free: per Newton iteration the stage residuals, (per pc_refresh)
            // the stage Jacobians scattered into the coupled diagonal blocks and their
            // block-Jacobi factor, the point data of every stage's state, then flexible GMRES
            // warm-started from the previous correction as in the reference
            const int ns = S->ns;
            const int bs = equation_bs(*S->E);
            const size_t NB = N * ns;
            CPG_HIP(hipMemsetAsync(k, 0, NB * sizeof(double), s));
            CPG_HIP(hipMemsetAsync(S->bigK.p, 0, NB * sizeof(double), s));
            hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, s, S->opscale.p, -dt);
            CPG_LAUNCH_CHECK();
            begin_step(*S, dt);
            bool converged = false;
            bool &factored = S->factored;
            for (int it = 0; it < S->sp.newton_maxit; it++) {
                auto t0 = clk::now();
                double rn2 = 0.0;
                for (int stg = 0; stg < ns; stg++) {
                    lincomb(*S, S->cu.p, u, &S->A[stg * ns], ns, k, s);
                    eq_residual(*S->E, S->cu.p, t + dt * S->c[stg], S->rhs.p, s);
                    const double rs = eq_newton_resid(*S->E, k + N * stg, S->rhs.p, dt, S->bigR.p + N * stg, nullptr, s);
                    rn2 += rs * rs;
                }
                const double rn = std::sqrt(rn2);
                st.ms_residual += ms_since(t0);
                st.last_resid = rn;
                CPG_CHECK(std::isfinite(rn), CPG_ERR_NONFINITE, "non-finite IRK Newton residual (it %d)", it);
                if (rn < S->sp.newton_tol) { converged = true; break; }
                st.newton_iters++;
                if ((S->sp.pc_refresh == 0 || !factored) && S->Cpat) {
                    // the stage Jacobians, then the coupled DILU formed from them
                    t0 = clk::now();
                    cpg_bsr Js[3] = {S->J, S->Jst[0], S->Jst[1]};
                    for (int stg = 0; stg < ns; stg++) {
                        lincomb(*S, S->cu.p, u, &S->A[stg * ns], ns, k, s);
                        eq_jacobian(*S->E, S->cu.p, t + dt * S->c[stg], 1.0, 0, Js[stg]->blocks.p, s);
                    }
                    CPG_HIP(hipStreamSynchronize(s));
                    st.ms_jacobian += ms_since(t0);
                    t0 = clk::now();
                    CoupledBlocks cb{};
                    for (int stg = 0; stg < ns; stg++) cb.J[stg] = Js[stg]->blocks.p;
                    cb.mass = equation_mass(*S->E);
                    for (int j = 0; j < 9; j++) cb.coef[j] = -dt * S->A[j];
                    cb.bs = bs;
                    cb.n = S->n;
                    cb.nnzb = S->J->nnzb;
                    check(guard([&] { prec_factor_coupled(S->P, cb, s); }));
                    st.ms_factor += ms_since(t0);
                    mark_factored(*S, dt);
                } else if (S->sp.pc_refresh == 0 || !factored) {
                    t0 = clk::now();
                    int nnzb = 0;
                    check(cpg_equation_info(S->E, nullptr, nullptr, &nnzb));
                    IrkCoef3 adt{};
                    for (int stg = 0; stg < ns; stg++) {
                        for (int tt = 0; tt < ns; tt++) adt.a[tt] = -dt * S->A[stg * ns + tt];
                        lincomb(*S, S->cu.p, u, &S->A[stg * ns], ns, k, s);
                        eq_jacobian(*S->E, S->cu.p, t + dt * S->c[stg], 1.0, 0, S->J->blocks.p, s);
                        const size_t cnt = (size_t)S->n * S->bl * ns * S->bl;
                        hipLaunchKernelGGL(k_irk_diag_scatter, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, ns, stg, S->n,
                                           S->bl, bs, S->J->rowptr.p, S->J->blocks.p, equation_mass(*S->E), adt,
                                           S->Dc->blocks.p);
                        CPG_LAUNCH_CHECK();
                    }
                    CPG_HIP(hipStreamSynchronize(s));
                    st.ms_jacobian += ms_since(t0);
                    t0 = clk::now();
                    check(cpg_prec_factor(S->P, S->Dc, s));
                    st.ms_factor += ms_since(t0);
                    mark_factored(*S, dt);
                }
                t0 = clk::now();
                for (int stg = 0; stg < ns; stg++) {
                    lincomb(*S, S->cu.p, u, &S->A[stg * ns], ns, k, s);
                    eq_newton_prepare(*S->E, S->cu.p, t + dt * S->c[stg], s, S->irkop.vol[stg], S->irkop.edg[stg]);
                }
                double tol = S->sp.rtol;
                int maxit = S->sp.maxit;
                check(guard([&] {
                    gmres_flexible_op(S->K, LinOp{irk_linop, &S->irkop}, PcOp{irk_pcop, &S->irkpc}, S->bigR.p,
                                      S->bigK.p, false, &tol, &maxit, s);
                }));
                st.ms_gmres += ms_since(t0);
                st.gmres_iters += maxit;
                st.linear_solves++;
                if (S->verbose)
                    std::fprintf(stderr, "cpolyg: IRK Newton iteration %d, residual norm = %.6e, GMRES iterations %d\n",
                                 it, rn, maxit);
                hipLaunchKernelGGL(k_axpy, dim3(ceil_div(NB, 256)), dim3(256), 0, s, k, -1.0, S->bigK.p, NB);
                CPG_LAUNCH_CHECK();
            }
            lincomb(*S, unp1, u, S->b, ns, k, s);
            CPG_HIP(hipStreamSynchronize(s));
            st.ms_total = ms_since(t_all);
            if (!converged) {
                st.status = CPG_ERR_NOT_CONVERGED;
                set_error("IRK Newton did not reach %g in %d iterations (last %g)", S->sp.newton_tol,
                          S->sp.newton_maxit, st.last_resid);
                throw Error{CPG_ERR_NOT_CONVERGED};
            }
            check_state(*S, unp1, t + dt, s, st);
            end_step(*S, st.gmres_iters);
            return;
        }
        if (S->method == CPG_STEP_IRK) {
            // IRK::advance (src/TimeIntegration.cpp:188-281): all stages in one Newton loop;
            // bigK is not reset between Newton iterations, so each GMRES starts from the
            // previous correction.
            const int ns = S->ns;
            const int bs = equation_bs(*S->E);
            int nnzb = 0;
            check(cpg_equation_info(S->E, nullptr, nullptr, &nnzb));
            CPG_HIP(hipMemsetAsync(k, 0, N * ns * sizeof(double), s));
            CPG_HIP(hipMemsetAsync(S->bigK.p, 0, N * ns * sizeof(double), s));
            bool converged = false;
            for (int it = 0; it < S->sp.newton_maxit; it++) {
                auto t0 = clk::now();
                double rn2 = 0.0;
                for (int stg = 0; stg < ns; stg++) {
                    lincomb(*S, S->cu.p, u, &S->A[stg * ns], ns, k, s);
                    eq_residual(*S->E, S->cu.p, t + dt * S->c[stg], S->rhs.p, s);
                    const double rs = eq_newton_resid(*S->E, k + N * stg, S->rhs.p, dt, S->bigR.p + N * stg, nullptr, s);
                    rn2 += rs * rs;
                }
                const double rn = std::sqrt(rn2);
                st.ms_residual += ms_since(t0);
                st.last_resid = rn;
                CPG_CHECK(std::isfinite(rn), CPG_ERR_NONFINITE, "non-finite IRK Newton residual (it %d)", it);
                if (rn < S->sp.newton_tol) { converged = true; break; }
                st.newton_iters++;
                t0 = clk::now();
                for (int stg = 0; stg < ns; stg++) {
                    lincomb(*S, S->cu.p, u, &S->A[stg * ns], ns, k, s);
                    eq_jacobian(*S->E, S->cu.p, t + dt * S->c[stg], 1.0, 0, S->J->blocks.p, s);
                    IrkCoef cf{};
                    for (int sj = 0; sj < ns; sj++) cf.c[sj] = -dt * S->A[stg * ns + sj];
                    const size_t cnt = (size_t)ns * nnzb * S->bl * S->bl;
                    hipLaunchKernelGGL(k_irk_scatter, dim3(ceil_div(cnt, 256)), dim3(256), 0, s, ns, stg, nnzb, S->bl,
                                       bs, S->J->rowptr.p, S->slot_row.p, S->J->blocks.p, cf, equation_mass(*S->E),
                                       S->big->blocks.p);
                    CPG_LAUNCH_CHECK();
                }
                CPG_HIP(hipStreamSynchronize(s));
                st.ms_jacobian += ms_since(t0);
                t0 = clk::now();
                check(cpg_prec_factor(S->P, S->big, s));
                st.ms_factor += ms_since(t0);
                t0 = clk::now();
                double tol = S->sp.rtol;
                int maxit = S->sp.maxit;
                check(cpg_gmres(S->K, S->big, S->P, S->bigR.p, S->bigK.p, &tol, &maxit, s));
                st.ms_gmres += ms_since(t0);
                st.gmres_iters += maxit;
                st.linear_solves++;
                hipLaunchKernelGGL(k_axpy, dim3(ceil_div(N * ns, 256)), dim3(256), 0, s, k, -1.0, S->bigK.p, N * ns);
                CPG_LAUNCH_CHECK();
            }
            lincomb(*S, unp1, u, S->b, ns, k, s);
            CPG_HIP(hipStreamSynchronize(s));
            st.ms_total = ms_since(t_all);
            if (!converged) {
                st.status = CPG_ERR_NOT_CONVERGED;
                set_error("IRK Newton did not reach %g in %d iterations (last %g)", S->sp.newton_tol,
                          S->sp.newton_maxit, st.last_resid);
                throw Error{CPG_ERR_NOT_CONVERGED};
            }
            check_state(*S, unp1, t + dt, s, st);
            end_step(*S, st.gmres_iters);
            return;
        }
        const int ns = S->ns;
        bool all_converged = true;
        begin_step(*S, dt);
        CPG_HIP(hipMemsetAsync(k, 0, N * ns * sizeof(double), s));
        for (int stg = 0; stg < ns; stg++) {
            double *ks = k + N * stg;
            const double ts = t + dt * S->c[stg];
            bool converged = false;
            if (S->sp.newton_guess == 1 && stg > 0)
                CPG_HIP(hipMemcpyAsync(ks, ks - N, N * sizeof(double), hipMemcpyDeviceToDevice, s));
            for (int it = 0; it < S->sp.newton_maxit; it++) {
                auto t0 = clk::now();
                double rn = 0.0;
                for (int half = 0;; half++) {
                    lincomb(*S, S->cu.p, u, &S->A[stg * ns], stg + 1, k, s);
                    eq_residual(*S->E, S->cu.p, ts, S->rhs.p, s);
                    rn = eq_newton_resid(*S->E, ks, S->rhs.p, dt, S->r.p, nullptr, s);
                    if (std::isfinite(rn) || !S->sp.newton_backtrack || it == 0 || half >= 10) break;
                    // the last update k_s -= x led to a non-physical state: take half of it
                    hipLaunchKernelGGL(k_axpy, dim3(ceil_div(N, 256)), dim3(256), 0, s, ks, std::ldexp(1.0, -half - 1),
                                       S->x.p, N);
                    CPG_LAUNCH_CHECK();
                    st.backtracks++;
                }
                st.ms_residual += ms_since(t0);
                st.last_resid = rn;
                CPG_CHECK(std::isfinite(rn), CPG_ERR_NONFINITE, "non-finite Newton residual (stage %d, it %d)", stg, it);
                if (rn < S->sp.newton_tol) { converged = true; break; }
                st.newton_iters++;
                // with the matrix-free operator the matrix only feeds the preconditioner: pc_refresh 1
                // refreshes it at the first Newton iteration of a stage, 2 at that of the step
                bool refresh = true;
                if (S->sp.newton_operator == 1 && S->sp.pc_refresh == 1) refresh = it == 0;
                if (S->sp.newton_operator == 1 && S->sp.pc_refresh == 2) refresh = it == 0 && (stg == 0 || !S->factored);
                if (S->sp.newton_operator == 1 && S->sp.pc_refresh == 3) refresh = !S->factored;
                if (refresh) mark_factored(*S, dt);
                if (refresh) {
                    t0 = clk::now();
                    eq_jacobian(*S->E, S->cu.p, ts, -dt * S->A[stg * ns + stg], 1, S->J->blocks.p, s);
                    CPG_HIP(hipStreamSynchronize(s));
                    st.ms_jacobian += ms_since(t0);
                    t0 = clk::now();
                    check_ctx(cpg_prec_factor(S->P, S->J, s), stg, it, rn);
                    st.ms_factor += ms_since(t0);
                }
                t0 = clk::now();
                double tol = S->sp.rtol;
                int maxit = S->sp.maxit;
                // GMRES from x = 0 (Jacobian::solve, src/Jacobian.cpp:72) on the assembled matrix or
                // on the same matrix applied at the quadrature points (the Dirichlet states the
                // Jacobian assembly just computed are those of time ts)
                LinOp op{bsr_linop, S->J};
                if (S->sp.newton_operator == 1) {
                    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, s, S->opscale.p, -dt * S->A[stg * ns + stg]);
                    CPG_LAUNCH_CHECK();
                    eq_newton_prepare(*S->E, S->cu.p, ts, s);
                    op = LinOp{newton_linop, &S->nop};
                }
                check_ctx(guard([&] { gmres_solve(S->K, op, S->P, S->r.p, S->x.p, true, &tol, &maxit, s); }), stg, it,
                          rn);
                st.ms_gmres += ms_since(t0);
                st.gmres_iters += maxit;
                st.linear_solves++;
                if (S->verbose)
                    std::fprintf(stderr, "cpolyg: stage %d Newton iteration %d, residual norm = %.6e, GMRES iterations %d, "
                                 "backtracks %d\n", stg, it, rn, maxit, st.backtracks);
                hipLaunchKernelGGL(k_axpy, dim3(ceil_div(N, 256)), dim3(256), 0, s, ks, -1.0, S->x.p, N);
                CPG_LAUNCH_CHECK();
            }
            if (!converged) all_converged = false;
        }
        lincomb(*S, unp1, u, S->b, ns, k, s);
        CPG_HIP(hipStreamSynchronize(s));
        st.ms_total = ms_since(t_all);
        if (!all_converged) {
            st.status = CPG_ERR_NOT_CONVERGED;
            set_error("Newton did not reach %g in %d iterations (last %g)", S->sp.newton_tol, S->sp.newton_maxit,
                      st.last_resid);
            throw Error{CPG_ERR_NOT_CONVERGED};
        }
        check_state(*S, unp1, t + dt, s, st);
        end_step(*S, st.gmres_iters);
    });
    st.status = ret;
    if (stats) *stats = st;
    return ret;
}

}  // extern "C"
