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

void tableau(int method, int ns, double *A, double *b, double *c) {
    std::memset(A, 0, sizeof(double) * 25);
    if (method == CPG_STEP_DIRK) {
        if (ns == 1) { A[0] = 1; b[0] = 1; c[0] = 1; return; }
        if (ns == 2) {
            const double g = 1.0 - 1.0 / std::sqrt(2.0);
            A[0] = g; A[1] = 0; A[2] = 1 - g; A[3] = g;
            b[0] = 1 - g; b[1] = g; c[0] = g; c[1] = 1;
            return;
        }
        if (ns == 3) {
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
    CPG_CHECK(false, CPG_ERR_UNSUPPORTED, "method %d with %d stages", method, ns);
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
    cpg_prec P = nullptr;
    cpg_krylov K = nullptr;
    DevBuf<double> k, cu, rhs, r, x, minv;
    ~cpg_stepper_s() {
        if (P) cpg_prec_destroy(P);
        if (K) cpg_krylov_destroy(K);
        if (J) cpg_bsr_destroy(J);
    }
};

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

}  // namespace

extern "C" {

int cpg_stepper_create(cpg_equation E, int method, int nstages, const cpg_solver_params *params, cpg_stepper *out) {
    return guard([&] {
        CPG_CHECK(E && params && out, CPG_ERR_ARG, "null argument");
        CPG_CHECK(method == CPG_STEP_DIRK || method == CPG_STEP_RK4, CPG_ERR_UNSUPPORTED,
                  "stepper method %d not available", method);
        auto *S = new cpg_stepper_s;
        try {
            S->E = E;
            S->method = method;
            S->ns = method == CPG_STEP_RK4 ? 4 : nstages;
            S->sp = *params;
            tableau(method, S->ns, S->A, S->b, S->c);
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
                check(cpg_krylov_create(n, bl, params->restart, params->gmres_mode, &S->K));
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
        if (S->method == CPG_STEP_RK4) {
            const int nc = (int)S->A[0], bs = (int)S->A[1];
            for (int q = 0; q < 4; q++) {
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
            lincomb(*S, unp1, u, S->b, 4, k, s);
            CPG_HIP(hipStreamSynchronize(s));
            st.ms_total = ms_since(t_all);
            return;
        }
        const int ns = S->ns;
        bool all_converged = true;
        CPG_HIP(hipMemsetAsync(k, 0, N * ns * sizeof(double), s));
        for (int stg = 0; stg < ns; stg++) {
            double *ks = k + N * stg;
            const double ts = t + dt * S->c[stg];
            bool converged = false;
            for (int it = 0; it < S->sp.newton_maxit; it++) {
                auto t0 = clk::now();
                lincomb(*S, S->cu.p, u, &S->A[stg * ns], stg + 1, k, s);
                eq_residual(*S->E, S->cu.p, ts, S->rhs.p, s);
                const double rn = eq_newton_resid(*S->E, ks, S->rhs.p, dt, S->r.p, nullptr, s);
                st.ms_residual += ms_since(t0);
                st.last_resid = rn;
                CPG_CHECK(std::isfinite(rn), CPG_ERR_NONFINITE, "non-finite Newton residual (stage %d, it %d)", stg, it);
                if (rn < S->sp.newton_tol) { converged = true; break; }
                st.newton_iters++;
                t0 = clk::now();
                eq_jacobian(*S->E, S->cu.p, ts, -dt * S->A[stg * ns + stg], 1, S->J->blocks.p, s);
                CPG_HIP(hipStreamSynchronize(s));
                st.ms_jacobian += ms_since(t0);
                t0 = clk::now();
                check(cpg_prec_factor(S->P, S->J, s));
                st.ms_factor += ms_since(t0);
                t0 = clk::now();
                vec_fill(S->x.p, 0.0, N, s);
                double tol = S->sp.rtol;
                int maxit = S->sp.maxit;
                check(cpg_gmres(S->K, S->J, S->P, S->r.p, S->x.p, &tol, &maxit, s));
                st.ms_gmres += ms_since(t0);
                st.gmres_iters += maxit;
                st.linear_solves++;
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
    });
    st.status = ret;
    if (stats) *stats = st;
    return ret;
}

}  // extern "C"
