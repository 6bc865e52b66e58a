// Restarted GMRES on the device (BlockMatrix::gmres, src/BlockMatrix.cpp:79-163).
//
// Vectors (Krylov basis V, and Z for the flexible variant) live in HBM; the
// Hessenberg column, Givens rotations and the small triangular solve run on the
// host in FP64 exactly as the reference does them (one D2H copy of i+2 doubles per
// Arnoldi step).  Classical Gram-Schmidt is two passes over V: a multi-dot
// h = V^T w (one read of V and w), then w -= V h fused with the partial sums of
// ||w||^2.  All reductions are two-stage with a fixed order (deterministic).
#include "cpg_internal.hpp"

#include <cmath>
#include <cstring>

namespace cpg {

constexpr int KWG = 256;
constexpr int KPT = 8;  // elements per thread in the vector kernels
constexpr int KCH = KWG * KPT;

// partial[b*nv + j] = sum over chunk b of V[j] . w
__global__ __launch_bounds__(KWG) void k_mdot(const double *__restrict__ V, size_t N, int nv,
                                              const double *__restrict__ w, double *__restrict__ partial) {
    __shared__ double red[KWG / 64][64];
    const size_t base = (size_t)blockIdx.x * KCH + threadIdx.x;
    double wv[KPT];
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        wv[k] = e < N ? w[e] : 0.0;
    }
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j0 = 0; j0 < nv; j0 += 64) {
        const int jn = min(64, nv - j0);
        for (int jj = 0; jj < jn; jj++) {
            const double *v = V + (size_t)(j0 + jj) * N;
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < KPT; k++) {
                const size_t e = base + (size_t)k * KWG;
                if (e < N) s = fma(v[e], wv[k], s);
            }
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) red[wid][jj] = s;
        }
        __syncthreads();
        if (threadIdx.x < jn) {
            double s = 0.0;
            for (int q = 0; q < KWG / 64; q++) s += red[q][threadIdx.x];
            partial[(size_t)blockIdx.x * nv + j0 + threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// out[j] = sum_b partial[b*nv + j]; if sqrt_out, out[j] = sqrt(out[j])
__global__ __launch_bounds__(KWG) void k_reduce(const double *__restrict__ partial, int nb, int nv,
                                                double *__restrict__ out, int sqrt_out) {
    __shared__ double red[KWG / 64];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += KWG) s += partial[(size_t)b * nv + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < KWG / 64; q++) t += red[q];
        out[j] = sqrt_out ? sqrt(t) : t;
    }
}

// w -= sum_j h[j] V[j]; partial[b] = sum over chunk b of w^2
__global__ __launch_bounds__(KWG) void k_orth_norm(const double *__restrict__ V, size_t N, int nv,
                                                   const double *__restrict__ h, double *__restrict__ w,
                                                   double *__restrict__ partial) {
    __shared__ double red[KWG / 64];
    const size_t base = (size_t)blockIdx.x * KCH + threadIdx.x;
    double wv[KPT];
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        wv[k] = e < N ? w[e] : 0.0;
    }
    for (int j = 0; j < nv; j++) {
        const double hj = h[j];
        const double *v = V + (size_t)j * N;
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const size_t e = base + (size_t)k * KWG;
            if (e < N) wv[k] = fma(-hj, v[e], wv[k]);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        if (e < N) {
            w[e] = wv[k];
            s = fma(wv[k], wv[k], s);
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < KWG / 64; q++) t += red[q];
        partial[blockIdx.x] = t;
    }
}

// v = w * (1/norm) (0 if norm is 0)
__global__ void k_scale_dev(const double *__restrict__ w, double *__restrict__ v, size_t N,
                            const double *__restrict__ norm) {
    const double nv = *norm;
    const double inv = nv > 0.0 ? 1.0 / nv : 0.0;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) v[e] = w[e] * inv;
}

// v = w * inv
__global__ void k_scale_by(const double *__restrict__ w, double *__restrict__ v, size_t N, double inv) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) v[e] = w[e] * inv;
}

// r = r - b
__global__ void k_sub(double *__restrict__ r, const double *__restrict__ b, size_t N) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) r[e] -= b[e];
}

// x -= sum_k y[k] B[k]
__global__ __launch_bounds__(KWG) void k_update(const double *__restrict__ B, size_t N, int nk,
                                                const double *__restrict__ y, double *__restrict__ x) {
    const size_t base = (size_t)blockIdx.x * KCH + threadIdx.x;
    double xv[KPT];
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        xv[k] = e < N ? x[e] : 0.0;
    }
    for (int j = 0; j < nk; j++) {
        const double yj = y[j];
        const double *b = B + (size_t)j * N;
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const size_t e = base + (size_t)k * KWG;
            if (e < N) xv[k] = fma(-yj, b[e], xv[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        if (e < N) x[e] = xv[k];
    }
}

}  // namespace cpg

using namespace cpg;

struct cpg_krylov_s {
    int n = 0, bl = 0, m = 0, mode = 0;
    size_t N = 0;
    int nblk = 0;
    DevBuf<double> V, Z, r, partial, hdev, ydev, scal;
    std::vector<double> h, y, cs, sn, hcol;
    // kernel timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> evkind;  // 0 matvec, 1 apply
    size_t evused = 0;
    int n_matvec = 0, n_apply = 0;
    double ms_matvec = 0, ms_apply = 0;

    ~cpg_krylov_s() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }

    void mark(int kind, bool start, hipStream_t s) {
        if (!timing) return;
        if (evused == ev.size()) {
            hipEvent_t e;
            CPG_HIP(hipEventCreate(&e));
            ev.push_back(e);
            evkind.push_back(0);
        }
        evkind[evused] = kind;
        CPG_HIP(hipEventRecord(ev[evused], s));
        evused++;
        (void)start;
    }
    void harvest() {
        if (!timing) return;
        for (size_t k = 0; k + 1 < evused; k += 2) {
            float ms = 0;
            CPG_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            if (evkind[k] == 0) { ms_matvec += ms; n_matvec++; }
            else { ms_apply += ms; n_apply++; }
        }
        evused = 0;
    }
};

namespace cpg {

struct KOps {
    cpg_krylov_s &K;
    cpg_bsr_s &A;
    cpg_prec P;
    hipStream_t s;

    void matvec(const double *x, double *y) {
        K.mark(0, true, s);
        bsr_matvec(A, x, y, s);
        K.mark(0, false, s);
    }
    void pc(double *x) {
        if (!P) return;
        K.mark(1, true, s);
        int st = cpg_prec_apply(P, x, s);
        if (st) throw Error{st};
        K.mark(1, false, s);
    }
    // out = M^-1 in (a copy when there is no preconditioner)
    void pc_to(const double *in, double *out) {
        if (!P) {
            CPG_HIP(hipMemcpyAsync(out, in, K.N * sizeof(double), hipMemcpyDeviceToDevice, s));
            return;
        }
        K.mark(1, true, s);
        prec_apply_to(P, in, out, s);
        K.mark(1, false, s);
    }
    // returns ||x|| (synchronous)
    double norm(const double *x) {
        hipLaunchKernelGGL(k_mdot, dim3(K.nblk), dim3(KWG), 0, s, x, K.N, 1, x, K.partial.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(KWG), 0, s, K.partial.p, K.nblk, 1, K.scal.p, 1);
        CPG_LAUNCH_CHECK();
        double v;
        CPG_HIP(hipMemcpyAsync(&v, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
        CPG_HIP(hipStreamSynchronize(s));
        return v;
    }
    void scale_into(const double *w, double *v, const double *norm_dev) {
        hipLaunchKernelGGL(k_scale_dev, dim3(ceil_div(K.N, 256)), dim3(256), 0, s, w, v, K.N, norm_dev);
        CPG_LAUNCH_CHECK();
    }
};

}  // namespace cpg

extern "C" {

int cpg_krylov_create(int n_rows, int bl, int restart, int mode, cpg_krylov *out) {
    return guard([&] {
        CPG_CHECK(out && n_rows >= 0 && bl > 0 && restart > 0, CPG_ERR_ARG, "bad arguments");
        CPG_CHECK(mode == CPG_GMRES_LEFT_REF || mode == CPG_GMRES_FLEXIBLE, CPG_ERR_ARG, "bad mode %d", mode);
        auto *K = new cpg_krylov_s;
        try {
            K->n = n_rows;
            K->bl = bl;
            K->m = restart;
            K->mode = mode;
            K->N = (size_t)n_rows * bl;
            K->nblk = std::max(1, ceil_div(K->N, KCH));
            K->V.alloc(K->N * (restart + 1));
            if (mode == CPG_GMRES_FLEXIBLE) K->Z.alloc(K->N * restart);
            K->r.alloc(K->N);
            K->partial.alloc((size_t)K->nblk * (restart + 1));
            K->hdev.alloc(restart + 2);
            K->ydev.alloc(restart + 1);
            K->scal.alloc(2);
            K->h.assign((size_t)restart * (restart + 1) / 2 + 1, 0.0);
            K->y.assign(restart + 1, 0.0);
            K->cs.assign(restart, 0.0);
            K->sn.assign(restart, 0.0);
            K->hcol.assign(restart + 2, 0.0);
        } catch (...) {
            delete K;
            throw;
        }
        *out = K;
    });
}

int cpg_krylov_destroy(cpg_krylov K) {
    return guard([&] { delete K; });
}

int cpg_gmres(cpg_krylov K, cpg_bsr A, cpg_prec P, const double *b, double *x, double *tol, int *maxit,
              void *stream) {
    return guard([&] {
        CPG_CHECK(K && A && b && x && tol && maxit, CPG_ERR_ARG, "null argument");
        CPG_CHECK(A->n == K->n && A->bl == K->bl, CPG_ERR_ARG, "matrix does not match the workspace");
        hipStream_t s = S(stream);
        KOps op{*K, *A, P, s};
        const size_t N = K->N;
        const int m = K->m;
        const bool left = K->mode == CPG_GMRES_LEFT_REF;
        double *V = K->V.p;
        double *r = K->r.p;
        const double tl = *tol;
        int mx = *maxit;

        CPG_HIP(hipMemcpyAsync(r, b, N * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (left) op.pc(r);
        const double nrm2b = op.norm(r);
        CPG_CHECK(std::isfinite(nrm2b), CPG_ERR_NONFINITE, "non-finite right-hand side norm");
        if (mx % m > 0) mx = (mx / m + 1) * m;
        if (nrm2b == 0.0) {
            vec_fill(x, 0.0, N, s);
            CPG_HIP(hipStreamSynchronize(s));
            *tol = 0.0;
            *maxit = 0;
            K->harvest();
            return;
        }
        auto &h = K->h;
        auto &y = K->y;
        auto &cs = K->cs;
        auto &sn = K->sn;
        int i = 0, j;
        y[0] = 0.0;
        for (j = 0; j < mx / m; j++) {
            // r = [P^-1](A x - b)
            op.matvec(x, r);
            hipLaunchKernelGGL(k_sub, dim3(ceil_div(N, 256)), dim3(256), 0, s, r, b, N);
            CPG_LAUNCH_CHECK();
            if (left) op.pc(r);
            const double beta = op.norm(r);
            CPG_CHECK(std::isfinite(beta), CPG_ERR_NONFINITE,
                      "non-finite residual norm at restart %d (previous cycle %d Arnoldi steps, |y_i| %g, ||b|| %g)", j,
                      i, std::fabs(y[i]), nrm2b);
            hipLaunchKernelGGL(k_scale_by, dim3(ceil_div(N, 256)), dim3(256), 0, s, r, V, N,
                               beta > 0.0 ? 1.0 / beta : 0.0);
            CPG_LAUNCH_CHECK();
            y[0] = beta;
            int uki = 0;
            for (i = 0; i < m; i++) {
                if (std::fabs(y[i]) < tl * nrm2b) break;
                double *vi = V + N * i;
                double *vi1 = V + N * (i + 1);
                const int u0i = uki;
                if (left) {
                    op.matvec(vi, vi1);
                    op.pc(vi1);
                } else {
                    double *zi = K->Z.p + N * i;
                    op.pc_to(vi, zi);
                    op.matvec(zi, vi1);
                }
                // classical Gram-Schmidt: h = V^T w ; w -= V h ; hnew = ||w||
                hipLaunchKernelGGL(k_mdot, dim3(K->nblk), dim3(KWG), 0, s, V, N, i + 1, vi1, K->partial.p);
                CPG_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_reduce, dim3(i + 1), dim3(KWG), 0, s, K->partial.p, K->nblk, i + 1, K->hdev.p, 0);
                CPG_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_orth_norm, dim3(K->nblk), dim3(KWG), 0, s, V, N, i + 1, K->hdev.p, vi1,
                                   K->partial.p);
                CPG_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_reduce, dim3(1), dim3(KWG), 0, s, K->partial.p, K->nblk, 1, K->hdev.p + i + 1, 1);
                CPG_LAUNCH_CHECK();
                op.scale_into(vi1, vi1, K->hdev.p + i + 1);
                CPG_HIP(hipMemcpyAsync(K->hcol.data(), K->hdev.p, (i + 2) * sizeof(double), hipMemcpyDeviceToHost, s));
                CPG_HIP(hipStreamSynchronize(s));
                for (int k = 0; k <= i; k++) h[u0i + k] = K->hcol[k];
                const double hnew = K->hcol[i + 1];
                bool hfin = std::isfinite(hnew);
                for (int k = 0; k <= i; k++) hfin = hfin && std::isfinite(K->hcol[k]);
                CPG_CHECK(hfin, CPG_ERR_NONFINITE,
                          "non-finite Hessenberg column (cycle %d, Arnoldi step %d, ||w|| %g, h_0 %g)", j, i, hnew,
                          K->hcol[0]);
                for (int k = 0; k < i; k++) {
                    const double tmp = cs[k] * h[uki] - sn[k] * h[uki + 1];
                    h[uki + 1] = sn[k] * h[uki] + cs[k] * h[uki + 1];
                    h[uki] = tmp;
                    ++uki;
                }
                const double rd = h[uki];
                const double dd = std::sqrt(rd * rd + hnew * hnew);
                cs[i] = rd / dd;
                sn[i] = -hnew / dd;
                h[uki] = dd;
                ++uki;
                y[i + 1] = sn[i] * y[i];
                y[i] = cs[i] * y[i];
            }
            // packed upper-triangular solve (column k starts at k(k+1)/2)
            double dmin = INFINITY;
            for (int c = i - 1; c >= 0; c--) {
                dmin = std::fmin(dmin, std::fabs(h[c * (c + 1) / 2 + c]));
                y[c] /= h[c * (c + 1) / 2 + c];
                const double yc = y[c];
                for (int rr = 0; rr < c; rr++) y[rr] -= h[c * (c + 1) / 2 + rr] * yc;
            }
            for (int c = 0; c < i; c++)
                CPG_CHECK(std::isfinite(y[c]), CPG_ERR_NONFINITE,
                          "GMRES least-squares solution non-finite (cycle %d, %d Arnoldi steps, min |R_kk| %g, "
                          "||b|| %g)", j, i, dmin, nrm2b);
            if (i > 0) {
                CPG_HIP(hipMemcpyAsync(K->ydev.p, y.data(), i * sizeof(double), hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_update, dim3(K->nblk), dim3(KWG), 0, s, left ? V : K->Z.p, N, i, K->ydev.p, x);
                CPG_LAUNCH_CHECK();
            }
            CPG_HIP(hipStreamSynchronize(s));
            if (std::fabs(y[i]) < tl * nrm2b) break;
        }
        *maxit = m * j + i;
        *tol = std::fabs(y[i]) / nrm2b;
        K->harvest();
        CPG_CHECK(std::isfinite(*tol), CPG_ERR_NONFINITE, "non-finite GMRES residual estimate");
    });
}

}  // extern "C"

// timing accessors used by the stepper
namespace cpg {
void krylov_timing(cpg_krylov K, bool on) { K->timing = on; K->n_matvec = K->n_apply = 0; K->ms_matvec = K->ms_apply = 0; }
void krylov_timing_read(cpg_krylov K, int *nm, double *msm, int *na, double *msa) {
    *nm = K->n_matvec; *msm = K->ms_matvec; *na = K->n_apply; *msa = K->ms_apply;
}
}  // namespace cpg
