// Batched implicit steps: many independent DG solves advanced together.
//
// The reference advances one mesh at a time (DIRK::advance / BackwardEuler::advance,
// src/TimeIntegration.cpp:31-151, each Newton system solved by BlockMatrix::gmres,
// src/BlockMatrix.cpp:79-163).  Here instances are grouped into classes of equal polynomial
// degree; a class is one equation over a union mesh whose connected components are the
// instances, so residual, Jacobian, BSR matvec, DILU factor and DILU sweeps of all
// instances of a class are single launches.  Each stage's Newton systems are solved by one
// batched flexible GMRES(m) over the concatenated vector of all classes:
//   * inner products and norms are segmented per instance (a chunk table maps every
//     2048-element chunk to one instance; two-stage fixed-order reductions);
//   * the Hessenberg columns, Givens rotations, least-squares solves and convergence flags
//     live on the device, one thread per instance, so a restart cycle has no host
//     synchronisation; an instance that converges stops updating its Krylov space and
//     solution while the others continue;
//   * a restart cycle is captured once into hipGraphs (start; blocks of 5 Arnoldi steps;
//     finish) and replayed; between blocks the host reads the instances' activity flags and
//     ends the cycle early when all have converged.
// Each instance has its own dt (its own CFL step): the Newton matrix M - dt_k a J and the
// Newton residual M k - dt_k b take a per-cell dt.
#include "cpg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

struct cpg_equation_s;

namespace cpg {
void eq_residual(const cpg_equation_s &E, const double *u, double t, double *b, hipStream_t s);
void eq_jacobian_cells(const cpg_equation_s &E, const double *u, double t, double scale, const double *cfac,
                       int add_mass, double *J, hipStream_t s);
void eq_newton_resid_cells(const cpg_equation_s &E, const double *kv, const double *b, double dt, const double *cfac,
                           double *r, hipStream_t s);
const double *eq_normc(const cpg_equation_s &E);
int equation_bs(const cpg_equation_s &E);
void tableau(int method, int ns, double *A, double *b, double *c, int dirk3);

constexpr int BWG = 256;
constexpr int BPT = 8;
constexpr int BCH = BWG * BPT;   // elements per chunk

struct Chunks {
    const int *c0, *c1, *inst;   // chunk k covers [c0[k], c1[k]) of instance inst[k]
};

// partial[k * ld + j] = sum over chunk k of V_j . w  (j < nv); V_j = V + j * N
__global__ __launch_bounds__(BWG) void k_bmdot(Chunks C, const int *__restrict__ active, const double *__restrict__ V,
                                               size_t N, int nv, const double *__restrict__ w,
                                               double *__restrict__ partial, int ld) {
    __shared__ double red[BWG / 64][64];
    const int k = blockIdx.x;
    if (active && !active[C.inst[k]]) return;   // converged instance: its partial sums are not read
    const int e0 = C.c0[k], e1 = C.c1[k];
    double wv[BPT];
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int e = e0 + threadIdx.x + q * BWG;
        wv[q] = e < e1 ? w[e] : 0.0;
    }
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j0 = 0; j0 < nv; j0 += 64) {
        const int jn = min(64, nv - j0);
        for (int jj = 0; jj < jn; jj++) {
            const double *v = V + (size_t)(j0 + jj) * N;
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < BPT; q++) {
                const int e = e0 + threadIdx.x + q * BWG;
                if (e < e1) s = fma(v[e], wv[q], s);
            }
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) red[wid][jj] = s;
        }
        __syncthreads();
        if (threadIdx.x < jn) {
            double s = 0.0;
            for (int q = 0; q < BWG / 64; q++) s += red[q][threadIdx.x];
            partial[(size_t)k * ld + j0 + threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// out[inst * ldo + j] = sum over the chunks of inst of partial[k * ld + j]  (grid: ninst x nv)
__global__ __launch_bounds__(64) void k_binst_sum(const int *__restrict__ ckptr, const double *__restrict__ partial,
                                                  int ld, double *__restrict__ out, int ldo, int sqrt_out) {
    const int inst = blockIdx.x, j = blockIdx.y;
    double s = 0.0;
    for (int k = ckptr[inst] + threadIdx.x; k < ckptr[inst + 1]; k += 64) s += partial[(size_t)k * ld + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (threadIdx.x == 0) out[(size_t)inst * ldo + j] = sqrt_out ? sqrt(s) : s;
}

// w -= sum_{j<nv} h[inst][j] V_j over each chunk; partial[k] = sum over chunk k of w^2
__global__ __launch_bounds__(BWG) void k_borth(Chunks C, const int *__restrict__ active, const double *__restrict__ V,
                                               size_t N, int nv, const double *__restrict__ H, int ldh,
                                               double *__restrict__ w, double *__restrict__ partial) {
    __shared__ double red[BWG / 64];
    const int k = blockIdx.x;
    if (active && !active[C.inst[k]]) return;
    const int e0 = C.c0[k], e1 = C.c1[k];
    const double *h = H + (size_t)C.inst[k] * ldh;
    double wv[BPT];
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int e = e0 + threadIdx.x + q * BWG;
        wv[q] = e < e1 ? w[e] : 0.0;
    }
    for (int j = 0; j < nv; j++) {
        const double hj = h[j];
        const double *v = V + (size_t)j * N;
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int e = e0 + threadIdx.x + q * BWG;
            if (e < e1) wv[q] = fma(-hj, v[e], wv[q]);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int e = e0 + threadIdx.x + q * BWG;
        if (e < e1) {
            w[e] = wv[q];
            s = fma(wv[q], wv[q], s);
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < BWG / 64; q++) t += red[q];
        partial[k] = t;
    }
}

// partial[k] = sum over chunk k of x^2
__global__ __launch_bounds__(BWG) void k_bnorm2(Chunks C, const double *__restrict__ x, double *__restrict__ partial) {
    __shared__ double red[BWG / 64];
    const int k = blockIdx.x;
    const int e0 = C.c0[k], e1 = C.c1[k];
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int e = e0 + threadIdx.x + q * BWG;
        if (e < e1) s = fma(x[e], x[e], s);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < BWG / 64; q++) t += red[q];
        partial[k] = t;
    }
}

// v = w / norm[inst] per chunk (0 when the norm is 0)
__global__ __launch_bounds__(BWG) void k_bscale(Chunks C, const int *__restrict__ active, const double *__restrict__ w,
                                                double *__restrict__ v, const double *__restrict__ norm, int ldn) {
    const int k = blockIdx.x;
    if (active && !active[C.inst[k]]) return;
    const int e0 = C.c0[k], e1 = C.c1[k];
    const double nv = norm[(size_t)C.inst[k] * ldn];
    const double inv = nv > 0.0 ? 1.0 / nv : 0.0;
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int e = e0 + threadIdx.x + q * BWG;
        if (e < e1) v[e] = w[e] * inv;
    }
}

// w = r - w
__global__ void k_bsub_from(const double *__restrict__ r, double *__restrict__ w, size_t N) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) w[e] = r[e] - w[e];
}

__global__ void k_bfill(double *__restrict__ x, double v, size_t N) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) x[e] = v;
}

__global__ void k_bfill_int(int *__restrict__ x, int v, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] = v;
}

// per-instance Krylov state
struct BState {
    double *H;        // ninst x m x (m+1): column i of instance q at H + (q*m + i)*(m+1)
    double *hcol;     // ninst x (m+1): the current Arnoldi column (h_0..h_i, then ||w||)
    double *cs, *sn;  // ninst x m
    double *g;        // ninst x (m+1)
    double *beta;     // ninst: initial residual norm of the cycle
    double *thr;      // ninst: absolute tolerance rtol * ||b_q||
    int *active;      // ninst: 1 while the instance is iterating in this cycle
    int *kdim;        // ninst: Krylov dimension used for the update
    int *conv;        // ninst: converged at the end of this cycle
    int *iters;       // ninst: Arnoldi steps taken (accumulated)
    double *resid;    // ninst: residual estimate at the end of the cycle
    int m;
};

// start of a cycle: g = beta e1; an instance whose residual is already below its threshold
// stays inactive (Krylov dimension 0)
__global__ void k_bcycle_init(int ninst, BState S) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ninst) return;
    const int m = S.m;
    const double b = S.beta[q];
    for (int r = 0; r <= m; r++) S.g[(size_t)q * (m + 1) + r] = 0.0;
    S.g[(size_t)q * (m + 1)] = b;
    const int act = (b >= S.thr[q] && b > 0.0) ? 1 : 0;
    S.active[q] = act;
    S.kdim[q] = 0;
    S.conv[q] = act ? 0 : 1;
    S.resid[q] = b;
}

// Arnoldi step i: store the Hessenberg column, apply the previous rotations, form the new
// one (c = rd/dd, s = -hnew/dd as in src/BlockMatrix.cpp:137-145), update g and test
// |g_{i+1}| < thr
__global__ void k_bgivens(int ninst, int i, BState S) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ninst || !S.active[q]) return;
    const int m = S.m;
    const double *hc = S.hcol + (size_t)q * (m + 1);
    double *h = S.H + ((size_t)q * m + i) * (m + 1);
    double *cs = S.cs + (size_t)q * m, *sn = S.sn + (size_t)q * m, *g = S.g + (size_t)q * (m + 1);
    for (int r = 0; r <= i; r++) h[r] = hc[r];
    const double hnew = hc[i + 1];
    for (int k = 0; k < i; k++) {
        const double t = cs[k] * h[k] - sn[k] * h[k + 1];
        h[k + 1] = sn[k] * h[k] + cs[k] * h[k + 1];
        h[k] = t;
    }
    const double rd = h[i];
    const double dd = sqrt(rd * rd + hnew * hnew);
    cs[i] = dd > 0.0 ? rd / dd : 1.0;
    sn[i] = dd > 0.0 ? -hnew / dd : 0.0;
    h[i] = dd;
    g[i + 1] = sn[i] * g[i];
    g[i] = cs[i] * g[i];
    S.iters[q]++;
    S.kdim[q] = i + 1;
    S.resid[q] = fabs(g[i + 1]);
    if (!(fabs(g[i + 1]) >= S.thr[q])) {   // converged (or non-finite: stop and report)
        S.active[q] = 0;
        S.conv[q] = 1;
    }
    if (i + 1 == m) S.active[q] = 0;
}

// least squares R y = g for each instance, y into ybuf (ninst x m)
__global__ void k_bsolve(int ninst, BState S, double *__restrict__ ybuf) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ninst) return;
    const int m = S.m, kd = S.kdim[q];
    double *y = ybuf + (size_t)q * m;
    const double *g = S.g + (size_t)q * (m + 1);
    for (int r = 0; r < m; r++) y[r] = r < kd ? g[r] : 0.0;
    for (int c = kd - 1; c >= 0; c--) {
        const double *h = S.H + ((size_t)q * m + c) * (m + 1);
        y[c] /= h[c];
        for (int r = 0; r < c; r++) y[r] -= h[r] * y[c];
    }
}

// x += sum_j y[inst][j] Z_j over each chunk (j < kdim[inst])
__global__ __launch_bounds__(BWG) void k_bupdate(Chunks C, const double *__restrict__ Z, size_t N, int m,
                                                 const double *__restrict__ ybuf, const int *__restrict__ kdim,
                                                 double *__restrict__ x) {
    const int k = blockIdx.x;
    const int q = C.inst[k];
    const int e0 = C.c0[k], e1 = C.c1[k];
    const int kd = kdim[q];
    if (kd == 0) return;
    const double *y = ybuf + (size_t)q * m;
    double xv[BPT];
#pragma unroll
    for (int t = 0; t < BPT; t++) {
        const int e = e0 + threadIdx.x + t * BWG;
        xv[t] = e < e1 ? x[e] : 0.0;
    }
    for (int j = 0; j < kd; j++) {
        const double yj = y[j];
        const double *z = Z + (size_t)j * N;
#pragma unroll
        for (int t = 0; t < BPT; t++) {
            const int e = e0 + threadIdx.x + t * BWG;
            if (e < e1) xv[t] = fma(yj, z[e], xv[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < BPT; t++) {
        const int e = e0 + threadIdx.x + t * BWG;
        if (e < e1) x[e] = xv[t];
    }
}

// Newton norm per instance: max_c sqrt(sum over the instance's cells of normc[i*nc + c])
// (MeshFn::L2Norm, src/MeshFn.cpp:299-312, per instance)
__global__ __launch_bounds__(256) void k_inst_l2(const int *__restrict__ cptr, int nc, const double *__restrict__ normc,
                                                 double *__restrict__ out) {
    __shared__ double red[4][4];
    const int q = blockIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = cptr[q] + threadIdx.x; i < cptr[q + 1]; i += 256)
        for (int c = 0; c < nc; c++) s[c] += normc[(size_t)i * nc + c];
    for (int c = 0; c < nc; c++) {
        double v = s[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0.0;
        for (int c = 0; c < nc; c++) {
            const double v = sqrt(red[0][c] + red[1][c] + red[2][c] + red[3][c]);
            if (!(v <= mx)) mx = v;   // NaN propagates
        }
        out[q] = mx;
    }
}

// zero the segments of instances whose flag is set (converged instances get a zero rhs)
__global__ __launch_bounds__(BWG) void k_bzero_done(Chunks C, const int *__restrict__ done, double *__restrict__ x) {
    const int k = blockIdx.x;
    if (!done[C.inst[k]]) return;
    const int e0 = C.c0[k], e1 = C.c1[k];
#pragma unroll
    for (int t = 0; t < BPT; t++) {
        const int e = e0 + threadIdx.x + t * BWG;
        if (e < e1) x[e] = 0.0;
    }
}

struct LinCombB {
    const double *v[6];
    double a[6];
    int nv;
};

__global__ void k_blincomb(LinCombB L, double *__restrict__ out, size_t N) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    double s = 0.0;
    for (int j = 0; j < L.nv; j++) s = fma(L.a[j], L.v[j][e], s);
    out[e] = s;
}

__global__ void k_baxpy(double *__restrict__ x, double a, const double *__restrict__ y, size_t N) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) x[e] = fma(a, y[e], x[e]);
}

}  // namespace cpg

using namespace cpg;

struct cpg_batch_s {
    struct Class {
        cpg_equation E = nullptr;
        int n = 0, bl = 0, nnzb = 0, nc = 1;
        size_t off = 0;            // offset of the class in the batch vector
        int inst0 = 0, ninst = 0;  // instances inst0 .. inst0+ninst-1
        std::vector<int> cptr;     // instance cell ranges (local to the class)
        cpg_bsr J = nullptr;
        cpg_prec P = nullptr;
        DevBuf<double> cdt;        // per-cell dt
        DevBuf<int> d_cptr;
    };
    std::vector<Class> cls;
    int ninst = 0;
    size_t N = 0;
    int ns = 1, m = 30;
    cpg_solver_params sp{};
    double A[25], b[5], c[5];
    std::vector<double> dt;        // per instance
    // chunk table
    int nchunk = 0;
    DevBuf<int> ck0, ck1, ckinst, ckptr;
    // vectors
    DevBuf<double> u, k, cu, rhs, r, x, w, V, Z, partial, ybuf;
    // per-instance state
    DevBuf<double> H, hcol, cs, sn, g, beta, thr, resid, nres;
    DevBuf<int> active, kdim, conv, iters, done;
    std::vector<int> h_conv, h_iters, h_done;
    std::vector<double> h_resid, h_nres;
    hipStream_t s = nullptr;
    bool use_graph = true;
    int kblock = 5;                         // Arnoldi steps per captured block
    hipGraphExec_t g_start = nullptr, g_finish = nullptr;
    std::vector<hipGraphExec_t> g_steps;

    Chunks chunks() const { return Chunks{ck0.p, ck1.p, ckinst.p}; }
    BState state() {
        return BState{H.p, hcol.p, cs.p, sn.p, g.p, beta.p, thr.p, active.p, kdim.p, conv.p, iters.p, resid.p, m};
    }
    ~cpg_batch_s() {
        if (g_start) (void)hipGraphExecDestroy(g_start);
        if (g_finish) (void)hipGraphExecDestroy(g_finish);
        for (auto g : g_steps)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto &C : cls) {
            if (C.P) cpg_prec_destroy(C.P);
            if (C.J) cpg_bsr_destroy(C.J);
        }
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

int ck(int st) {
    if (st) throw Error{st};
    return st;
}

// segmented norms of x into out (ninst entries)
void seg_norm(cpg_batch_s &B, const double *x, double *out) {
    hipLaunchKernelGGL(k_bnorm2, dim3(B.nchunk), dim3(BWG), 0, B.s, B.chunks(), x, B.partial.p);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_binst_sum, dim3(B.ninst, 1), dim3(64), 0, B.s, B.ckptr.p, B.partial.p, 1, out, 1, 1);
    CPG_LAUNCH_CHECK();
}

// y = A x over every class (A = the Newton matrices)
void batch_matvec(cpg_batch_s &B, const double *x, double *y) {
    for (auto &C : B.cls) bsr_matvec(*C.J, x + C.off, y + C.off, B.s);
}

// A restart cycle of the batched flexible GMRES for A x = r (x updated in place), in three
// parts so that the host can stop a cycle early once every instance has converged:
// start (residual, beta, V_0), Arnoldi steps i0..i1-1, finish (least squares, x update).
void gmres_start(cpg_batch_s &B) {
    const size_t N = B.N;
    BState S = B.state();
    double *w = B.w.p;
    // w = r - A x ; beta = ||w|| per instance ; V_0 = w / beta
    batch_matvec(B, B.x.p, w);
    hipLaunchKernelGGL(k_bsub_from, dim3(ceil_div(N, 256)), dim3(256), 0, B.s, B.r.p, w, N);
    CPG_LAUNCH_CHECK();
    seg_norm(B, w, B.beta.p);
    hipLaunchKernelGGL(k_bcycle_init, dim3(ceil_div(B.ninst, 64)), dim3(64), 0, B.s, B.ninst, S);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bscale, dim3(B.nchunk), dim3(BWG), 0, B.s, B.chunks(), nullptr, w, B.V.p, B.beta.p, 1);
    CPG_LAUNCH_CHECK();
}

void gmres_steps(cpg_batch_s &B, int i0, int i1) {
    const size_t N = B.N;
    const int m = B.m;
    BState S = B.state();
    double *V = B.V.p, *Z = B.Z.p, *w = B.w.p;
    const int *act = B.active.p;
    for (int i = i0; i < i1; i++) {
        double *vi = V + N * i, *vi1 = V + N * (i + 1), *zi = Z + N * i;
        for (auto &C : B.cls) prec_apply_to(C.P, vi + C.off, zi + C.off, B.s);   // z_i = M^-1 v_i
        batch_matvec(B, zi, w);                                                // w = A z_i
        // classical Gram-Schmidt: h = V^T w ; w -= V h ; ||w||  (segmented per instance)
        hipLaunchKernelGGL(k_bmdot, dim3(B.nchunk), dim3(BWG), 0, B.s, B.chunks(), act, V, N, i + 1, w, B.partial.p,
                           m + 1);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_binst_sum, dim3(B.ninst, i + 1), dim3(64), 0, B.s, B.ckptr.p, B.partial.p, m + 1, B.hcol.p,
                           m + 1, 0);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_borth, dim3(B.nchunk), dim3(BWG), 0, B.s, B.chunks(), act, V, N, i + 1, B.hcol.p, m + 1, w,
                           B.partial.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_binst_sum, dim3(B.ninst, 1), dim3(64), 0, B.s, B.ckptr.p, B.partial.p, 1,
                           B.hcol.p + i + 1, m + 1, 1);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bscale, dim3(B.nchunk), dim3(BWG), 0, B.s, B.chunks(), act, w, vi1, B.hcol.p + i + 1,
                           m + 1);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bgivens, dim3(ceil_div(B.ninst, 64)), dim3(64), 0, B.s, B.ninst, i, S);
        CPG_LAUNCH_CHECK();
    }
}

void gmres_finish(cpg_batch_s &B) {
    BState S = B.state();
    hipLaunchKernelGGL(k_bsolve, dim3(ceil_div(B.ninst, 64)), dim3(64), 0, B.s, B.ninst, S, B.ybuf.p);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bupdate, dim3(B.nchunk), dim3(BWG), 0, B.s, B.chunks(), B.Z.p, B.N, B.m, B.ybuf.p, B.kdim.p,
                       B.x.p);
    CPG_LAUNCH_CHECK();
}

// run `body` on B.s, from a graph captured at its first use when graphs are on
template <class Fn>
void run_part(cpg_batch_s &B, hipGraphExec_t &exec, Fn &&body) {
    if (!B.use_graph) {
        body();
        return;
    }
    if (!exec) {
        CPG_HIP(hipStreamBeginCapture(B.s, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(B.s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        CPG_HIP(hipStreamEndCapture(B.s, &g));
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CPG_HIP(e);
    }
    CPG_HIP(hipGraphLaunch(exec, B.s));
}

bool any_active(cpg_batch_s &B) {
    CPG_HIP(hipMemcpyAsync(B.h_conv.data(), B.active.p, B.ninst * sizeof(int), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipStreamSynchronize(B.s));
    for (int q = 0; q < B.ninst; q++)
        if (B.h_conv[q]) return true;
    return false;
}

// solve A x = r for every instance (x = 0 start, FGMRES tolerance rtol * ||r_q||)
void batch_solve(cpg_batch_s &B, cpg_step_stats &st) {
    const size_t N = B.N;
    hipLaunchKernelGGL(k_bfill, dim3(ceil_div(N, 256)), dim3(256), 0, B.s, B.x.p, 0.0, N);
    CPG_LAUNCH_CHECK();
    seg_norm(B, B.r.p, B.thr.p);
    // thr = rtol * ||r_q|| (FGMRES tolerance relative to the right-hand side of each instance)
    std::vector<double> nb(B.ninst);
    CPG_HIP(hipMemcpyAsync(nb.data(), B.thr.p, B.ninst * sizeof(double), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipStreamSynchronize(B.s));
    for (auto &v : nb) {
        CPG_CHECK(std::isfinite(v), CPG_ERR_NONFINITE, "non-finite right-hand side norm in the batch");
        v *= B.sp.rtol;
    }
    CPG_HIP(hipMemcpyAsync(B.thr.p, nb.data(), B.ninst * sizeof(double), hipMemcpyHostToDevice, B.s));
    hipLaunchKernelGGL(k_bfill_int, dim3(ceil_div(B.ninst, 64)), dim3(64), 0, B.s, B.iters.p, 0, B.ninst);
    CPG_LAUNCH_CHECK();
    const int maxcyc = std::max(1, (B.sp.maxit + B.m - 1) / B.m);
    int cyc = 0;
    const int nblk = (B.m + B.kblock - 1) / B.kblock;
    if (B.g_steps.empty()) B.g_steps.assign(nblk, nullptr);
    for (; cyc < maxcyc; cyc++) {
        run_part(B, B.g_start, [&] { gmres_start(B); });
        // Arnoldi steps in blocks of kblock, stopping once every instance has converged
        for (int bk = 0; bk < nblk && any_active(B); bk++) {
            const int i0 = bk * B.kblock, i1 = std::min(B.m, i0 + B.kblock);
            run_part(B, B.g_steps[bk], [&] { gmres_steps(B, i0, i1); });
        }
        run_part(B, B.g_finish, [&] { gmres_finish(B); });
        CPG_HIP(hipMemcpyAsync(B.h_conv.data(), B.conv.p, B.ninst * sizeof(int), hipMemcpyDeviceToHost, B.s));
        CPG_HIP(hipStreamSynchronize(B.s));
        bool all = true;
        for (int q = 0; q < B.ninst; q++) all = all && B.h_conv[q];
        if (all) break;
    }
    CPG_HIP(hipMemcpyAsync(B.h_iters.data(), B.iters.p, B.ninst * sizeof(int), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipMemcpyAsync(B.h_resid.data(), B.resid.p, B.ninst * sizeof(double), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipStreamSynchronize(B.s));
    int mx = 0;
    for (int q = 0; q < B.ninst; q++) {
        mx = std::max(mx, B.h_iters[q]);
        CPG_CHECK(std::isfinite(B.h_resid[q]), CPG_ERR_NONFINITE, "non-finite GMRES residual in batch instance %d", q);
    }
    st.gmres_iters += mx;
    st.linear_solves++;
    CPG_CHECK(cyc < maxcyc, CPG_ERR_NOT_CONVERGED, "batched GMRES did not converge in %d cycles of %d", maxcyc, B.m);
}

}  // namespace

extern "C" {

int cpg_batch_create(int nclass, const cpg_equation *eqs, const int *ninst, const int *const *cell_ptr, int nstages,
                     const cpg_solver_params *params, int use_graph, cpg_batch *out) {
    return guard([&] {
        CPG_CHECK(nclass > 0 && eqs && ninst && cell_ptr && params && out, CPG_ERR_ARG, "null argument");
        CPG_CHECK(nstages >= 1 && nstages <= 5, CPG_ERR_ARG, "bad stage count %d", nstages);
        CPG_CHECK(params->restart > 0 && params->restart <= 200 && params->maxit > 0 && params->newton_maxit > 0,
                  CPG_ERR_ARG, "bad solver parameters");
        CPG_CHECK(params->pc_kind == CPG_PC_DILU || params->pc_kind == CPG_PC_BJACOBI || params->pc_kind == CPG_PC_NONE,
                  CPG_ERR_ARG, "batched solves take DILU, block-Jacobi or no preconditioner");
        auto *B = new cpg_batch_s;
        try {
            B->sp = *params;
            B->ns = nstages;
            B->m = params->restart;
            tableau(CPG_STEP_DIRK, nstages, B->A, B->b, B->c, params->dirk_tableau);
            B->use_graph = use_graph != 0;
            CPG_HIP(hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking));
            size_t off = 0;
            std::vector<int> c0, c1, cinst, ckptr(1, 0);
            for (int c = 0; c < nclass; c++) {
                cpg_batch_s::Class C;
                C.E = eqs[c];
                CPG_CHECK(C.E && ninst[c] > 0 && cell_ptr[c], CPG_ERR_ARG, "class %d: bad equation or instances", c);
                ck(cpg_equation_info(C.E, &C.n, &C.bl, &C.nnzb));
                C.nc = C.bl / equation_bs(*C.E);
                C.cptr.assign(cell_ptr[c], cell_ptr[c] + ninst[c] + 1);
                CPG_CHECK(C.cptr[0] == 0 && C.cptr[ninst[c]] == C.n, CPG_ERR_ARG,
                          "class %d: instance cell ranges do not span the %d cells", c, C.n);
                for (int q = 0; q < ninst[c]; q++)
                    CPG_CHECK(C.cptr[q + 1] > C.cptr[q], CPG_ERR_ARG, "class %d: empty instance %d", c, q);
                C.off = off;
                C.inst0 = B->ninst;
                C.ninst = ninst[c];
                for (int q = 0; q < ninst[c]; q++) {
                    const int e0 = (int)off + C.cptr[q] * C.bl, e1 = (int)off + C.cptr[q + 1] * C.bl;
                    for (int e = e0; e < e1; e += BCH) {
                        c0.push_back(e);
                        c1.push_back(std::min(e + BCH, e1));
                        cinst.push_back(B->ninst + q);
                    }
                    ckptr.push_back((int)c0.size());
                }
                B->ninst += ninst[c];
                off += (size_t)C.n * C.bl;
                CPG_CHECK(off < (size_t)1 << 31, CPG_ERR_UNSUPPORTED, "batch vector longer than 2^31");
                ck(cpg_jacobian_create(C.E, &C.J));
                ck(cpg_prec_create(C.J, params->pc_kind, params->precision, &C.P));
                C.cdt.alloc(C.n);
                C.d_cptr.upload(C.cptr);
                // launch shapes of this class's block size: matvec (FP64) and sweeps (factor precision)
                int best = 0;
                ck(cpg_stream_autotune(C.J, CPG_FP64, 3, &best, nullptr, B->s));
                if (params->precision != CPG_FP64)   // the stream's storage (BF16 for BF16 with FP32 D^-1)
                    ck(cpg_stream_autotune(C.J, params->precision == CPG_BF16_D32 ? CPG_BF16 : params->precision, 3,
                                           &best, nullptr, B->s));
                B->cls.push_back(std::move(C));
            }
            B->N = off;
            B->nchunk = (int)c0.size();
            B->ck0.upload(c0);
            B->ck1.upload(c1);
            B->ckinst.upload(cinst);
            B->ckptr.upload(ckptr);
            const size_t N = B->N;
            const int m = B->m, ni = B->ninst;
            B->u.alloc(N);
            B->k.alloc(N * B->ns);
            B->cu.alloc(N);
            B->rhs.alloc(N);
            B->r.alloc(N);
            B->x.alloc(N);
            B->w.alloc(N);
            B->V.alloc(N * (m + 1));
            B->Z.alloc(N * m);
            B->partial.alloc((size_t)B->nchunk * (m + 1));
            B->ybuf.alloc((size_t)ni * m);
            B->H.alloc((size_t)ni * m * (m + 1));
            B->hcol.alloc((size_t)ni * (m + 1));
            B->cs.alloc((size_t)ni * m);
            B->sn.alloc((size_t)ni * m);
            B->g.alloc((size_t)ni * (m + 1));
            B->beta.alloc(ni);
            B->thr.alloc(ni);
            B->resid.alloc(ni);
            B->nres.alloc(ni);
            B->active.alloc(ni);
            B->kdim.alloc(ni);
            B->conv.alloc(ni);
            B->iters.alloc(ni + 1);
            B->done.alloc(ni);
            B->h_conv.assign(ni, 0);
            B->h_iters.assign(ni, 0);
            B->h_done.assign(ni, 0);
            B->h_resid.assign(ni, 0.0);
            B->h_nres.assign(ni, 0.0);
            B->dt.assign(ni, 0.0);
            CPG_HIP(hipStreamSynchronize(B->s));
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

int cpg_batch_destroy(cpg_batch B) {
    return guard([&] { delete B; });
}

int cpg_batch_info(cpg_batch B, int *ninst, int *nchunk, int *graph_captured) {
    return guard([&] {
        CPG_CHECK(B, CPG_ERR_ARG, "null batch");
        if (ninst) *ninst = B->ninst;
        if (nchunk) *nchunk = B->nchunk;
        if (graph_captured) *graph_captured = B->g_start != nullptr;
    });
}

int cpg_batch_advance(cpg_batch B, const double *const *u, const double *dt, double t, double *const *unp1,
                      cpg_step_stats *stats, int *inst_gmres_iters, void *stream) {
    cpg_step_stats st{};
    int ret = guard([&] {
        CPG_CHECK(B && u && dt && unp1, CPG_ERR_ARG, "null argument");
        for (int q = 0; q < B->ninst; q++) CPG_CHECK(dt[q] > 0.0 && std::isfinite(dt[q]), CPG_ERR_ARG, "bad dt[%d]", q);
        hipStream_t us = S(stream);
        // the batch runs on its own stream (graph capture needs a non-default stream): order it
        // after the caller's stream
        CPG_HIP(hipStreamSynchronize(us));
        const size_t N = B->N;
        const int ns = B->ns;
        const int gN = ceil_div(N, 256);
        // per-cell dt and the batch copy of u
        for (auto &C : B->cls) {
            std::vector<double> cd(C.n);
            for (int q = 0; q < C.ninst; q++)
                for (int i = C.cptr[q]; i < C.cptr[q + 1]; i++) cd[i] = dt[C.inst0 + q];
            CPG_HIP(hipMemcpyAsync(C.cdt.p, cd.data(), C.n * sizeof(double), hipMemcpyHostToDevice, B->s));
            const int ci = (int)(&C - B->cls.data());
            CPG_CHECK(u[ci] && unp1[ci], CPG_ERR_ARG, "null vector for class %d", ci);
            CPG_HIP(hipMemcpyAsync(B->u.p + C.off, u[ci], (size_t)C.n * C.bl * sizeof(double), hipMemcpyDeviceToDevice,
                                   B->s));
        }
        hipLaunchKernelGGL(k_bfill, dim3(ceil_div(N * ns, 256)), dim3(256), 0, B->s, B->k.p, 0.0, N * ns);
        CPG_LAUNCH_CHECK();
        if (inst_gmres_iters) std::fill(inst_gmres_iters, inst_gmres_iters + B->ninst, 0);
        bool all_conv = true;
        for (int stg = 0; stg < ns; stg++) {
            double *ks = B->k.p + N * stg;
            const double ts_off = B->c[stg];
            std::fill(B->h_done.begin(), B->h_done.end(), 0);
            bool converged = false;
            for (int it = 0; it < B->sp.newton_maxit; it++) {
                // cu = u + sum_j A_sj k_j ; r = M k_s - dt_q b(cu) per instance
                LinCombB L{};
                L.v[L.nv] = B->u.p;
                L.a[L.nv++] = 1.0;
                for (int j = 0; j <= stg; j++)
                    if (B->A[stg * ns + j] != 0.0) {
                        L.v[L.nv] = B->k.p + N * j;
                        L.a[L.nv++] = B->A[stg * ns + j];
                    }
                hipLaunchKernelGGL(k_blincomb, dim3(gN), dim3(256), 0, B->s, L, B->cu.p, N);
                CPG_LAUNCH_CHECK();
                for (auto &C : B->cls) {
                    // one stage time per class launch, t + c_s dt of the class's first instance
                    // (exact for steady fields with time-independent Dirichlet data)
                    eq_residual(*C.E, B->cu.p + C.off, t + ts_off * dt[C.inst0], B->rhs.p + C.off, B->s);
                    eq_newton_resid_cells(*C.E, ks + C.off, B->rhs.p + C.off, 1.0, C.cdt.p, B->r.p + C.off, B->s);
                    hipLaunchKernelGGL(k_inst_l2, dim3(C.ninst), dim3(256), 0, B->s, C.d_cptr.p, C.nc,
                                       eq_normc(*C.E), B->nres.p + C.inst0);
                    CPG_LAUNCH_CHECK();
                }
                CPG_HIP(hipMemcpyAsync(B->h_nres.data(), B->nres.p, B->ninst * sizeof(double), hipMemcpyDeviceToHost,
                                       B->s));
                CPG_HIP(hipStreamSynchronize(B->s));
                bool all = true;
                double worst = 0.0;
                for (int q = 0; q < B->ninst; q++) {
                    CPG_CHECK(std::isfinite(B->h_nres[q]), CPG_ERR_NONFINITE,
                              "non-finite Newton residual (instance %d, stage %d, it %d)", q, stg, it);
                    B->h_done[q] = B->h_nres[q] < B->sp.newton_tol;
                    all = all && B->h_done[q];
                    worst = std::max(worst, B->h_nres[q]);
                }
                st.last_resid = worst;
                if (all) { converged = true; break; }
                st.newton_iters++;
                // B = M - dt_q a_ss J(cu), factor, solve B x = r (converged instances: r = 0), k_s -= x
                for (auto &C : B->cls) {
                    eq_jacobian_cells(*C.E, B->cu.p + C.off, t + ts_off * dt[C.inst0], -B->A[stg * ns + stg], C.cdt.p, 1,
                                      C.J->blocks.p, B->s);
                    ck(cpg_prec_factor(C.P, C.J, B->s));
                }
                CPG_HIP(hipMemcpyAsync(B->done.p, B->h_done.data(), B->ninst * sizeof(int), hipMemcpyHostToDevice, B->s));
                hipLaunchKernelGGL(k_bzero_done, dim3(B->nchunk), dim3(BWG), 0, B->s, B->chunks(), B->done.p, B->r.p);
                CPG_LAUNCH_CHECK();
                batch_solve(*B, st);
                if (inst_gmres_iters)
                    for (int q = 0; q < B->ninst; q++) inst_gmres_iters[q] += B->h_iters[q];
                hipLaunchKernelGGL(k_baxpy, dim3(gN), dim3(256), 0, B->s, ks, -1.0, B->x.p, N);
                CPG_LAUNCH_CHECK();
            }
            if (!converged) all_conv = false;
        }
        // unp1 = u + sum_j b_j k_j
        LinCombB L{};
        L.v[L.nv] = B->u.p;
        L.a[L.nv++] = 1.0;
        for (int j = 0; j < ns; j++)
            if (B->b[j] != 0.0) {
                L.v[L.nv] = B->k.p + N * j;
                L.a[L.nv++] = B->b[j];
            }
        hipLaunchKernelGGL(k_blincomb, dim3(gN), dim3(256), 0, B->s, L, B->w.p, N);
        CPG_LAUNCH_CHECK();
        for (auto &C : B->cls) {
            const int ci = (int)(&C - B->cls.data());
            CPG_HIP(hipMemcpyAsync(unp1[ci], B->w.p + C.off, (size_t)C.n * C.bl * sizeof(double),
                                   hipMemcpyDeviceToDevice, B->s));
        }
        CPG_HIP(hipStreamSynchronize(B->s));
        if (!all_conv) {
            st.status = CPG_ERR_NOT_CONVERGED;
            set_error("batched Newton did not reach %g in %d iterations (worst instance %g)", B->sp.newton_tol,
                      B->sp.newton_maxit, st.last_resid);
            throw Error{CPG_ERR_NOT_CONVERGED};
        }
    });
    st.status = ret;
    if (stats) *stats = st;
    return ret;
}

}  // extern "C"
