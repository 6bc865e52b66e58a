// Restarted GMRES on the device (BlockMatrix::gmres, src/BlockMatrix.cpp:79-163).
//
// Vectors (Krylov basis V, and Z for the flexible variant) live in HBM.  Classical
// Gram-Schmidt is two passes over V: a multi-dot h = V^T w (one read of V and w), then
// w -= V h fused with the partial sums of ||w||^2.  All reductions are two-stage with a
// fixed order (deterministic).  The operator is an abstraction (LinOp): the BSR matvec of an
// assembled matrix, or any other y = A x the caller supplies (the stepper's matrix-free
// Newton operator).
//
// Two drivers:
//  * CPG_GMRES_LEFT_REF, the reference algorithm: left preconditioning, tolerance relative to
//    ||P^-1 b||; the Hessenberg column comes to the host once per Arnoldi step (i+2 doubles)
//    and the Givens rotations and triangular solve run there in FP64 in the reference's order
//    (iteration-count parity).
//  * CPG_GMRES_FLEXIBLE: right-preconditioned flexible GMRES with everything on the device.
//    The Hessenberg column, Givens rotations (c = rd/dd, s = -hnew/dd as in
//    src/BlockMatrix.cpp:137-145) and the convergence flag are FP64 kernels of one thread, the
//    least squares one wavefront; a restart cycle is a start part, blocks of KBLK Arnoldi steps and a finish
//    part, each captured once into a hipGraph and replayed for every later solve with the same
//    operands.  The host launches block b+1 before it reads the flag that block b left, so the
//    device never waits on it; once the flag reads "converged" every kernel of a replayed step
//    returns at once (the operator and preconditioner launches take the flag too).  With kernel
//    timing on, the same sequence runs eagerly with events around every operator and
//    preconditioner application.
#include "cpg_internal.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace cpg {

constexpr int KWG = 256;
constexpr int KPT = 8;  // elements per thread in the vector kernels
constexpr int KCH = KWG * KPT;
// Arnoldi steps per captured block (flexible driver).  Each graph launch costs a ~13 us gap before
// its first kernel, and the look-ahead block queued after the step that converges runs as no-ops
// (~15 launches of ~4.5 us per step): two steps per block balance the two for ~19-step solves.
constexpr int KBLK = 2;

// partial[b*nv + j] = sum over chunk b of V[j] . w.  Two basis vectors at a time, the next two
// vectors' loads issued before this pair's products and wave reductions (one vector at a time,
// every vector cost a full memory round trip, and the reductions stalled the next loads); the
// summation order of every product is that of one vector at a time.
__global__ __launch_bounds__(KWG) void k_mdot(const double *__restrict__ V, size_t N, int nv,
                                              const double *__restrict__ w, double *__restrict__ partial,
                                              const int *__restrict__ active) {
    __shared__ double red[KWG / 64][64];
    if (active && !*active) return;
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
        // software pipeline over pairs of vectors: the next pair's loads are issued before this
        // pair's products and reductions
        double cur[2][KPT], nxt[2][KPT];
        auto load = [&](int jj, double (&buf)[2][KPT]) {
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const bool in = jj + u < jn;
                const double *v = V + (size_t)(j0 + (in ? jj + u : 0)) * N;
#pragma unroll
                for (int k = 0; k < KPT; k++) {
                    const size_t e = base + (size_t)k * KWG;
                    buf[u][k] = (in && e < N) ? v[e] : 0.0;
                }
            }
        };
        load(0, cur);
        for (int jj = 0; jj < jn; jj += 2) {
            if (jj + 2 < jn) load(jj + 2, nxt);
            double s[2] = {0.0, 0.0};
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int k = 0; k < KPT; k++) s[u] = fma(cur[u][k], wv[k], s[u]);
            for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                for (int u = 0; u < 2; u++) s[u] += __shfl_xor(s[u], off);
            if (lane == 0)
#pragma unroll
                for (int u = 0; u < 2; u++)
                    if (jj + u < jn) red[wid][jj + u] = s[u];
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int k = 0; k < KPT; k++) cur[u][k] = nxt[u][k];
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
                                                double *__restrict__ out, int sqrt_out,
                                                const int *__restrict__ active) {
    __shared__ double red[KWG / 64];
    if (active && !*active) return;
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
                                                   double *__restrict__ partial, const int *__restrict__ active) {
    __shared__ double red[KWG / 64];
    if (active && !*active) return;
    const size_t base = (size_t)blockIdx.x * KCH + threadIdx.x;
    double wv[KPT];
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        wv[k] = e < N ? w[e] : 0.0;
    }
    // four basis vectors' loads in flight together; w is updated vector by vector in order
    for (int j0 = 0; j0 < nv; j0 += 4) {
        double vv[4][KPT];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (j0 + u < nv) {
                const double *v = V + (size_t)(j0 + u) * N;
#pragma unroll
                for (int k = 0; k < KPT; k++) {
                    const size_t e = base + (size_t)k * KWG;
                    vv[u][k] = e < N ? v[e] : 0.0;
                }
            }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (j0 + u < nv) {
                const double hj = h[j0 + u];
#pragma unroll
                for (int k = 0; k < KPT; k++) wv[k] = fma(-hj, vv[u][k], wv[k]);
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
                            const double *__restrict__ norm, const int *__restrict__ active) {
    if (active && !*active) return;
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

// r = -b
__global__ void k_neg(double *__restrict__ r, const double *__restrict__ b, size_t N) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) r[e] = -b[e];
}

// x -= sum_k y[k] B[k] for k < nk, with nk from the host or (kdim non-null) from the device
__global__ __launch_bounds__(KWG) void k_update(const double *__restrict__ B, size_t N, int nk,
                                                const int *__restrict__ kdim, const double *__restrict__ y,
                                                double *__restrict__ x) {
    if (kdim) nk = *kdim;
    if (nk == 0) return;
    const size_t base = (size_t)blockIdx.x * KCH + threadIdx.x;
    double xv[KPT];
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        xv[k] = e < N ? x[e] : 0.0;
    }
    for (int j0 = 0; j0 < nk; j0 += 4) {
        double bv[4][KPT];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (j0 + u < nk) {
                const double *b = B + (size_t)(j0 + u) * N;
#pragma unroll
                for (int k = 0; k < KPT; k++) {
                    const size_t e = base + (size_t)k * KWG;
                    bv[u][k] = e < N ? b[e] : 0.0;
                }
            }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (j0 + u < nk) {
                const double yj = y[j0 + u];
#pragma unroll
                for (int k = 0; k < KPT; k++) xv[k] = fma(-yj, bv[u][k], xv[k]);
            }
    }
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        if (e < N) x[e] = xv[k];
    }
}

// ---- device-resident least-squares state of the flexible driver -----------------------
struct FgState {
    double *H;      // m columns of m+1: column i at H + i (m+1)
    double *hcol;   // the current Arnoldi column: h_0..h_i, then ||w|| at i+1
    double *cs, *sn;
    double *g;      // m+1
    double *y;      // m
    double *scal;   // [0] ||b||  [1] beta of the cycle  [2] absolute tolerance  [3] residual estimate
    int *flags;     // [0] active  [1] Krylov dimension  [2] converged  [3] Arnoldi steps (all cycles)
    int m;
};

__global__ void k_fg_setup(FgState S, double rtol) {
    S.scal[2] = rtol * S.scal[0];
    S.flags[3] = 0;
}

// start of a cycle: g = beta e1; nothing to do when beta is already below the tolerance
__global__ void k_fg_cycle_init(FgState S) {
    const double b = S.scal[1];
    for (int r = 0; r <= S.m; r++) S.g[r] = 0.0;
    S.g[0] = b;
    const int act = (b >= S.scal[2] && b > 0.0) ? 1 : 0;
    S.flags[0] = act;
    S.flags[1] = 0;
    S.flags[2] = act ? 0 : 1;
    S.scal[3] = b;
}

// Arnoldi step i: store the column (h_0..h_i from hcol, hnew = ||w||), apply the previous
// rotations, form the new one, update g and test |g_{i+1}| against the tolerance (a non-finite
// estimate also ends the cycle).  One thread.
__device__ inline void fg_givens(int i, const FgState &S, double hnew) {
    const int m = S.m;
    double *h = S.H + (size_t)i * (m + 1);
    for (int r = 0; r <= i; r++) h[r] = S.hcol[r];
    for (int k = 0; k < i; k++) {
        const double t = S.cs[k] * h[k] - S.sn[k] * h[k + 1];
        h[k + 1] = S.sn[k] * h[k] + S.cs[k] * h[k + 1];
        h[k] = t;
    }
    const double rd = h[i];
    const double dd = sqrt(rd * rd + hnew * hnew);
    S.cs[i] = dd > 0.0 ? rd / dd : 1.0;
    S.sn[i] = dd > 0.0 ? -hnew / dd : 0.0;
    h[i] = dd;
    S.g[i + 1] = S.sn[i] * S.g[i];
    S.g[i] = S.cs[i] * S.g[i];
    S.flags[3]++;
    S.flags[1] = i + 1;
    S.scal[3] = fabs(S.g[i + 1]);
    if (!(fabs(S.g[i + 1]) >= S.scal[2])) {
        S.flags[0] = 0;
        S.flags[2] = 1;
    }
    if (i + 1 == m) S.flags[0] = 0;
}

// Step i's tail in one launch: ||w||^2 is the sum of the orthogonalisation's chunk partial sums,
// which every workgroup adds in k_reduce's order (so all hold the same norm, equal to the one
// k_reduce gives); each workgroup scales its chunk, v_{i+1} = w / ||w||; workgroup 0 then stores
// ||w|| in the Hessenberg column and runs the Givens update (fg_givens), and writes the resulting active
// flag to `hflag` (host memory the driver polls; null: none).  A workgroup that starts after
// workgroup 0 cleared the flag skips its chunk; the flag is cleared only when the cycle ends at
// this step, and v_{i+1} is then never read.
__global__ __launch_bounds__(KWG) void k_fg_norm_step(const double *__restrict__ partial, int nb, double *v, size_t N,
                                                      int i, FgState S, int *hflag) {
    __shared__ double red[KWG / 64];
    if (!S.flags[0]) {
        if (blockIdx.x == 0 && threadIdx.x == 0 && hflag) *hflag = 0;
        return;
    }
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += KWG) s += partial[b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < KWG / 64; q++) t += red[q];
    const double nrm = sqrt(t);
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    const size_t base = (size_t)blockIdx.x * KCH + threadIdx.x;
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        if (e < N) v[e] = v[e] * inv;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S.hcol[i + 1] = nrm;
        fg_givens(i, S, nrm);
        if (hflag) *hflag = S.flags[0];
    }
}

// R y = g over the Krylov dimension of the cycle.  Up to 64 columns: one wavefront, the
// triangle staged in LDS, lane r holding y_r (the operations and their order per entry are the
// serial back substitution's); longer ones: one thread.
__global__ __launch_bounds__(64) void k_fg_solve(FgState S) {
    __shared__ double R[64 * 64];
    const int kd = S.flags[1];
    const int lane = threadIdx.x;
    if (kd > 64) {
        if (lane == 0) {
            for (int r = 0; r < kd; r++) S.y[r] = S.g[r];
            for (int c = kd - 1; c >= 0; c--) {
                const double *h = S.H + (size_t)c * (S.m + 1);
                S.y[c] /= h[c];
                for (int r = 0; r < c; r++) S.y[r] -= h[r] * S.y[c];
            }
        }
        return;
    }
    for (int c = 0; c < kd; c++) {
        const double *h = S.H + (size_t)c * (S.m + 1);
        for (int r = lane; r <= c; r += 64) R[c * 64 + r] = h[r];
    }
    double y = lane < kd ? S.g[lane] : 0.0;
    __syncthreads();
    for (int c = kd - 1; c >= 0; c--) {
        const double yc = __shfl(y, c) / R[c * 64 + c];
        if (lane == c) y = yc;
        if (lane < c) y -= R[c * 64 + lane] * yc;
    }
    if (lane < kd) S.y[lane] = y;
}

}  // namespace cpg

using namespace cpg;

struct cpg_krylov_s {
    int n = 0, bl = 0, m = 0, mode = 0;
    size_t N = 0;
    int nblk = 0;
    DevBuf<double> V, Z, r, partial, hdev, ydev, scal;
    std::vector<double> h, y, cs, sn, hcol;
    // flexible driver: device least-squares state, captured graphs, look-ahead flags
    DevBuf<double> fH, fcs, fsn, fg, fy;
    DevBuf<int> fflags;
    int *h_flags = nullptr;   // pinned: one active flag per block, then the final 4 flags
    int *d_flags = nullptr;   // the device's address of h_flags (the step kernels write the block flags)
    std::vector<hipEvent_t> fev;
    // the flexible driver runs on its own stream (capture is not possible on the null stream),
    // ordered after the caller's stream by an event
    hipStream_t kst = nullptr;
    hipEvent_t cev = nullptr;
    struct Sig {
        const double *b = nullptr;
        double *x = nullptr;
        const void *fn = nullptr, *ctx = nullptr, *P = nullptr;
        uint64_t agen = 0, pgen = 0;   // generations of a BSR operator and of P (0 for other kinds)
        bool operator==(const Sig &o) const {
            return b == o.b && x == o.x && fn == o.fn && ctx == o.ctx && P == o.P && agen == o.agen && pgen == o.pgen;
        }
    } sig;
    hipGraphExec_t g_start0 = nullptr, g_start = nullptr, g_finish = nullptr;
    std::vector<hipGraphExec_t> g_blocks;
    bool use_graph = true;   // CPOLYG_GMRES_GRAPH=0 at creation: launch the flexible cycles eagerly
    // kernel timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> evkind;  // 0 operator, 1 preconditioner
    size_t evused = 0;
    // per marker pair: a dispatch-timestamp pair (launch_timed) that replaces it where the timed
    // launch took it
    std::vector<hipEvent_t> dev;
    std::vector<char> dused;
    int n_matvec = 0, n_apply = 0;
    double ms_matvec = 0, ms_apply = 0;

    FgState fstate() {
        return FgState{fH.p, hdev.p, fcs.p, fsn.p, fg.p, fy.p, scal.p, fflags.p, m};
    }
    void drop_graphs() {
        for (auto *g : {&g_start0, &g_start, &g_finish})
            if (*g) {
                (void)hipGraphExecDestroy(*g);
                *g = nullptr;
            }
        for (auto &g : g_blocks)
            if (g) (void)hipGraphExecDestroy(g);
        g_blocks.assign(g_blocks.size(), nullptr);
    }
    ~cpg_krylov_s() {
        drop_graphs();
        for (auto e : ev) (void)hipEventDestroy(e);
        for (auto e : dev) (void)hipEventDestroy(e);
        for (auto e : fev) (void)hipEventDestroy(e);
        if (h_flags) (void)hipHostFree(h_flags);
        if (cev) (void)hipEventDestroy(cev);
        if (kst) (void)hipStreamDestroy(kst);
    }

    void mark(int kind, hipStream_t s) {
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
    }
    // before an operator / preconditioner application (after its opening mark): its timed launch,
    // if it has exactly one, takes the next dispatch-timestamp pair
    void arm() {
        if (!timing) return;
        const size_t p = evused / 2;
        while (dev.size() < 2 * (p + 1)) {
            hipEvent_t e;
            CPG_HIP(hipEventCreate(&e));
            dev.push_back(e);
        }
        if (dused.size() < p + 1) dused.resize(p + 1);
        dispatch_timing() = DispatchTiming{dev[2 * p], dev[2 * p + 1], false};
    }
    void disarm() {
        if (!timing) return;
        dused[evused / 2] = dispatch_timing().used;
        dispatch_timing() = DispatchTiming{};
    }
    void harvest() {
        if (!timing) return;
        for (size_t k = 0; k + 1 < evused; k += 2) {
            float ms = 0;
            if (k / 2 < dused.size() && dused[k / 2]) CPG_HIP(hipEventElapsedTime(&ms, dev[k], dev[k + 1]));
            else CPG_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            if (evkind[k] == 0) { ms_matvec += ms; n_matvec++; }
            else { ms_apply += ms; n_apply++; }
        }
        evused = 0;
    }
};

namespace cpg {

struct KOps {
    cpg_krylov_s &K;
    const LinOp &A;
    cpg_prec P;
    hipStream_t s;
    const PcOp *M = nullptr;   // preconditioner as an operator (flexible driver), instead of P

    void matvec(const double *x, double *y, const int *active = nullptr) {
        K.mark(0, s);
        K.arm();
        A.fn(A.ctx, x, y, active, s);
        K.disarm();
        K.mark(0, s);
    }
    void pc(double *x) {
        if (!P) return;
        K.mark(1, s);
        K.arm();
        int st = cpg_prec_apply(P, x, s);
        K.disarm();
        if (st) throw Error{st};
        K.mark(1, s);
    }
    // out = M^-1 in (a copy when there is no preconditioner)
    void pc_to(const double *in, double *out, const int *active = nullptr) {
        if (M) {
            K.mark(1, s);
            M->fn(M->ctx, in, out, active, s);
            K.mark(1, s);
            return;
        }
        if (!P) {
            CPG_HIP(hipMemcpyAsync(out, in, K.N * sizeof(double), hipMemcpyDeviceToDevice, s));
            return;
        }
        K.mark(1, s);
        K.arm();
        prec_apply_to(P, in, out, s, active);
        K.disarm();
        K.mark(1, s);
    }
    // ||x|| into dst (device), two fixed-order stages
    void norm_dev(const double *x, double *dst) {
        hipLaunchKernelGGL(k_mdot, dim3(K.nblk), dim3(KWG), 0, s, x, K.N, 1, x, K.partial.p, nullptr);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(KWG), 0, s, K.partial.p, K.nblk, 1, dst, 1, nullptr);
        CPG_LAUNCH_CHECK();
    }
    // returns ||x|| (synchronous)
    double norm(const double *x) {
        norm_dev(x, K.scal.p);
        double v;
        CPG_HIP(hipMemcpyAsync(&v, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
        CPG_HIP(hipStreamSynchronize(s));
        return v;
    }
};

// ---- reference driver: left preconditioning, host Givens -------------------------------
void gmres_left_ref(cpg_krylov_s *K, const LinOp &A, cpg_prec P, const double *b, double *x, double *tol, int *maxit,
                    hipStream_t s) {
    KOps op{*K, A, P, s};
    const size_t N = K->N;
    const int m = K->m;
    double *V = K->V.p;
    double *r = K->r.p;
    const double tl = *tol;
    int mx = *maxit;

    CPG_HIP(hipMemcpyAsync(r, b, N * sizeof(double), hipMemcpyDeviceToDevice, s));
    op.pc(r);
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
        // r = P^-1 (A x - b)
        op.matvec(x, r);
        hipLaunchKernelGGL(k_sub, dim3(ceil_div(N, 256)), dim3(256), 0, s, r, b, N);
        CPG_LAUNCH_CHECK();
        op.pc(r);
        const double beta = op.norm(r);
        CPG_CHECK(std::isfinite(beta), CPG_ERR_NONFINITE,
                  "non-finite residual norm at restart %d (previous cycle %d Arnoldi steps, |y_i| %g, ||b|| %g)", j, i,
                  std::fabs(y[i]), nrm2b);
        hipLaunchKernelGGL(k_scale_by, dim3(ceil_div(N, 256)), dim3(256), 0, s, r, V, N, beta > 0.0 ? 1.0 / beta : 0.0);
        CPG_LAUNCH_CHECK();
        y[0] = beta;
        int uki = 0;
        for (i = 0; i < m; i++) {
            if (std::fabs(y[i]) < tl * nrm2b) break;
            double *vi = V + N * i;
            double *vi1 = V + N * (i + 1);
            const int u0i = uki;
            op.matvec(vi, vi1);
            op.pc(vi1);
            // classical Gram-Schmidt: h = V^T w ; w -= V h ; hnew = ||w||
            hipLaunchKernelGGL(k_mdot, dim3(K->nblk), dim3(KWG), 0, s, V, N, i + 1, vi1, K->partial.p, nullptr);
            CPG_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_reduce, dim3(i + 1), dim3(KWG), 0, s, K->partial.p, K->nblk, i + 1, K->hdev.p, 0,
                               nullptr);
            CPG_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_orth_norm, dim3(K->nblk), dim3(KWG), 0, s, V, N, i + 1, K->hdev.p, vi1, K->partial.p,
                               nullptr);
            CPG_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(KWG), 0, s, K->partial.p, K->nblk, 1, K->hdev.p + i + 1, 1,
                               nullptr);
            CPG_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_scale_dev, dim3(ceil_div(N, 256)), dim3(256), 0, s, vi1, vi1, N, K->hdev.p + i + 1,
                               nullptr);
            CPG_LAUNCH_CHECK();
            CPG_HIP(hipMemcpyAsync(K->hcol.data(), K->hdev.p, (i + 2) * sizeof(double), hipMemcpyDeviceToHost, s));
            CPG_HIP(hipStreamSynchronize(s));
            for (int k = 0; k <= i; k++) h[u0i + k] = K->hcol[k];
            const double hnew = K->hcol[i + 1];
            bool hfin = std::isfinite(hnew);
            for (int k = 0; k <= i; k++) hfin = hfin && std::isfinite(K->hcol[k]);
            CPG_CHECK(hfin, CPG_ERR_NONFINITE, "non-finite Hessenberg column (cycle %d, Arnoldi step %d, ||w|| %g, h_0 %g)",
                      j, i, hnew, K->hcol[0]);
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
                      "GMRES least-squares solution non-finite (cycle %d, %d Arnoldi steps, min |R_kk| %g, ||b|| %g)", j,
                      i, dmin, nrm2b);
        if (i > 0) {
            CPG_HIP(hipMemcpyAsync(K->ydev.p, y.data(), i * sizeof(double), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_update, dim3(K->nblk), dim3(KWG), 0, s, V, N, i, nullptr, K->ydev.p, x);
            CPG_LAUNCH_CHECK();
        }
        CPG_HIP(hipStreamSynchronize(s));
        if (std::fabs(y[i]) < tl * nrm2b) break;
    }
    *maxit = m * j + i;
    *tol = std::fabs(y[i]) / nrm2b;
    K->harvest();
    prec_flow_check(P);
    CPG_CHECK(std::isfinite(*tol), CPG_ERR_NONFINITE, "non-finite GMRES residual estimate");
}

// ---- flexible driver: everything on the device, cycles replayed from graphs --------------
namespace {

// r = A x - b (or -b when x is zero); beta = ||r||; V_0 = r / beta; g = beta e1
void fg_start(cpg_krylov_s &K, KOps &op, const double *b, double *x, bool x_zero) {
    const size_t N = K.N;
    hipStream_t s = op.s;
    if (x_zero) {
        vec_fill(x, 0.0, N, s);
        hipLaunchKernelGGL(k_neg, dim3(ceil_div(N, 256)), dim3(256), 0, s, K.r.p, b, N);
    } else {
        op.matvec(x, K.r.p);
        hipLaunchKernelGGL(k_sub, dim3(ceil_div(N, 256)), dim3(256), 0, s, K.r.p, b, N);
    }
    CPG_LAUNCH_CHECK();
    op.norm_dev(K.r.p, K.scal.p + 1);
    hipLaunchKernelGGL(k_fg_cycle_init, dim3(1), dim3(1), 0, s, K.fstate());
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scale_dev, dim3(ceil_div(N, 256)), dim3(256), 0, s, K.r.p, K.V.p, N, K.scal.p + 1, nullptr);
    CPG_LAUNCH_CHECK();
}

// Arnoldi steps i0 .. i1-1: z_i = M^-1 v_i ; w = A z_i ; CGS ; v_{i+1} = w / ||w|| ; Givens
void fg_steps(cpg_krylov_s &K, KOps &op, int i0, int i1) {
    const size_t N = K.N;
    hipStream_t s = op.s;
    const int *act = K.fflags.p;
    double *V = K.V.p;
    for (int i = i0; i < i1; i++) {
        double *vi = V + N * i, *vi1 = V + N * (i + 1), *zi = K.Z.p + N * i;
        op.pc_to(vi, zi, act);
        op.matvec(zi, vi1, act);
        hipLaunchKernelGGL(k_mdot, dim3(K.nblk), dim3(KWG), 0, s, V, N, i + 1, vi1, K.partial.p, act);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_reduce, dim3(i + 1), dim3(KWG), 0, s, K.partial.p, K.nblk, i + 1, K.hdev.p, 0, act);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_orth_norm, dim3(K.nblk), dim3(KWG), 0, s, V, N, i + 1, K.hdev.p, vi1, K.partial.p, act);
        CPG_LAUNCH_CHECK();
        // ||w||, v_{i+1} = w / ||w||, Givens, and the active flag into the host's look-ahead slot
        hipLaunchKernelGGL(k_fg_norm_step, dim3(K.nblk), dim3(KWG), 0, s, K.partial.p, K.nblk, vi1, N, i, K.fstate(),
                           K.d_flags ? K.d_flags + i / KBLK : nullptr);
        CPG_LAUNCH_CHECK();
    }
}

// y = R^-1 g ; x -= Z y
void fg_finish(cpg_krylov_s &K, KOps &op, double *x) {
    hipLaunchKernelGGL(k_fg_solve, dim3(1), dim3(64), 0, op.s, K.fstate());
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_update, dim3(K.nblk), dim3(KWG), 0, op.s, K.Z.p, K.N, 0, K.fflags.p + 1, K.fy.p, x);
    CPG_LAUNCH_CHECK();
}

// run `body` on s from a graph captured at its first use (eagerly when timing)
template <class Fn>
void run_part(cpg_krylov_s &K, hipStream_t s, hipGraphExec_t &exec, Fn &&body) {
    if (K.timing || !K.use_graph) {
        body();
        return;
    }
    if (!exec) {
        CPG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        CPG_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CPG_HIP(e);
    }
    CPG_HIP(hipGraphLaunch(exec, s));
}

}  // namespace

void gmres_flexible(cpg_krylov_s *K, const LinOp &A, cpg_prec P, const PcOp *M, const double *b, double *x,
                    bool x_zero, double *tol, int *maxit, hipStream_t caller) {
    // everything below runs on K->kst after the caller's queued work; the host waits for it
    // before returning, so later work on the caller's stream is ordered after it
    CPG_HIP(hipEventRecord(K->cev, caller));
    CPG_HIP(hipStreamWaitEvent(K->kst, K->cev, 0));
    hipStream_t s = K->kst;
    KOps op{*K, A, P, s, M};
    const int m = K->m;
    // the graphs hold device pointers of the operator and the preconditioner: they are replayed
    // only for the same handles with the same buffers
    const cpg_krylov_s::Sig sig{b, x, reinterpret_cast<const void *>(A.fn), A.ctx,
                                M ? reinterpret_cast<const void *>(M->ctx) : P,
                                A.fn == bsr_linop ? static_cast<const cpg_bsr_s *>(A.ctx)->gen : 0,
                                M ? 0 : prec_generation(P)};
    if (!(sig == K->sig)) {
        K->drop_graphs();
        K->sig = sig;
    }
    // ||b|| and the absolute tolerance (eager, read once)
    op.norm_dev(b, K->scal.p);
    hipLaunchKernelGGL(k_fg_setup, dim3(1), dim3(1), 0, s, K->fstate(), *tol);
    CPG_LAUNCH_CHECK();
    double nb = 0.0;
    CPG_HIP(hipMemcpyAsync(&nb, K->scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    CPG_HIP(hipStreamSynchronize(s));
    CPG_CHECK(std::isfinite(nb), CPG_ERR_NONFINITE, "non-finite right-hand side norm");
    if (nb == 0.0) {
        vec_fill(x, 0.0, K->N, s);
        CPG_HIP(hipStreamSynchronize(s));
        *tol = 0.0;
        *maxit = 0;
        K->harvest();
        return;
    }
    const int maxcyc = std::max(1, (*maxit + m - 1) / m);
    const int nblk = (m + KBLK - 1) / KBLK;
    int *hf = K->h_flags;   // [0..nblk) per-block active flags, [nblk..nblk+4) final flags
    bool conv = false;
    for (int cyc = 0; cyc < maxcyc && !conv; cyc++) {
        const bool zero = x_zero && cyc == 0;
        run_part(*K, s, zero ? K->g_start0 : K->g_start, [&] { fg_start(*K, op, b, x, zero); });
        for (int bk = 0; bk < nblk; bk++) {
            const int i0 = bk * KBLK, i1 = std::min(m, i0 + KBLK);
            run_part(*K, s, K->g_blocks[bk], [&] { fg_steps(*K, op, i0, i1); });
            CPG_HIP(hipEventRecord(K->fev[bk], s));
            // look-ahead: block bk is queued; stop if block bk-1 already ended the cycle
            if (bk > 0) {
                CPG_HIP(hipEventSynchronize(K->fev[bk - 1]));
                if (!hf[bk - 1]) break;
            }
        }
        run_part(*K, s, K->g_finish, [&] { fg_finish(*K, op, x); });
        CPG_HIP(hipMemcpyAsync(hf + nblk, K->fflags.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        CPG_HIP(hipStreamSynchronize(s));
        conv = hf[nblk + 2] != 0;
    }
    double resid = 0.0;
    CPG_HIP(hipMemcpyAsync(&resid, K->scal.p + 3, sizeof(double), hipMemcpyDeviceToHost, s));
    CPG_HIP(hipStreamSynchronize(s));
    *maxit = hf[nblk + 3];
    *tol = resid / nb;
    K->harvest();
    prec_flow_check(P);
    CPG_CHECK(std::isfinite(*tol), CPG_ERR_NONFINITE,
              "non-finite GMRES residual estimate (%d Arnoldi steps, ||b|| %g)", *maxit, nb);
}

// solve A x = b with the driver of K's mode; x holds the initial guess unless x_zero
void gmres_solve(cpg_krylov K, const LinOp &A, cpg_prec P, const double *b, double *x, bool x_zero, double *tol,
                 int *maxit, hipStream_t s) {
    if (K->mode == CPG_GMRES_FLEXIBLE) {
        gmres_flexible(K, A, P, nullptr, b, x, x_zero, tol, maxit, s);
        return;
    }
    if (x_zero) vec_fill(x, 0.0, K->N, s);
    gmres_left_ref(K, A, P, b, x, tol, maxit, s);
}

void gmres_flexible_op(cpg_krylov K, const LinOp &A, const PcOp &M, const double *b, double *x, bool x_zero,
                       double *tol, int *maxit, hipStream_t s) {
    CPG_CHECK(K->mode == CPG_GMRES_FLEXIBLE, CPG_ERR_ARG, "an operator preconditioner needs flexible GMRES");
    gmres_flexible(K, A, nullptr, &M, b, x, x_zero, tol, maxit, s);
}

// partial[b] = sum over chunk b of (t - b)^2
__global__ __launch_bounds__(KWG) void k_diff_sq(const double *__restrict__ t, const double *__restrict__ b,
                                                 size_t N, double *__restrict__ partial) {
    __shared__ double red[KWG / 64];
    const size_t base = (size_t)blockIdx.x * KCH + threadIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < KPT; k++) {
        const size_t e = base + (size_t)k * KWG;
        if (e < N) {
            const double d = t[e] - b[e];
            acc = fma(d, d, acc);
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = 0.0;
        for (int w = 0; w < KWG / 64; w++) q += red[w];
        partial[blockIdx.x] = q;
    }
}

// x = (omega Db + x) - omega t: the Jacobi update of BlockMatrix::jacobi (src/BlockMatrix.cpp:73)
__global__ void k_jacobi_update(double *__restrict__ x, const double *__restrict__ Db, const double *__restrict__ t,
                                size_t N, double omega) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) x[e] = (omega * Db[e] + x[e]) - omega * t[e];
}

void bsr_linop(void *ctx, const double *x, double *y, const int *active, hipStream_t s) {
    bsr_matvec(*static_cast<const cpg_bsr_s *>(ctx), x, y, s, active);
}

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
            if (const char *g = std::getenv("CPOLYG_GMRES_GRAPH")) K->use_graph = !(g[0] == '0');
            K->N = (size_t)n_rows * bl;
            K->nblk = std::max(1, ceil_div(K->N, KCH));
            K->V.alloc(K->N * (restart + 1));
            if (mode == CPG_GMRES_FLEXIBLE) K->Z.alloc(K->N * restart);
            K->r.alloc(K->N);
            K->partial.alloc((size_t)K->nblk * (restart + 1));
            K->hdev.alloc(restart + 2);
            K->ydev.alloc(restart + 1);
            K->scal.alloc(4);
            K->h.assign((size_t)restart * (restart + 1) / 2 + 1, 0.0);
            K->y.assign(restart + 1, 0.0);
            K->cs.assign(restart, 0.0);
            K->sn.assign(restart, 0.0);
            K->hcol.assign(restart + 2, 0.0);
            if (mode == CPG_GMRES_FLEXIBLE) {
                const int nb = (restart + KBLK - 1) / KBLK;
                K->fH.alloc((size_t)restart * (restart + 1));
                K->fcs.alloc(restart);
                K->fsn.alloc(restart);
                K->fg.alloc(restart + 1);
                K->fy.alloc(restart);
                K->fflags.alloc(4);
                CPG_HIP(hipHostMalloc((void **)&K->h_flags, sizeof(int) * (nb + 4), hipHostMallocMapped));
                std::memset(K->h_flags, 0, sizeof(int) * (nb + 4));
                CPG_HIP(hipHostGetDevicePointer((void **)&K->d_flags, K->h_flags, 0));
                K->fev.assign(nb, nullptr);
                for (auto &e : K->fev) CPG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                K->g_blocks.assign(nb, nullptr);
                CPG_HIP(hipStreamCreateWithFlags(&K->kst, hipStreamNonBlocking));
                CPG_HIP(hipEventCreateWithFlags(&K->cev, hipEventDisableTiming));
            }
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
        const LinOp op{bsr_linop, A};
        gmres_solve(K, op, P, b, x, false, tol, maxit, S(stream));
    });
}

int cpg_jacobi(cpg_krylov K, cpg_bsr A, cpg_prec P, const double *b, double *x, double *tol, int *maxit,
               void *stream) {
    return guard([&] {
        CPG_CHECK(K && A && b && x && tol && maxit, CPG_ERR_ARG, "null argument");
        CPG_CHECK(A->n == K->n && A->bl == K->bl, CPG_ERR_ARG, "matrix does not match the workspace");
        CPG_CHECK(*maxit >= 0, CPG_ERR_ARG, "negative iteration limit");
        hipStream_t s = S(stream);
        const size_t N = K->N;
        double *Db = K->V.p, *t = K->V.p + N;   // V holds restart + 1 >= 2 vectors
        const double omega = 1.0;
        // Dinvb = P^-1 b
        CPG_HIP(hipMemcpyAsync(Db, b, N * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (P) {
            const int st = cpg_prec_apply(P, Db, s);
            if (st) throw Error{st};
        }
        double rn = 0.0;   // the reference's norm of an empty residual when no iteration runs
        int it = 0;
        for (; it < *maxit; it++) {
            bsr_matvec(*A, x, t, s, nullptr);   // t = A x
            hipLaunchKernelGGL(k_diff_sq, dim3(K->nblk), dim3(KWG), 0, s, t, b, N, K->partial.p);
            CPG_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(KWG), 0, s, K->partial.p, K->nblk, 1, K->scal.p, 1, nullptr);
            CPG_LAUNCH_CHECK();
            CPG_HIP(hipMemcpyAsync(&rn, K->scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
            CPG_HIP(hipStreamSynchronize(s));
            if (rn < *tol) break;   // ||A x - b||_2, absolute
            if (P) {
                const int st = cpg_prec_apply(P, t, s);
                if (st) throw Error{st};
            }
            hipLaunchKernelGGL(k_jacobi_update, dim3(ceil_div(N, 256)), dim3(256), 0, s, x, Db, t, N, omega);
            CPG_LAUNCH_CHECK();
        }
        CPG_HIP(hipStreamSynchronize(s));
        *tol = rn;
        *maxit = it;
    });
}

}  // extern "C"

// timing accessors used by the stepper
namespace cpg {
DispatchTiming &dispatch_timing() {
    static thread_local DispatchTiming t;
    return t;
}

void krylov_timing(cpg_krylov K, bool on) {
    K->timing = on;
    K->n_matvec = K->n_apply = 0;
    K->ms_matvec = K->ms_apply = 0;
}
void krylov_timing_read(cpg_krylov K, int *nm, double *msm, int *na, double *msa) {
    *nm = K->n_matvec; *msm = K->ms_matvec; *na = K->n_apply; *msa = K->ms_apply;
}
}  // namespace cpg
