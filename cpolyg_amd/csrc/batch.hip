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
#include "rowstream.hpp"

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

// ---------------------------------------------------------------------------------------
// Persistent batched flexible GMRES: one cooperative launch solves every instance's Newton
// system, all restart cycles included.
//
// Instance q is solved by a team of nwg_q workgroups (wg0_q .. wg0_q + nwg_q - 1), sized by the
// instance's bytes per Arnoldi step times its Arnoldi count in the previous solve, so that the
// teams finish together.  A team runs the GMRES of its instance alone: the DILU sweeps take the
// instance's rows of each colour level (a team barrier after every level), the matvec its block
// rows, and the Gram-Schmidt passes its vector segment in fixed chunks of PCH elements.  The
// partial sums of a reduction are written per chunk; after the barrier every workgroup of the
// team adds them in chunk order and runs the Givens rotations and the least-squares update on
// its own LDS copy of the Hessenberg matrix, so every workgroup takes the same decisions without
// a broadcast, and the results do not depend on how many workgroups a team has.  Instances
// that converge leave the kernel while the others go on.
//
// Team barriers are an arrival counter and a generation word per instance (agent-scope
// acquire/release, so the writes of workgroups on other XCDs are visible after them).  The
// launch is cooperative: every workgroup is resident, so every barrier completes.  A barrier
// that waits longer than spin_max sleeps nevertheless gives up, sets the abort flag, which
// every other wait also checks, and every workgroup then leaves the kernel; the host reports
// the error.
namespace pgm {

constexpr int PT = 256;           // threads per workgroup
constexpr int PW = PT / 64;       // wavefronts per workgroup
constexpr int PEL = 4;            // elements per thread and chunk
constexpr int PCH = PT * PEL;     // elements per chunk

struct Task {
    int row, b0, b1;   // class-local block row, packed blocks [b0, b1)
};

struct Cls {
    const double *A;           // Newton matrix blocks (FP64, column-major)
    const int *rowptr, *colind;
    const void *F;             // factor stream in the batch's precision
    const float *scale;        // BF16 per-block scales
    const int *scol;           // factor-stream source tags; null for block Jacobi (block i of row i)
    long long off;             // offset of the class in the batch vector
    int bl;
};

struct Inst {
    int cls, wg0, nwg;
    int r0, r1;                // class-local block rows
    int e0, len;               // batch-vector segment [e0, e0 + len) (the batch vector is below 2^31)
    int ck0, nck;              // the segment's chunks in the chunk-indexed partial sums
    int ph0, nfw, nbw;         // forward phases, then backward phases, from ph0 in the phase table
    int tm0, tm1;              // tasks of the matvec (block rows with their BSR blocks)
};

struct Args {
    const Cls *cls;
    const Inst *inst;
    const int *wg_inst;        // instance of each workgroup
    const int2 *phase;         // tasks [x, y) of each phase
    const Task *tasks;
    const double *r;           // right-hand side
    double *x, *V, *Z, *w, *y; // solution, Krylov bases (m + 1 and m vectors of N), work vectors
    long long N;
    double *partial;           // 2 x nchunk x (m + 1)
    int nchunk;
    unsigned *bar;             // 2 per instance: arrivals, generation
    int *abort_flag;
    int *iters, *conv;         // per instance: Arnoldi steps, converged
    double *resid;             // per instance: residual estimate at exit
    unsigned long long *clk;   // per instance: the team's first workgroup's start and end (wall clock)
    int m, maxcyc, pc;         // pc: 0 none, 1 factor streams
    double rtol;
    unsigned spin_max;
};

// Team barrier; false when it gave up (abort).  Every thread of the workgroup calls it.
__device__ inline bool team_sync(unsigned *bar, int nwg, int *abort_flag, unsigned spin_max, int *okl) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        if (nwg > 1) {
            unsigned *cnt = bar, *gen = bar + 1;
            const unsigned g0 = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned arrived = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == (unsigned)nwg - 1u) {
                __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(gen, g0 + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned k = 0;
                while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                    if (++k >= spin_max || __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                if (!ok) __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        *okl = ok;
    }
    __syncthreads();
    return *okl != 0;
}

// out[j] = sum over the instance's chunks k = 0 .. nck-1 (in that order per lane, then a fixed
// butterfly) of P[(ck0 + k) * ld + j], j < nj; identical in every workgroup of the team
__device__ inline void team_sum(const double *P, int ck0, int nck, int ld, int nj, double *out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = wv; j < nj; j += PW) {
        double s = 0.0;
        for (int k = lane; k < nck; k += 64) s += P[(size_t)(ck0 + k) * ld + j];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) out[j] = s;
    }
    __syncthreads();
}

// the workgroup's sum of v over its threads, in a fixed order; valid in thread 0
__device__ inline double wg_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < PW; k++) s += red[k];
    __syncthreads();
    return s;
}

template <class T>
__device__ inline double ld_elem(const T *p) {
    return (double)*p;
}
template <>
__device__ inline double ld_elem<bf16>(const bf16 *p) {
    return (double)__uint_as_float((uint32_t)p->v << 16);
}

// sources of a packed block by its tag (stream_source): tag >= 0 c1 * s1[tag], -1 - u c0 * s0[u],
// -1 - (kSrc2Tag + k) c2 * s2[k]
struct Src {
    const double *s0, *s1, *s2;
    double c0, c1, c2;
};

__device__ inline const double *src_of(int tag, int bl, const double *s0, const double *s1, const double *s2,
                                      double c0, double c1, double c2, double &cf) {
    const int u = -1 - tag;
    const double *base = s0;
    int row = u;
    cf = c0;
    if (tag >= 0) {
        base = s1;
        row = tag;
        cf = c1;
    } else if (u >= kSrc2Tag) {
        base = s2;
        row = u - kSrc2Tag;
        cf = c2;
    }
    return base + (size_t)row * bl;
}

// out_row = (base_row +) sum over the row's blocks of coef * block * source, for the tasks [t0, t1)
// of a phase (task t: row tasks[t].row from the blocks [b0, b1) of F, block b's source by tags[b],
// or the row itself from S.s1 when tags is null).  A wavefront takes R = 64 / bl rows at a time,
// lane (row slot, r) forms output entry r: the lanes of a slot read a column of a block together
// (bl consecutive elements) and the column's source value (one address), eight columns per round
// with two partial sums, and the next block's source tag is loaded before the current block is
// used.  Several rows per wavefront keep more loads in flight than one row per wavefront with
// wider loads, which matters more here: the team shares the GPU with the other teams and holds
// few wavefronts per CU.  One routine for every block size keeps the kernel's register count that
// of its vector passes.
template <class T>
__device__ void phase_rows(int bl, const T *F, const float *scale, const int *tags, const Task *tasks, int t0, int t1,
                           const Src S, const double *base, double *out, int gw, int nwave) {
    const double *s0 = S.s0, *s1 = S.s1, *s2 = S.s2;
    const double c0 = S.c0, c1 = S.c1, c2 = S.c2;
    const int R = 64 / bl;
    const int lane = threadIdx.x & 63;
    const int rs = lane / bl, r = lane - rs * bl;
    if (rs >= R) return;
    const size_t bb = (size_t)bl * bl;
    const int bl8 = bl & ~7, bl4 = bl & ~3;
    for (int tb = t0 + gw * R; tb < t1; tb += nwave * R) {
        const int t = tb + rs;
        if (t >= t1) break;
        const Task k = tasks[t];
        double acc = 0.0;
        int tag = tags ? tags[k.b0] : k.row;
        for (int b = k.b0; b < k.b1; b++) {
            const int tagn = (tags && b + 1 < k.b1) ? tags[b + 1] : k.row;
            double cf;
            const double *src = src_of(tag, bl, s0, s1, s2, c0, c1, c2, cf);
            const T *Fb = F + (size_t)b * bb + r;
            double p0 = 0.0, p1 = 0.0;
            int c = 0;
            for (; c < bl8; c += 8) {
                double f[8];
#pragma unroll
                for (int u = 0; u < 8; u++) f[u] = ld_elem(Fb + (size_t)(c + u) * bl);
#pragma unroll
                for (int u = 0; u < 8; u += 2) {
                    p0 = fma(f[u], src[c + u], p0);
                    p1 = fma(f[u + 1], src[c + u + 1], p1);
                }
            }
            for (; c < bl4; c += 4) {
                const double f0 = ld_elem(Fb + (size_t)c * bl), f1 = ld_elem(Fb + (size_t)(c + 1) * bl);
                const double f2 = ld_elem(Fb + (size_t)(c + 2) * bl), f3 = ld_elem(Fb + (size_t)(c + 3) * bl);
                p0 = fma(f0, src[c], p0);
                p1 = fma(f1, src[c + 1], p1);
                p0 = fma(f2, src[c + 2], p0);
                p1 = fma(f3, src[c + 3], p1);
            }
            for (; c < bl; c++) p0 = fma(ld_elem(Fb + (size_t)c * bl), src[c], p0);
            double part = p0 + p1;
            if constexpr (sizeof(T) == 2) part *= (double)scale[b];
            acc = fma(cf, part, acc);
            tag = tagn;
        }
        out[(size_t)k.row * bl + r] = base ? base[(size_t)k.row * bl + r] + acc : acc;
    }
}

// what a workgroup knows of its team
struct Team {
    Inst I;
    Cls C;
    int q, wl, gw, nwave, ld;
    unsigned *bar;
    int *okl;
};

__device__ inline bool tsync(const Args &a, const Team &t) {
    return team_sync(t.bar, t.I.nwg, a.abort_flag, a.spin_max, t.okl);
}

// the workgroup's chunk k of the instance: [c0, c1) of the batch vector
__device__ inline void chunk_range(const Team &t, int k, int &c0, int &c1) {
    c0 = t.I.e0 + k * PCH;
    const int end = t.I.e0 + t.I.len;
    c1 = c0 + PCH < end ? c0 + PCH : end;
}

// w = A v (v, w batch vectors), then a team barrier
__device__ bool t_matvec(const Args &a, const Team &t, const double *v) {
    phase_rows<double>(t.C.bl, t.C.A, nullptr, t.C.colind, a.tasks, t.I.tm0, t.I.tm1,
                       Src{nullptr, v + t.C.off, nullptr, 0.0, 1.0, 0.0}, nullptr, a.w + t.C.off, t.gw, t.nwave);
    return tsync(a, t);
}

// z = s M^-1 w: the forward sweep on w, the backward sweep with s on its D^-1 sources, so that
// z = M^-1 (s w); a team barrier after every level
template <class T>
__device__ bool t_precondition(const Args &a, const Team &t, double s, double *z) {
    const int tid = threadIdx.x;
    if (a.pc == 0) {
        for (int k = t.wl; k < t.I.nck; k += t.I.nwg) {
            int c0, c1;
            chunk_range(t, k, c0, c1);
            for (int e = c0 + tid; e < c1; e += PT) z[e] = s * a.w[e];
        }
        return tsync(a, t);
    }
    const T *F = reinterpret_cast<const T *>(t.C.F);
    double *wc = a.w + t.C.off, *yc = a.y + t.C.off, *zc = z + t.C.off;
    bool ok = true;
    for (int p = 0; ok && p < t.I.nfw; p++) {   // y_i = w_i - sum L_ik y_k (sources without lower blocks: w)
        const int2 ph = a.phase[t.I.ph0 + p];
        phase_rows<T>(t.C.bl, F, t.C.scale, t.C.scol, a.tasks, ph.x, ph.y, Src{nullptr, yc, wc, 0.0, -1.0, -1.0}, wc,
                      yc, t.gw, t.nwave);
        ok = tsync(a, t);
    }
    const Src Sb = t.C.scol ? Src{yc, zc, wc, s, -1.0, s} : Src{nullptr, wc, nullptr, 0.0, s, 0.0};
    for (int p = 0; ok && p < t.I.nbw; p++) {   // z_i = s D_i^-1 y_i - sum DU_ij z_j (block Jacobi: s D_i^-1 w_i)
        const int2 ph = a.phase[t.I.ph0 + t.I.nfw + p];
        phase_rows<T>(t.C.bl, F, t.C.scale, t.C.scol, a.tasks, ph.x, ph.y, Sb, nullptr, zc, t.gw, t.nwave);
        ok = tsync(a, t);
    }
    return ok;
}

// P[chunk] = sum over the chunk of w^2 for the workgroup's chunks; with sub, w = r - w first
__device__ void t_norm_parts(const Args &a, const Team &t, bool sub, double *P, double *red) {
    const int tid = threadIdx.x;
    for (int k = t.wl; k < t.I.nck; k += t.I.nwg) {
        int c0, c1;
        chunk_range(t, k, c0, c1);
        double s = 0.0;
        for (int e = c0 + tid; e < c1; e += PT) {
            const double v = sub ? a.r[e] - a.w[e] : a.w[e];
            if (sub) a.w[e] = v;
            s = fma(v, v, s);
        }
        s = wg_sum(s, red);
        if (tid == 0) P[(size_t)(t.I.ck0 + k) * t.ld] = s;
    }
}

// P[chunk][j] = V_j . w over the workgroup's chunks, j <= i
__device__ void t_mdot(const Args &a, const Team &t, int i, double *P, double *red) {
    const int tid = threadIdx.x;
    const int ld = t.ld;
    for (int k = t.wl; k < t.I.nck; k += t.I.nwg) {
        int c0, c1;
        chunk_range(t, k, c0, c1);
        double wv[PEL];
#pragma unroll
        for (int u = 0; u < PEL; u++) {
            const int e = c0 + tid + u * PT;
            wv[u] = e < c1 ? a.w[e] : 0.0;
        }
        for (int j = 0; j <= i; j++) {
            const double *v = a.V + (size_t)j * a.N;
            double sj = 0.0;
#pragma unroll
            for (int u = 0; u < PEL; u++) {
                const int e = c0 + tid + u * PT;
                if (e < c1) sj = fma(v[e], wv[u], sj);
            }
            for (int off = 32; off > 0; off >>= 1) sj += __shfl_xor(sj, off);
            if ((tid & 63) == 0) red[(tid >> 6) * ld + j] = sj;
        }
        __syncthreads();
        for (int j = tid; j <= i; j += PT) {
            double sj = 0.0;
            for (int v2 = 0; v2 < PW; v2++) sj += red[v2 * ld + j];
            P[(size_t)(t.I.ck0 + k) * ld + j] = sj;
        }
        __syncthreads();
    }
}

// w -= sum_{j<=i} h_j V_j over the workgroup's chunks; P[chunk] = sum of the new w^2
__device__ void t_orth(const Args &a, const Team &t, int i, const double *h, double *P, double *red) {
    const int tid = threadIdx.x;
    for (int k = t.wl; k < t.I.nck; k += t.I.nwg) {
        int c0, c1;
        chunk_range(t, k, c0, c1);
        double wv[PEL];
#pragma unroll
        for (int u = 0; u < PEL; u++) {
            const int e = c0 + tid + u * PT;
            wv[u] = e < c1 ? a.w[e] : 0.0;
        }
        for (int j = 0; j <= i; j++) {
            const double hj = h[j];
            const double *v = a.V + (size_t)j * a.N;
#pragma unroll
            for (int u = 0; u < PEL; u++) {
                const int e = c0 + tid + u * PT;
                if (e < c1) wv[u] = fma(-hj, v[e], wv[u]);
            }
        }
        double sq = 0.0;
#pragma unroll
        for (int u = 0; u < PEL; u++) {
            const int e = c0 + tid + u * PT;
            if (e < c1) {
                a.w[e] = wv[u];
                sq = fma(wv[u], wv[u], sq);
            }
        }
        sq = wg_sum(sq, red);
        if (tid == 0) P[(size_t)(t.I.ck0 + k) * t.ld] = sq;
    }
}

// dst = f * w over the workgroup's chunks
__device__ void t_scale_w(const Args &a, const Team &t, double f, double *dst) {
    for (int k = t.wl; k < t.I.nck; k += t.I.nwg) {
        int c0, c1;
        chunk_range(t, k, c0, c1);
        for (int e = c0 + (int)threadIdx.x; e < c1; e += PT) dst[e] = a.w[e] * f;
    }
}

template <class T>
__global__ __launch_bounds__(PT, 5) void k_batch_fgmres(Args a) {
    extern __shared__ double sm[];
    __shared__ int okl, flag;
    const int m = a.m;
    double *H = sm;                          // column i at H + i * (m + 1)
    double *cs = H + (size_t)m * (m + 1);
    double *sn = cs + m;
    double *g = sn + m;                      // m + 1
    double *hv = g + m + 1;                  // m + 2: the current column, reduction results
    double *yv = hv + m + 2;                 // m
    double *red = yv + m;                    // PW x (m + 1)
    const int tid = threadIdx.x;
    Team t;
    t.q = a.wg_inst[blockIdx.x];
    t.I = a.inst[t.q];
    t.C = a.cls[t.I.cls];
    t.wl = blockIdx.x - t.I.wg0;
    t.gw = t.wl * PW + (tid >> 6);
    t.nwave = t.I.nwg * PW;
    t.ld = m + 1;
    t.bar = a.bar + 2 * t.q;
    t.okl = &okl;
    const int ld = t.ld;
    if (t.wl == 0 && tid == 0) a.clk[2 * t.q] = wall_clock64();
    double *P0 = a.partial, *P1 = a.partial + (size_t)a.nchunk * ld;
    int rb = 0;   // reductions so far: the parity picks the partial-sum buffer

    // x = 0, w = r, ||r||^2
    for (int k = t.wl; k < t.I.nck; k += t.I.nwg) {
        int c0, c1;
        chunk_range(t, k, c0, c1);
        for (int e = c0 + tid; e < c1; e += PT) {
            a.x[e] = 0.0;
            a.w[e] = a.r[e];
        }
    }
    t_norm_parts(a, t, false, P0, red);
    bool ok = tsync(a, t);
    rb++;
    team_sum(P0, t.I.ck0, t.I.nck, ld, 1, hv);
    const double bnorm = sqrt(hv[0]);
    const double thr = a.rtol * bnorm;
    double beta = bnorm, res = bnorm;
    int its = 0, conv = 0;
    for (int cyc = 0; ok && cyc < a.maxcyc; cyc++) {
        if (cyc > 0) {   // w = r - A x
            ok = t_matvec(a, t, a.x);
            if (!ok) break;
            double *P = (rb & 1) ? P1 : P0;
            t_norm_parts(a, t, true, P, red);
            ok = tsync(a, t);
            rb++;
            if (!ok) break;
            team_sum(P, t.I.ck0, t.I.nck, ld, 1, hv);
            beta = sqrt(hv[0]);
        }
        res = beta;
        if (!(beta >= thr && beta > 0.0)) {   // converged, zero, or non-finite (reported by res)
            conv = 1;
            break;
        }
        t_scale_w(a, t, 1.0 / beta, a.V);
        if (tid == 0) {
            for (int r = 0; r <= m; r++) g[r] = 0.0;
            g[0] = beta;
        }
        int kd = 0;
        double s = 1.0 / beta;
        for (int i = 0; i < m; i++) {
            double *zi = a.Z + (size_t)i * a.N;
            ok = t_precondition<T>(a, t, s, zi);
            if (ok) ok = t_matvec(a, t, zi);
            if (!ok) break;
            // h_j = V_j . w  (j <= i)
            double *P = (rb & 1) ? P1 : P0;
            t_mdot(a, t, i, P, red);
            ok = tsync(a, t);
            rb++;
            if (!ok) break;
            team_sum(P, t.I.ck0, t.I.nck, ld, i + 1, hv);
            // w -= sum_j h_j V_j ; ||w||
            P = (rb & 1) ? P1 : P0;
            t_orth(a, t, i, hv, P, red);
            ok = tsync(a, t);
            rb++;
            if (!ok) break;
            team_sum(P, t.I.ck0, t.I.nck, ld, 1, hv + i + 1);
            const double nrm = sqrt(hv[i + 1]);
            t_scale_w(a, t, nrm > 0.0 ? 1.0 / nrm : 0.0, a.V + (size_t)(i + 1) * a.N);
            if (tid == 0) {
                // rotations as src/BlockMatrix.cpp:137-145: c = rd/dd, s = -hnew/dd
                double *h = H + (size_t)i * ld;
                for (int r = 0; r <= i; r++) h[r] = hv[r];
                for (int k = 0; k < i; k++) {
                    const double tk = cs[k] * h[k] - sn[k] * h[k + 1];
                    h[k + 1] = sn[k] * h[k] + cs[k] * h[k + 1];
                    h[k] = tk;
                }
                const double rd = h[i];
                const double dd = sqrt(rd * rd + nrm * nrm);
                cs[i] = dd > 0.0 ? rd / dd : 1.0;
                sn[i] = dd > 0.0 ? -nrm / dd : 0.0;
                h[i] = dd;
                g[i + 1] = sn[i] * g[i];
                g[i] = cs[i] * g[i];
                flag = !(fabs(g[i + 1]) >= thr) ? 1 : (i + 1 == m ? 2 : 0);
            }
            __syncthreads();
            kd = i + 1;
            const int f = flag;
            if (f) {
                conv = f == 1;
                break;
            }
            s = nrm > 0.0 ? 1.0 / nrm : 0.0;
        }
        if (!ok) break;
        // y = R^-1 g ; x += Z y
        if (tid == 0) {
            for (int r = 0; r < kd; r++) yv[r] = g[r];
            for (int c = kd - 1; c >= 0; c--) {
                const double *h = H + (size_t)c * ld;
                yv[c] /= h[c];
                for (int r = 0; r < c; r++) yv[r] -= h[r] * yv[c];
            }
        }
        __syncthreads();
        for (int k = t.wl; k < t.I.nck; k += t.I.nwg) {
            int c0, c1;
            chunk_range(t, k, c0, c1);
            for (int e = c0 + tid; e < c1; e += PT) {
                double xv = a.x[e];
                for (int j = 0; j < kd; j++) xv = fma(yv[j], a.Z[(size_t)j * a.N + e], xv);
                a.x[e] = xv;
            }
        }
        its += kd;
        res = fabs(g[kd]);
        if (conv) break;
        ok = tsync(a, t);   // the next cycle's matvec reads x of every chunk
    }
    if (ok && t.wl == 0 && tid == 0) {
        a.iters[t.q] = its;
        a.resid[t.q] = res;
        a.conv[t.q] = conv;
        a.clk[2 * t.q + 1] = wall_clock64();
    }
}

}  // namespace pgm

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
    int mode = CPG_BATCH_GRAPH;
    bool use_graph = true;
    int kblock = 5;                         // Arnoldi steps per captured block
    // persistent mode (CPG_BATCH_PERSISTENT): the cooperative solver's tables
    struct Persist {
        DevBuf<pgm::Cls> cls;
        DevBuf<pgm::Inst> inst;
        DevBuf<int> wg_inst;
        DevBuf<int2> phase;
        DevBuf<pgm::Task> tasks;
        DevBuf<double> y, partial;
        DevBuf<unsigned> bar;
        DevBuf<int> abort_flag;
        std::vector<pgm::Inst> h_inst;      // teams as last uploaded
        std::vector<double> bytes;          // bytes per Arnoldi step of each instance
        std::vector<int> cap;               // most workgroups an instance can use
        std::vector<int> last_iters;        // Arnoldi steps of each instance in the last solve
        // the same per solve slot (stage, Newton iteration): a step's solves differ from one another
        // more than from the same solve of the previous step
        std::vector<std::vector<int>> slot_iters;
        // measured cost of each instance: team time x team size per Arnoldi step (0: not yet)
        std::vector<double> cost;
        DevBuf<unsigned long long> clk;
        std::vector<unsigned long long> h_clk;
        int nchunk = 0, grid = 0, occupancy = 0, cus = 0;
        double clock_khz = 1.0;             // device wall clock
        int wg_per_cu = 0;                  // 0: autotune over the first solves
        int max_team = 0;                   // most workgroups per team (0: no limit but cap)
        unsigned spin_max = 1u << 22;
        size_t lds = 0;
        // autotuner: seconds per (instance bytes x Arnoldi steps) of each candidate
        std::vector<int> cand;
        std::vector<double> cand_cost;
        int trial = 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
    } P;
    hipGraphExec_t g_start = nullptr, g_finish = nullptr;
    std::vector<hipGraphExec_t> g_steps;

    Chunks chunks() const { return Chunks{ck0.p, ck1.p, ckinst.p}; }
    BState state() {
        return BState{H.p, hcol.p, cs.p, sn.p, g.p, beta.p, thr.p, active.p, kdim.p, conv.p, iters.p, resid.p, m};
    }
    ~cpg_batch_s() {
        if (P.e0) (void)hipEventDestroy(P.e0);
        if (P.e1) (void)hipEventDestroy(P.e1);
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

// ---- persistent mode -------------------------------------------------------------------

size_t persist_lds(int m) {
    return (size_t)((size_t)m * (m + 1) + 2 * m + (m + 1) + (m + 2) + m + pgm::PW * (m + 1)) * sizeof(double);
}

template <class T>
const void *persist_kernel() {
    return reinterpret_cast<const void *>(pgm::k_batch_fgmres<T>);
}

const void *persist_kernel_of(int precision) {
    return precision == CPG_FP64 ? persist_kernel<double>() : precision == CPG_FP32 ? persist_kernel<float>()
                                                                                    : persist_kernel<bf16>();
}

// tables of the persistent solver: classes, instances with their chunks and sweep phases
void persist_setup(cpg_batch_s &B) {
    auto &P = B.P;
    std::vector<pgm::Cls> cls;
    std::vector<pgm::Inst> inst;
    std::vector<int2> phase;
    std::vector<pgm::Task> tasks;
    int ck = 0;
    const size_t elem = B.sp.precision == CPG_FP64 ? 8 : B.sp.precision == CPG_FP32 ? 4 : 2;
    for (size_t c = 0; c < B.cls.size(); c++) {
        auto &C = B.cls[c];
        const PrecTables t = prec_tables(C.P);
        CPG_CHECK(!t.diag32, CPG_ERR_UNSUPPORTED, "persistent batched solves: no BF16 factors with FP32 inverses");
        CPG_CHECK(C.bl <= 64, CPG_ERR_UNSUPPORTED, "persistent batched solves: block size %d > 64", C.bl);
        const bool dilu = t.kind == CPG_PC_DILU;
        pgm::Cls k{};
        k.A = C.J->blocks.p;
        k.rowptr = C.J->rowptr.p;
        k.colind = C.J->colind.p;
        k.F = t.F;
        k.scale = t.scale;
        k.scol = dilu ? t.scol : nullptr;
        k.off = (long long)C.off;
        k.bl = C.bl;
        cls.push_back(k);
        const auto &rp = C.J->h_rowptr;
        // instance of every row, then the instance's rows of each sweep level in sweep order
        std::vector<int> iof(C.n);
        for (int q = 0; q < C.ninst; q++)
            for (int i = C.cptr[q]; i < C.cptr[q + 1]; i++) iof[i] = q;
        std::vector<std::vector<std::vector<pgm::Task>>> ph(C.ninst);   // [instance][phase] tasks
        std::vector<int> nfw(C.ninst, 0);
        auto levels = [&](const std::vector<int> &lev, const std::vector<int> &rows, const std::vector<int> &ptr,
                          size_t l0) {
            for (size_t l = l0; l + 1 < lev.size(); l++) {
                std::vector<std::vector<pgm::Task>> cur(C.ninst);
                for (int p = lev[l]; p < lev[l + 1]; p++) {
                    const int i = rows[p];
                    cur[iof[i]].push_back(pgm::Task{i, ptr[p], ptr[p + 1]});
                }
                for (int q = 0; q < C.ninst; q++)
                    if (!cur[q].empty()) ph[q].push_back(std::move(cur[q]));
            }
        };
        if (dilu) {
            // forward level 0 is y = w (its rows have no lower blocks; their dependants read w)
            levels(*t.flev, *t.frows, *t.fptr, 1);
            for (int q = 0; q < C.ninst; q++) nfw[q] = (int)ph[q].size();
            levels(*t.blev, *t.brows, *t.bptr, 0);
        } else if (t.kind == CPG_PC_BJACOBI) {
            for (int q = 0; q < C.ninst; q++) {
                std::vector<pgm::Task> v;
                for (int i = C.cptr[q]; i < C.cptr[q + 1]; i++) v.push_back(pgm::Task{i, i, i + 1});
                ph[q].push_back(std::move(v));
            }
        }
        for (int q = 0; q < C.ninst; q++) {
            pgm::Inst I{};
            I.cls = (int)c;
            I.r0 = C.cptr[q];
            I.r1 = C.cptr[q + 1];
            I.e0 = (int)C.off + I.r0 * C.bl;
            I.len = (I.r1 - I.r0) * C.bl;
            I.ck0 = ck;
            I.nck = ceil_div(I.len, pgm::PCH);
            ck += I.nck;
            I.tm0 = (int)tasks.size();
            for (int i = I.r0; i < I.r1; i++) tasks.push_back(pgm::Task{i, rp[i], rp[i + 1]});
            I.tm1 = (int)tasks.size();
            I.ph0 = (int)phase.size();
            I.nfw = nfw[q];
            I.nbw = (int)ph[q].size() - nfw[q];
            for (auto &v : ph[q]) {
                phase.push_back(int2{(int)tasks.size(), (int)(tasks.size() + v.size())});
                tasks.insert(tasks.end(), v.begin(), v.end());
            }
            inst.push_back(I);
            const double nnzb = rp[I.r1] - rp[I.r0];
            const double bb = (double)C.bl * C.bl;
            const double fblocks = dilu ? nnzb : t.kind == CPG_PC_BJACOBI ? (double)(I.r1 - I.r0) : 0.0;
            P.bytes.push_back(nnzb * bb * 8.0 + fblocks * bb * elem + (double)I.len * 8.0 * (B.m / 2.0 + 8.0));
            P.cap.push_back(std::max(1, ceil_div(I.r1 - I.r0, pgm::PW * (64 / C.bl))));
        }
    }
    CPG_CHECK(tasks.size() < (size_t)1 << 31, CPG_ERR_UNSUPPORTED, "too many sweep tasks");
    P.nchunk = ck;
    P.cls.upload(cls);
    P.h_inst = inst;
    P.inst.upload(inst);
    if (!phase.empty()) P.phase.upload(phase);
    if (!tasks.empty()) P.tasks.upload(tasks);
    P.y.alloc(B.N);
    P.partial.alloc((size_t)2 * ck * (B.m + 1));
    P.bar.alloc((size_t)2 * B.ninst);
    CPG_HIP(hipMemset(P.bar.p, 0, P.bar.n * sizeof(unsigned)));
    P.abort_flag.alloc(1);
    P.last_iters.assign(B.ninst, 1);
    P.cost.assign(B.ninst, 0.0);
    P.clk.alloc((size_t)2 * B.ninst);
    P.h_clk.assign((size_t)2 * B.ninst, 0);
    P.lds = persist_lds(B.m);
    CPG_CHECK(P.lds <= 160 * 1024, CPG_ERR_UNSUPPORTED, "persistent batched solver: %zu bytes of LDS at restart %d",
              P.lds, B.m);
    const void *k = persist_kernel_of(B.sp.precision);
    if (P.lds > 64 * 1024)
        CPG_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds));
    int dev = 0, coop = 0;
    CPG_HIP(hipGetDevice(&dev));
    CPG_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    CPG_CHECK(coop, CPG_ERR_UNSUPPORTED, "the device has no cooperative launch");
    CPG_HIP(hipDeviceGetAttribute(&P.cus, hipDeviceAttributeMultiprocessorCount, dev));
    int khz = 0;
    CPG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    P.clock_khz = khz > 0 ? khz : 1.0;
    CPG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&P.occupancy, k, pgm::PT, P.lds));
    CPG_CHECK(P.occupancy > 0, CPG_ERR_UNSUPPORTED, "persistent batched solver: no occupancy at restart %d", B.m);
    for (int c : {1, 2, 4, P.occupancy})
        if (c <= P.occupancy && std::find(P.cand.begin(), P.cand.end(), c) == P.cand.end()) P.cand.push_back(c);
    P.cand_cost.assign(P.cand.size(), 0.0);
    CPG_HIP(hipEventCreate(&P.e0));
    CPG_HIP(hipEventCreate(&P.e1));
}

// teams for a grid of about wg_per_cu workgroups per CU.  Instance q's work in this solve is
// predicted as its Arnoldi count in the last solve times its cost per step (the team time x team
// size per step measured in the last solve; before any measurement its bytes per step), and its
// time as that work over its team size.  Every team starts with one workgroup and the next
// workgroup goes to the team with the longest predicted time below its cap (cap[q], max_team),
// which minimises the longest team under that model; all teams run in the one launch.
void persist_plan(cpg_batch_s &B, int wg_per_cu, int slot) {
    auto &P = B.P;
    const std::vector<int> &its =
        slot >= 0 && slot < (int)P.slot_iters.size() && !P.slot_iters[slot].empty() ? P.slot_iters[slot] : P.last_iters;
    const int W = std::max(B.ninst, std::min(wg_per_cu, P.occupancy) * P.cus);
    bool measured = true;
    for (int q = 0; q < B.ninst; q++) measured = measured && P.cost[q] > 0.0;
    std::vector<double> work(B.ninst);
    std::vector<int> n(B.ninst, 1), lim(B.ninst);
    for (int q = 0; q < B.ninst; q++) {
        work[q] = (measured ? P.cost[q] : P.bytes[q]) * std::max(1, its[q]);
        lim[q] = P.max_team > 0 ? std::min(P.cap[q], P.max_team) : P.cap[q];
    }
    std::vector<std::pair<double, int>> heap;
    for (int q = 0; q < B.ninst; q++)
        if (n[q] < lim[q]) heap.push_back({work[q] / n[q], q});
    std::make_heap(heap.begin(), heap.end());
    for (int used = B.ninst; used < W && !heap.empty(); used++) {
        std::pop_heap(heap.begin(), heap.end());
        const int q = heap.back().second;
        heap.pop_back();
        n[q]++;
        if (n[q] < lim[q]) {
            heap.push_back({work[q] / n[q], q});
            std::push_heap(heap.begin(), heap.end());
        }
    }
    std::vector<pgm::Inst> inst = P.h_inst;
    std::vector<int> wg_inst;
    int g0 = 0;
    for (int q = 0; q < B.ninst; q++) {
        inst[q].wg0 = g0;
        inst[q].nwg = n[q];
        for (int k = 0; k < n[q]; k++) wg_inst.push_back(q);
        g0 += n[q];
    }
    CPG_CHECK(g0 <= P.occupancy * P.cus, CPG_ERR_UNSUPPORTED,
              "persistent batched solver: %d instances need %d resident workgroups, the device holds %d", B.ninst,
              g0, P.occupancy * P.cus);
    bool same = g0 == P.grid;
    for (int q = 0; same && q < B.ninst; q++) same = inst[q].wg0 == P.h_inst[q].wg0 && inst[q].nwg == P.h_inst[q].nwg;
    if (same) return;
    P.h_inst = inst;
    P.grid = g0;
    P.inst.upload(inst, B.s);
    P.wg_inst.upload(wg_inst, B.s);
}

template <class T>
void persist_launch(cpg_batch_s &B) {
    auto &P = B.P;
    pgm::Args a{};
    a.cls = P.cls.p;
    a.inst = P.inst.p;
    a.wg_inst = P.wg_inst.p;
    a.phase = P.phase.p;
    a.tasks = P.tasks.p;
    a.r = B.r.p;
    a.x = B.x.p;
    a.V = B.V.p;
    a.Z = B.Z.p;
    a.w = B.w.p;
    a.y = P.y.p;
    a.N = (long long)B.N;
    a.partial = P.partial.p;
    a.nchunk = P.nchunk;
    a.bar = P.bar.p;
    a.abort_flag = P.abort_flag.p;
    a.iters = B.iters.p;
    a.conv = B.conv.p;
    a.resid = B.resid.p;
    a.clk = P.clk.p;
    a.m = B.m;
    a.maxcyc = std::max(1, (B.sp.maxit + B.m - 1) / B.m);
    a.pc = B.sp.pc_kind == CPG_PC_NONE ? 0 : 1;
    a.rtol = B.sp.rtol;
    a.spin_max = P.spin_max;
    void *args[] = {&a};
    CPG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(pgm::k_batch_fgmres<T>), dim3(P.grid),
                                       dim3(pgm::PT), args, (unsigned)P.lds, B.s));
}

// solve A x = r for every instance in one cooperative launch (x = 0 start, FGMRES tolerance
// rtol * ||r_q||)
// slot: the solve's place in the step (stage x Newton iteration), keying the Arnoldi-count history
void batch_solve_persistent(cpg_batch_s &B, cpg_step_stats &st, int slot) {
    auto &P = B.P;
    // launch shape: the fixed choice, or the next autotuner candidate, or the best one measured
    int wpc = P.wg_per_cu;
    const bool tuning = wpc <= 0 && P.trial < (int)P.cand.size();
    if (wpc <= 0) {
        if (tuning) {
            wpc = P.cand[P.trial];
        } else {
            size_t best = 0;
            for (size_t c = 1; c < P.cand.size(); c++)
                if (P.cand_cost[c] < P.cand_cost[best]) best = c;
            wpc = P.cand[best];
        }
    }
    wpc = std::min(wpc, P.occupancy);
    persist_plan(B, wpc, slot);
    CPG_HIP(hipMemsetAsync(P.abort_flag.p, 0, sizeof(int), B.s));
    CPG_HIP(hipEventRecord(P.e0, B.s));
    switch (B.sp.precision) {
    case CPG_FP64: persist_launch<double>(B); break;
    case CPG_FP32: persist_launch<float>(B); break;
    default: persist_launch<bf16>(B); break;
    }
    CPG_HIP(hipEventRecord(P.e1, B.s));
    int abort_flag = 0;
    CPG_HIP(hipMemcpyAsync(B.h_iters.data(), B.iters.p, B.ninst * sizeof(int), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipMemcpyAsync(B.h_resid.data(), B.resid.p, B.ninst * sizeof(double), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipMemcpyAsync(B.h_conv.data(), B.conv.p, B.ninst * sizeof(int), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipMemcpyAsync(&abort_flag, P.abort_flag.p, sizeof(int), hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipMemcpyAsync(P.h_clk.data(), P.clk.p, P.h_clk.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, B.s));
    CPG_HIP(hipStreamSynchronize(B.s));
    if (abort_flag) {
        // the barriers' counters are left mid-count: start the next solve from zero
        CPG_HIP(hipMemsetAsync(P.bar.p, 0, P.bar.n * sizeof(unsigned), B.s));
        CPG_HIP(hipStreamSynchronize(B.s));
        CPG_CHECK(false, CPG_ERR_HIP, "persistent batched solve: a team barrier waited past its limit (%u rounds)",
                  P.spin_max);
    }
    int mx = 0;
    double work = 0.0;
    for (int q = 0; q < B.ninst; q++) {
        mx = std::max(mx, B.h_iters[q]);
        work += P.bytes[q] * B.h_iters[q];
        P.last_iters[q] = B.h_iters[q];
        if (slot >= 0) {
            if ((int)P.slot_iters.size() <= slot) P.slot_iters.resize(slot + 1);
            P.slot_iters[slot].assign(B.h_iters.begin(), B.h_iters.end());
        }
        if (B.h_iters[q] > 0 && P.h_clk[2 * q + 1] > P.h_clk[2 * q])
            P.cost[q] = (double)(P.h_clk[2 * q + 1] - P.h_clk[2 * q]) * P.h_inst[q].nwg / B.h_iters[q];
        CPG_CHECK(std::isfinite(B.h_resid[q]), CPG_ERR_NONFINITE, "non-finite GMRES residual in batch instance %d", q);
    }
    if (tuning) {
        float ms = 0.f;
        CPG_HIP(hipEventElapsedTime(&ms, P.e0, P.e1));
        P.cand_cost[P.trial] = work > 0.0 ? ms / work : 0.0;
        P.trial++;
    }
    st.gmres_iters += mx;
    st.linear_solves++;
    for (int q = 0; q < B.ninst; q++)
        CPG_CHECK(B.h_conv[q], CPG_ERR_NOT_CONVERGED, "batched GMRES did not converge in %d cycles of %d (instance %d)",
                  std::max(1, (B.sp.maxit + B.m - 1) / B.m), B.m, q);
}

}  // namespace

extern "C" {

int cpg_batch_create(int nclass, const cpg_equation *eqs, const int *ninst, const int *const *cell_ptr, int nstages,
                     const cpg_solver_params *params, int mode, cpg_batch *out) {
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
            CPG_CHECK(mode >= CPG_BATCH_EAGER && mode <= CPG_BATCH_PERSISTENT, CPG_ERR_ARG, "bad batch mode %d", mode);
            B->mode = mode;
            B->use_graph = mode == CPG_BATCH_GRAPH;
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
                ck(cpg_prec_set_flow(C.P, params->pc_level_launches ? 0 : -1, -1));
                C.cdt.alloc(C.n);
                C.d_cptr.upload(C.cptr);
                // launch shapes of this class's block size: matvec (FP64) and sweeps (factor precision)
                int best = 0;
                if (mode != CPG_BATCH_PERSISTENT) {
                    ck(cpg_stream_autotune(C.J, CPG_FP64, 3, &best, nullptr, B->s));
                    if (params->precision != CPG_FP64)   // the stream's storage (BF16 for BF16 with FP32 D^-1)
                        ck(cpg_stream_autotune(C.J, params->precision == CPG_BF16_D32 ? CPG_BF16 : params->precision,
                                               3, &best, nullptr, B->s));
                }
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
            if (mode == CPG_BATCH_PERSISTENT) persist_setup(*B);
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

int cpg_batch_set_teams(cpg_batch B, int wg_per_cu, int max_team, unsigned spin_limit) {
    return guard([&] {
        CPG_CHECK(B && wg_per_cu >= 0 && max_team >= 0, CPG_ERR_ARG, "bad argument");
        CPG_CHECK(B->mode == CPG_BATCH_PERSISTENT, CPG_ERR_ARG, "not a persistent batch");
        B->P.wg_per_cu = wg_per_cu;
        B->P.max_team = max_team;
        B->P.spin_max = spin_limit ? spin_limit : 1u << 22;
    });
}

int cpg_batch_teams(cpg_batch B, int *grid, int *wg_per_cu, int *team_wg, double *team_ms) {
    return guard([&] {
        CPG_CHECK(B, CPG_ERR_ARG, "null batch");
        CPG_CHECK(B->mode == CPG_BATCH_PERSISTENT, CPG_ERR_ARG, "not a persistent batch");
        if (grid) *grid = B->P.grid;
        if (wg_per_cu) {
            int w = B->P.wg_per_cu;
            if (w <= 0 && !B->P.cand.empty()) {
                size_t best = 0;
                for (size_t c = 1; c < B->P.cand.size(); c++)
                    if (B->P.cand_cost[c] < B->P.cand_cost[best]) best = c;
                w = B->P.trial < (int)B->P.cand.size() ? 0 : B->P.cand[best];
            }
            *wg_per_cu = w;
        }
        if (team_wg)
            for (int q = 0; q < B->ninst; q++) team_wg[q] = B->P.h_inst[q].nwg;
        if (team_ms)
            for (int q = 0; q < B->ninst; q++) {
                const auto t0 = B->P.h_clk[2 * q], t1 = B->P.h_clk[2 * q + 1];
                team_ms[q] = t1 > t0 ? (double)(t1 - t0) / B->P.clock_khz : 0.0;
            }
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
                if (B->mode == CPG_BATCH_PERSISTENT) batch_solve_persistent(*B, st, stg * B->sp.newton_maxit + it);
                else batch_solve(*B, st);
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
