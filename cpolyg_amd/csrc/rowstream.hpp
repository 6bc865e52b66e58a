// Row-stream kernel: out_i = base_i + sum_b B_b * v_b over the packed blocks of one block row.
//
// Every bandwidth-bound sweep of the implicit solver has this shape: the BSR matvec
// (BlockMatrix::matvec, src/BlockMatrix.cpp:28-49), the forward and backward sweeps of block
// DILU (BlockILU0::solve, src/Preconditioners.cpp:125-158) and the block-Jacobi apply
// (BlockJacobi::solve, src/Preconditioners.cpp:39-47).  The blocks of one output row are
// stored back to back ("packed": bl x bl column-major each, in the order the row consumes
// them), so a row is a single contiguous stream of nb*bl columns of bl elements.
//
// One workgroup of WPR wavefronts owns one output row:
//   1. gather: the nb source sub-vectors v_b (bl values each) are read once into LDS.  A
//      block's source (bcol entry, see stream_source) is src1[j] times c1, src0[u], or
//      src2[k] times c2; the BF16 per-block power-of-two scale is folded in here, so the
//      stream loop does none of this work;
//   2. stream: a group of TPC = bl/VW lanes covers one column with 16-byte loads (VW
//      elements of one column per lane); the row's columns are dealt to G = 4 * 64/TPC
//      logical groups (column cg to group cg mod G), each wavefront load is a contiguous
//      ~1 KB segment of the stream, and every lane owns VW fixed rows of 4/WPR logical groups
//      and keeps their partial sums in registers.  Loads go out in chunks of four per lane,
//      the next chunk issued before the current one is consumed and the first one before
//      the gather: the short rows of the DILU sweeps (3-5 blocks) are then one round trip
//      to memory instead of one per column a lane owns;
//   3. reduce: the G partial vectors are summed through LDS in a fixed order and added to
//      base_i.  No atomics, and the logical groups do not depend on WPR, so the result is
//      bitwise the same for every launch shape the autotuner may pick.
// Storage types: FP64, FP32 (FP64 accumulation) and BF16 (FP32 accumulation of the
// per-lane partial sums, FP64 final sum).
#pragma once

#include "cpg_internal.hpp"

namespace cpg {

struct bf16 {
    uint16_t v;
};

// packed streams: item w (blockIdx.x) writes row rows[w] (row0 + w if rows is null) from the
// packed blocks [bptr[w], bptr[w+1]) (block row0 + w alone if bptr is null); bcol[b] selects the source
// of packed block b (the row itself from src1 if bcol is null)
struct RowStream {
    const int *rows;
    const int *bptr;
    const int *bcol;
    const int *active = nullptr;   // device flag: when it reads 0 the launch does nothing
    int row0 = 0;                  // rows null: item w writes row row0 + w (a contiguous level)
    // set: item w's row result z is followed by the packed block at dptr[w] (D^-1 of a DILU row
    // without upper blocks) and the launch writes that block times z instead of z
    const int *dptr = nullptr;
};

// Packed source of a block (bcol entry): j >= 0 reads c1 * src1[j]; -1 - u with u < kSrc2Tag
// reads src0[u]; -1 - (kSrc2Tag + k) reads c2 * src2[k]
constexpr int kSrc2Tag = 1 << 30;

// COH: the sources are read at device scope (vectors that other workgroups of the same launch write)
template <class XT, bool COH = false>
__device__ inline double stream_source(int j, int c, int bl, const double *__restrict__ src0, const XT *src1,
                                       double c1, const XT *src2, double c2) {
    if constexpr (sizeof(XT) == 8) {
        // one load from a selected pointer: no branch between the column index and its load
        const int u = -1 - j;
        const bool s1 = j >= 0, s2 = !s1 && u >= kSrc2Tag;
        const double *base = s1 ? (const double *)src1 : s2 ? (const double *)src2 : src0;
        const int r = s1 ? j : s2 ? u - kSrc2Tag : u;
        const double cf = s1 ? c1 : s2 ? c2 : 1.0;
        if constexpr (COH)
            return cf * __hip_atomic_load(base + (size_t)r * bl + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return cf * base[(size_t)r * bl + c];
    }
    if (j >= 0) return c1 * (double)src1[(size_t)j * bl + c];
    const int u = -1 - j;
    if (u < kSrc2Tag) return src0[(size_t)u * bl + c];
    return c2 * (double)src2[(size_t)(u - kSrc2Tag) * bl + c];
}

template <class T>
struct StreamAcc {
    using type = double;
};
template <>
struct StreamAcc<bf16> {
    using type = float;
};

// elements per lane load: the widest power of two <= 16 bytes that divides bl
template <class T, int BL>
constexpr int stream_vw() {
    int w = 16 / (int)sizeof(T);
    while (w > 1 && BL % w) w >>= 1;
    return w;
}

typedef uint32_t stream_u4 __attribute__((ext_vector_type(4)));
typedef uint32_t stream_u2 __attribute__((ext_vector_type(2)));

// One lane's VW elements of a column as raw bytes (16 at most) in a 4-dword register tuple.
// The blocks are read once per sweep: non-temporal loads keep them from evicting the
// gathered vectors from the caches.
template <class T, int VW>
__device__ inline stream_u4 raw_load(const T *__restrict__ p) {
    constexpr int B = (int)sizeof(T) * VW;
    stream_u4 r = {0u, 0u, 0u, 0u};
    if constexpr (B == 16) {
        r = __builtin_nontemporal_load(reinterpret_cast<const stream_u4 *>(p));
    } else if constexpr (B == 8) {
        const stream_u2 v = __builtin_nontemporal_load(reinterpret_cast<const stream_u2 *>(p));
        r.x = v.x;
        r.y = v.y;
    } else if constexpr (B == 4) {
        r.x = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p));
    } else {
        r.x = (uint32_t)reinterpret_cast<const uint16_t *>(p)[0];
    }
    return r;
}
// the VW elements of a raw_load as accumulator values (BF16: the high half of an FP32 word,
// element 2k the low half of word k)
template <class T, int VW, class A>
__device__ inline void raw_convert(stream_u4 r, A (&a)[VW]) {
    uint32_t u[4] = {r.x, r.y, r.z, r.w};
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int k = 0; k < VW; k++) {
            const uint64_t b = ((uint64_t)u[2 * k + 1] << 32) | u[2 * k];
            double d;
            __builtin_memcpy(&d, &b, 8);
            a[k] = (A)d;
        }
    } else if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < VW; k++) a[k] = (A)__uint_as_float(u[k]);
    } else {
#pragma unroll
        for (int k = 0; k < VW; k++) {
            const uint32_t bits = (k & 1) ? (u[k >> 1] & 0xffff0000u) : (u[k >> 1] << 16);
            a[k] = (A)__uint_as_float(bits);
        }
    }
}

// LDS bytes of k_rowstream for rows of at most nbmax blocks
template <class T, int BL, int WPR>
inline size_t rowstream_lds(int nbmax) {
    using A = typename StreamAcc<T>::type;
    constexpr int TPC = BL / stream_vw<T, BL>();
    constexpr int G = (64 / TPC) * 4;
    return (size_t)(G + (nbmax > 0 ? nbmax : 1)) * BL * sizeof(A);
}

// Hooks of a sweep whose rows wait for the rows they read (none: the level launches order them).
// coherent: the row's sources and its result are read and written at device scope.
// row(): the item's row i, first packed block b0 and block count nb; tag(): the source tag of
// packed block b0 + b; has_d() / dblock(): the fused D^-1 block (RowStream::dptr).  coh_src /
// coh_out: the row's sources are read / its result written at device scope.
struct NoFlow {
    static constexpr bool coh_src = false, coh_out = false;
    __device__ void row(const RowStream &S, int w, int &i, int &b0, int &nb) const {
        i = S.rows ? S.rows[w] : S.row0 + w;
        b0 = S.bptr ? S.bptr[w] : S.row0 + w;
        nb = S.bptr ? S.bptr[w + 1] - b0 : 1;
    }
    __device__ int tag(const RowStream &S, int b0, int b, int i) const { return S.bcol ? S.bcol[b0 + b] : i; }
    __device__ bool has_d(const RowStream &S) const { return S.dptr != nullptr; }
    __device__ int dblock(const RowStream &S, int w) const { return S.dptr[w]; }
    __device__ void prefetch(int, int) {}
    __device__ void wait(int, int, int, const int *, int, int) {}
    __device__ void publish(int, int) {}
};

template <class XT, bool COH>
__device__ inline void stream_store(XT *p, XT v) {
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// Item w of the sweep S (one workgroup of WPR wavefronts).  out may alias src1 and src2 (rows
// written by one sweep are never sources of the same level, and a row that is its own source is
// gathered before it is written).  Exactly one of base64 / base32 may be non-null.
template <class T, int BL, int WPR, class XT, class Flow>
__device__ __forceinline__ void rowstream_row(const RowStream &S, int w, const T *__restrict__ F,
                                              const float *__restrict__ scale, const double *__restrict__ src0,
                                              const XT *src1, double c1, const XT *src2, double c2,
                                              const double *__restrict__ base64, const float *__restrict__ base32,
                                              XT *out, Flow &fl) {
    using A = typename StreamAcc<T>::type;
    constexpr int VW = stream_vw<T, BL>();
    constexpr int TPC = BL / VW;
    static_assert(TPC <= 64, "a column must fit one wavefront");
    constexpr int GW = 64 / TPC;
    constexpr int G = GW * 4;       // logical groups (independent of WPR)
    constexpr int LPW = 4 / WPR;    // logical groups per lane
    constexpr int NT = 64 * WPR;
    constexpr size_t BB = (size_t)BL * BL;
    extern __shared__ double lds_d[];
    A *red = reinterpret_cast<A *>(lds_d);
    A *xs = red + G * BL;
    const int tid = threadIdx.x;
    int i, b0, nb;
    fl.row(S, w, i, b0, nb);
    const int ncol = nb * BL;
    const int lane = tid & 63, wv = tid >> 6;
    const int gl = lane / TPC;
    const int rr = (lane - gl * TPC) * VW;
    const int g0 = gl + GW * wv;
    const T *p = F + (size_t)b0 * BB + rr;
    // the stream phase takes the row in chunks of PF columns per logical-group slot: every
    // chunk's loads are issued together (the next chunk's before this one is consumed), and
    // the first chunk is issued before the gather, so its latency overlaps the gather's
    constexpr int PF = 4 / LPW;
    constexpr int CH = G * PF;   // columns per chunk
    stream_u4 cur[LPW][PF] = {};
    auto issue = [&](stream_u4 (&r)[LPW][PF], int base) {
#pragma unroll
        for (int k = 0; k < PF; k++)
#pragma unroll
            for (int l = 0; l < LPW; l++) {
                const int cg = base + g0 + GW * WPR * l + G * k;
                if (cg < ncol) r[l][k] = raw_load<T, VW>(p + (size_t)cg * BL);
            }
    };
    fl.prefetch(i, tid);
    if (gl < GW) issue(cur, 0);
    fl.wait(i, b0, nb, S.bcol, tid, NT);
    for (int e = tid; e < ncol; e += NT) {
        const int b = e / BL, c = e - b * BL;
        const int j = fl.tag(S, b0, b, i);
        double v = stream_source<XT, Flow::coh_src>(j, c, BL, src0, src1, c1, src2, c2);
        if constexpr (sizeof(T) == 2) v *= (double)scale[b0 + b];
        xs[e] = (A)v;
    }
    __syncthreads();
    if (gl < GW) {
        A acc[LPW][VW];
#pragma unroll
        for (int l = 0; l < LPW; l++)
#pragma unroll
            for (int v = 0; v < VW; v++) acc[l][v] = 0;
        for (int base = 0; base < ncol; base += CH) {
            stream_u4 nxt[LPW][PF] = {};
            if (base + CH < ncol) issue(nxt, base + CH);
#pragma unroll
            for (int k = 0; k < PF; k++)
#pragma unroll
                for (int l = 0; l < LPW; l++) {
                    const int cg = base + g0 + GW * WPR * l + G * k;
                    if (cg < ncol) {
                        A a[VW];
                        raw_convert<T, VW, A>(cur[l][k], a);
                        const A xv = xs[cg];
#pragma unroll
                        for (int v = 0; v < VW; v++) acc[l][v] = fma(a[v], xv, acc[l][v]);
                    }
                }
#pragma unroll
            for (int k = 0; k < PF; k++)
#pragma unroll
                for (int l = 0; l < LPW; l++) cur[l][k] = nxt[l][k];
        }
#pragma unroll
        for (int l = 0; l < LPW; l++)
#pragma unroll
            for (int v = 0; v < VW; v++) red[(g0 + GW * WPR * l) * BL + rr + v] = acc[l][v];
    }
    __syncthreads();
    if (!fl.has_d(S)) {
        for (int t = tid; t < BL; t += NT) {
            double s = base64 ? base64[(size_t)i * BL + t] : base32 ? (double)base32[(size_t)i * BL + t] : 0.0;
            double r = 0.0;
#pragma unroll
            for (int gg = 0; gg < G; gg++) r += (double)red[gg * BL + t];
            stream_store<XT, Flow::coh_out>(out + (size_t)i * BL + t, (XT)(s + r));
        }
        fl.publish(i, tid);
        return;
    }
    // the row result becomes the source of one more block: gathered (and BF16-scaled) exactly as
    // a separate launch would gather it, then streamed and reduced in the same order
    const int bd = fl.dblock(S, w);
    double dsc = 1.0;
    if constexpr (sizeof(T) == 2) dsc = (double)scale[bd];
    for (int t = tid; t < BL; t += NT) {
        double s = base64 ? base64[(size_t)i * BL + t] : base32 ? (double)base32[(size_t)i * BL + t] : 0.0;
        double r = 0.0;
#pragma unroll
        for (int gg = 0; gg < G; gg++) r += (double)red[gg * BL + t];
        xs[t] = (A)((s + r) * dsc);
    }
    __syncthreads();
    if (gl < GW) {
        A acc[LPW][VW];
#pragma unroll
        for (int l = 0; l < LPW; l++)
#pragma unroll
            for (int v = 0; v < VW; v++) acc[l][v] = 0;
        const T *pd = F + (size_t)bd * BB + rr;
        for (int base = 0; base < BL; base += G) {
#pragma unroll
            for (int l = 0; l < LPW; l++) {
                const int cg = base + g0 + GW * WPR * l;
                if (cg < BL) {
                    A a[VW];
                    raw_convert<T, VW, A>(raw_load<T, VW>(pd + (size_t)cg * BL), a);
                    const A xv = xs[cg];
#pragma unroll
                    for (int v = 0; v < VW; v++) acc[l][v] = fma(a[v], xv, acc[l][v]);
                }
            }
        }
#pragma unroll
        for (int l = 0; l < LPW; l++)
#pragma unroll
            for (int v = 0; v < VW; v++) red[(g0 + GW * WPR * l) * BL + rr + v] = acc[l][v];
    }
    __syncthreads();
    for (int t = tid; t < BL; t += NT) {
        double r = 0.0;
#pragma unroll
        for (int gg = 0; gg < G; gg++) r += (double)red[gg * BL + t];
        stream_store<XT, Flow::coh_out>(out + (size_t)i * BL + t, (XT)r);
    }
    fl.publish(i, tid);
}

// one launch per level: item w = blockIdx.x
template <class T, int BL, int WPR, class XT>
__global__ __launch_bounds__(64 * WPR) void k_rowstream(RowStream S, const T *__restrict__ F,
                                                        const float *__restrict__ scale,
                                                        const double *__restrict__ src0, const XT *src1, double c1,
                                                        const XT *src2, double c2,
                                                        const double *__restrict__ base64,
                                                        const float *__restrict__ base32, XT *out) {
    if (S.active && !*S.active) return;
    NoFlow fl;
    rowstream_row<T, BL, WPR, XT>(S, (int)blockIdx.x, F, scale, src0, src1, c1, src2, c2, base64, base32, out, fl);
}

// Any block size up to 256 (no compiled specialisation): one thread per row entry walks all
// of the row's columns.  Correct, not tuned.
template <class T, class XT>
__global__ __launch_bounds__(256) void k_rowstream_any(int bl, RowStream S, const T *__restrict__ F,
                                                       const float *__restrict__ scale,
                                                       const double *__restrict__ src0, const XT *src1, double c1,
                                                       const XT *src2, double c2,
                                                       const double *__restrict__ base64,
                                                       const float *__restrict__ base32, XT *out) {
    extern __shared__ double lds_d[];
    double *xs = lds_d;
    if (S.active && !*S.active) return;
    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    const int i = S.rows ? S.rows[w] : S.row0 + w;
    const int b0 = S.bptr ? S.bptr[w] : S.row0 + w;
    const int nb = S.bptr ? S.bptr[w + 1] - b0 : 1;
    const int ncol = nb * bl;
    for (int e = tid; e < ncol; e += 256) {
        const int b = e / bl, c = e - b * bl;
        const int j = S.bcol ? S.bcol[b0 + b] : i;
        double v = stream_source<XT>(j, c, bl, src0, src1, c1, src2, c2);
        if constexpr (sizeof(T) == 2) v *= (double)scale[b0 + b];
        xs[e] = v;
    }
    __syncthreads();
    const size_t bb = (size_t)bl * bl;
    for (int r = tid; r < bl; r += 256) {
        double s = base64 ? base64[(size_t)i * bl + r] : base32 ? (double)base32[(size_t)i * bl + r] : 0.0;
        const T *p = F + (size_t)b0 * bb + r;
        for (int c = 0; c < ncol; c++) {
            double a;
            if constexpr (sizeof(T) == 2) a = (double)__uint_as_float((uint32_t)p[(size_t)c * bl].v << 16);
            else a = (double)p[(size_t)c * bl];
            s = fma(a, xs[c], s);
        }
        out[(size_t)i * bl + r] = (XT)s;
    }
}

// Launch parameters of one row-stream sweep
template <class XT>
struct StreamArgs {
    RowStream S;
    int items = 0;
    int nbmax = 1;
    const double *src0 = nullptr;
    const XT *src1 = nullptr;
    double c1 = 1.0;
    const XT *src2 = nullptr;
    double c2 = 1.0;
    const double *base64 = nullptr;
    const float *base32 = nullptr;
    XT *out = nullptr;
};

template <class T, int BL, int WPR, class XT>
inline void launch_rowstream_wpr(const StreamArgs<XT> &a, const T *F, const float *scale, hipStream_t s) {
    const size_t lds = rowstream_lds<T, BL, WPR>(a.nbmax);
    if (lds > 64 * 1024)
        CPG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_rowstream<T, BL, WPR, XT>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_rowstream<T, BL, WPR, XT>), dim3(a.items), dim3(64 * WPR), lds, s, a.S, F, scale, a.src0,
                       a.src1, a.c1, a.src2, a.c2, a.base64, a.base32, a.out);
    CPG_LAUNCH_CHECK();
}

template <class T, int BL, class XT>
inline void launch_rowstream_bl(const StreamArgs<XT> &a, const T *F, const float *scale, int wpr, hipStream_t s) {
    switch (wpr) {
    case 1: launch_rowstream_wpr<T, BL, 1, XT>(a, F, scale, s); break;
    case 2: launch_rowstream_wpr<T, BL, 2, XT>(a, F, scale, s); break;
    default: launch_rowstream_wpr<T, BL, 4, XT>(a, F, scale, s); break;
    }
}

// block sizes with a compiled k_rowstream (the others run k_rowstream_any, which has no dptr)
inline bool rowstream_compiled(int bl) {
    return bl == 3 || bl == 6 || bl == 10 || bl == 12 || bl == 15 || bl == 24 || bl == 40 || bl == 60 || bl == 120;
}

// one sweep over a.items rows; wpr <= 0 takes the tuned value
template <class T, class XT>
inline void launch_rowstream(int bl, const StreamArgs<XT> &a, const T *F, const float *scale, hipStream_t s,
                             int wpr = 0) {
    if (a.items <= 0) return;
    if (wpr <= 0) wpr = stream_wpr((int)sizeof(T), bl);
    switch (bl) {
    case 3: launch_rowstream_bl<T, 3, XT>(a, F, scale, wpr, s); return;
    case 6: launch_rowstream_bl<T, 6, XT>(a, F, scale, wpr, s); return;
    case 10: launch_rowstream_bl<T, 10, XT>(a, F, scale, wpr, s); return;
    case 12: launch_rowstream_bl<T, 12, XT>(a, F, scale, wpr, s); return;
    case 15: launch_rowstream_bl<T, 15, XT>(a, F, scale, wpr, s); return;
    case 24: launch_rowstream_bl<T, 24, XT>(a, F, scale, wpr, s); return;
    case 40: launch_rowstream_bl<T, 40, XT>(a, F, scale, wpr, s); return;
    case 60: launch_rowstream_bl<T, 60, XT>(a, F, scale, wpr, s); return;
    case 120: launch_rowstream_bl<T, 120, XT>(a, F, scale, wpr, s); return;
    default: break;
    }
    CPG_CHECK(bl <= 256, CPG_ERR_UNSUPPORTED, "block size %d > 256", bl);
    CPG_CHECK(!a.S.dptr, CPG_ERR_UNSUPPORTED, "block size %d: no fused D^-1 step", bl);
    const size_t lds = (size_t)a.nbmax * bl * sizeof(double);
    if (lds > 64 * 1024)
        CPG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_rowstream_any<T, XT>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    CPG_CHECK(lds <= 160 * 1024, CPG_ERR_UNSUPPORTED, "row of %d blocks of %d too long", a.nbmax, bl);
    hipLaunchKernelGGL((k_rowstream_any<T, XT>), dim3(a.items), dim3(256), lds, s, bl, a.S, F, scale, a.src0, a.src1,
                       a.c1, a.src2, a.c2, a.base64, a.base32, a.out);
    CPG_LAUNCH_CHECK();
}

}  // namespace cpg
