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
//      block's source is src1[j] times c1 (j = bcol >= 0) or src0[-1-j] (bcol < 0); the
//      BF16 per-block power-of-two scale is folded in here, so the stream loop does none of
//      this work;
//   2. stream: a group of TPC = bl/VW lanes covers one column with 16-byte loads (VW
//      elements of one column per lane); the row's columns are dealt to G = 4 * 64/TPC
//      logical groups (column cg to group cg mod G), each wavefront load is a contiguous
//      ~1 KB segment of the stream, and every lane owns VW fixed rows of 4/WPR logical groups
//      and keeps their partial sums in registers;
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

// packed streams: item w (blockIdx.x) writes row rows[w] (w if rows is null) from the packed
// blocks [bptr[w], bptr[w+1]) (block w alone if bptr is null); bcol[b] selects the source
// of packed block b (the row itself from src1 if bcol is null)
struct RowStream {
    const int *rows;
    const int *bptr;
    const int *bcol;
    const int *active = nullptr;   // device flag: when it reads 0 the launch does nothing
};

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

typedef double stream_d2 __attribute__((ext_vector_type(2)));
typedef float stream_f4 __attribute__((ext_vector_type(4)));
typedef float stream_f2 __attribute__((ext_vector_type(2)));
typedef uint32_t stream_u4 __attribute__((ext_vector_type(4)));
typedef uint32_t stream_u2 __attribute__((ext_vector_type(2)));

// the blocks are read once per sweep: non-temporal loads keep them from evicting the
// gathered vectors from the caches
template <class T, int VW, class A>
__device__ inline void stream_load(const T *__restrict__ p, A (&a)[VW]) {
    if constexpr (sizeof(T) == 8) {
        if constexpr (VW == 2) {
            const stream_d2 v = __builtin_nontemporal_load(reinterpret_cast<const stream_d2 *>(p));
            a[0] = (A)v.x;
            a[1] = (A)v.y;
        } else {
            a[0] = (A)__builtin_nontemporal_load(reinterpret_cast<const double *>(p));
        }
    } else if constexpr (sizeof(T) == 4) {
        if constexpr (VW == 4) {
            const stream_f4 v = __builtin_nontemporal_load(reinterpret_cast<const stream_f4 *>(p));
            a[0] = (A)v.x; a[1] = (A)v.y; a[2] = (A)v.z; a[3] = (A)v.w;
        } else if constexpr (VW == 2) {
            const stream_f2 v = __builtin_nontemporal_load(reinterpret_cast<const stream_f2 *>(p));
            a[0] = (A)v.x; a[1] = (A)v.y;
        } else {
            a[0] = (A)__builtin_nontemporal_load(reinterpret_cast<const float *>(p));
        }
    } else {
        // BF16: the high half of an FP32 word; element 2k is the low half of word k
        constexpr int NW = (VW + 1) / 2;
        uint32_t u[NW];
        if constexpr (VW == 8) {
            const stream_u4 v = __builtin_nontemporal_load(reinterpret_cast<const stream_u4 *>(p));
            u[0] = v.x; u[1] = v.y; u[2] = v.z; u[3] = v.w;
        } else if constexpr (VW == 4) {
            const stream_u2 v = __builtin_nontemporal_load(reinterpret_cast<const stream_u2 *>(p));
            u[0] = v.x; u[1] = v.y;
        } else if constexpr (VW == 2) {
            u[0] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p));
        } else {
            u[0] = (uint32_t)reinterpret_cast<const uint16_t *>(p)[0];
        }
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

// out may alias src1 (rows written by one launch are never sources of the same launch, and a
// row that is its own source is gathered before it is written).  Exactly one of base64 /
// base32 may be non-null.
template <class T, int BL, int WPR, class XT>
__global__ __launch_bounds__(64 * WPR) void k_rowstream(RowStream S, const T *__restrict__ F,
                                                        const float *__restrict__ scale,
                                                        const double *__restrict__ src0, const XT *src1, double c1,
                                                        const double *__restrict__ base64,
                                                        const float *__restrict__ base32, XT *out) {
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
    if (S.active && !*S.active) return;
    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    const int i = S.rows ? S.rows[w] : w;
    const int b0 = S.bptr ? S.bptr[w] : w;
    const int nb = S.bptr ? S.bptr[w + 1] - b0 : 1;
    const int ncol = nb * BL;
    for (int e = tid; e < ncol; e += NT) {
        const int b = e / BL, c = e - b * BL;
        const int j = S.bcol ? S.bcol[b0 + b] : i;
        double v = j >= 0 ? c1 * (double)src1[(size_t)j * BL + c] : src0[(size_t)(-1 - j) * BL + c];
        if constexpr (sizeof(T) == 2) v *= (double)scale[b0 + b];
        xs[e] = (A)v;
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    const int gl = lane / TPC;
    const int rr = (lane - gl * TPC) * VW;
    if (gl < GW) {
        // logical group of slot l of this lane: gl + GW (wv + WPR l)
        A acc[LPW][VW];
#pragma unroll
        for (int l = 0; l < LPW; l++)
#pragma unroll
            for (int v = 0; v < VW; v++) acc[l][v] = 0;
        const T *p = F + (size_t)b0 * BB + rr;
        const int g0 = gl + GW * wv;
        if constexpr (LPW == 1) {
#pragma unroll 4
            for (int cg = g0; cg < ncol; cg += G) {
                A a[VW];
                stream_load<T, VW, A>(p + (size_t)cg * BL, a);
                const A xv = xs[cg];
#pragma unroll
                for (int v = 0; v < VW; v++) acc[0][v] = fma(a[v], xv, acc[0][v]);
            }
        } else {
#pragma unroll 2
            for (int base = 0; base < ncol; base += G) {
#pragma unroll
                for (int l = 0; l < LPW; l++) {
                    const int cg = base + g0 + GW * WPR * l;
                    if (cg < ncol) {
                        A a[VW];
                        stream_load<T, VW, A>(p + (size_t)cg * BL, a);
                        const A xv = xs[cg];
#pragma unroll
                        for (int v = 0; v < VW; v++) acc[l][v] = fma(a[v], xv, acc[l][v]);
                    }
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
        double s = base64 ? base64[(size_t)i * BL + t] : base32 ? (double)base32[(size_t)i * BL + t] : 0.0;
        double r = 0.0;
#pragma unroll
        for (int gg = 0; gg < G; gg++) r += (double)red[gg * BL + t];
        out[(size_t)i * BL + t] = (XT)(s + r);
    }
}

// Any block size up to 256 (no compiled specialisation): one thread per row entry walks all
// of the row's columns.  Correct, not tuned.
template <class T, class XT>
__global__ __launch_bounds__(256) void k_rowstream_any(int bl, RowStream S, const T *__restrict__ F,
                                                       const float *__restrict__ scale,
                                                       const double *__restrict__ src0, const XT *src1, double c1,
                                                       const double *__restrict__ base64,
                                                       const float *__restrict__ base32, XT *out) {
    extern __shared__ double lds_d[];
    double *xs = lds_d;
    if (S.active && !*S.active) return;
    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    const int i = S.rows ? S.rows[w] : w;
    const int b0 = S.bptr ? S.bptr[w] : w;
    const int nb = S.bptr ? S.bptr[w + 1] - b0 : 1;
    const int ncol = nb * bl;
    for (int e = tid; e < ncol; e += 256) {
        const int b = e / bl, c = e - b * bl;
        const int j = S.bcol ? S.bcol[b0 + b] : i;
        double v = j >= 0 ? c1 * (double)src1[(size_t)j * bl + c] : src0[(size_t)(-1 - j) * bl + c];
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
                       a.src1, a.c1, a.base64, a.base32, a.out);
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
    const size_t lds = (size_t)a.nbmax * bl * sizeof(double);
    if (lds > 64 * 1024)
        CPG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_rowstream_any<T, XT>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    CPG_CHECK(lds <= 160 * 1024, CPG_ERR_UNSUPPORTED, "row of %d blocks of %d too long", a.nbmax, bl);
    hipLaunchKernelGGL((k_rowstream_any<T, XT>), dim3(a.items), dim3(256), lds, s, bl, a.S, F, scale, a.src0, a.src1,
                       a.c1, a.base64, a.base32, a.out);
    CPG_LAUNCH_CHECK();
}

}  // namespace cpg
