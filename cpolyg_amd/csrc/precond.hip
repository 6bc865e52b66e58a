// Block DILU ("BlockILU0") and block-Jacobi preconditioners on the device.
//
// Reference: BlockILU0 (src/Preconditioners.cpp:49-158) is block DILU:
//   D_j = A_jj - sum_{i<j} A_ji D_i^-1 A_ij,  L_ji = A_ji D_i^-1,  DU_ij = D_i^-1 A_ij,
// applied as x <- (I + DU)^-1 D^-1 (I + L)^-1 x.  Rows are processed in levels.  For a
// colour-ordered matrix (every neighbour pair i<j has colour(i) < colour(j)) the levels are
// contiguous row ranges, one per colour (see segments()), and the backward sweep takes them
// from the last; otherwise a row's forward level is one more than the deepest lower
// neighbour's and its backward level one more than the deepest upper neighbour's.
//
// Factor (left-looking, one workgroup per row of a forward level):
//   AD = A_ii - sum_k L_ik A_ki with L_ik = A_ik D_k^-1 (k lower neighbour),
//   D_i^-1 by Gauss-Jordan with partial pivoting in LDS, DU_ij = D_i^-1 A_ij.
// The factor runs in FP64 and rounds once when storing to the requested precision;
// an FP64 copy of D^-1 is kept for the L products of later levels.
//
// Storage: the factor blocks are packed in the order the sweeps read them (rowstream.hpp):
// the forward stream holds, level by level and row by row, the L blocks of each row; the
// backward stream, level by level (from the last), each row's D^-1 followed by its DU
// blocks.  pos[slot] maps a slot of A to its packed block.
// Apply (row-stream launches, one per level and direction; forward level 0 is y = x and the
// last forward level also forms x = D^-1 y for its rows, see apply_t):
//   forward  y_i = x_i - sum_k L_ik y_k,
//   backward x_i = D_i^-1 y_i - sum_j DU_ij x_j,
// which is (I + DU)^-1 D^-1 (I + L)^-1 x with the D^-1 product moved into the backward
// sweep, so every row of either sweep is one independent stream of blocks.
#include "cpg_internal.hpp"
#include "rowstream.hpp"

#include <climits>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace cpg {

// ---- storage types -----------------------------------------------------------------

__host__ __device__ inline float bf16_to_f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// round to nearest even; NaN stays NaN
__host__ __device__ inline uint16_t f_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

// round to nearest even straight from FP64: the value is first rounded to FP32 by rounding to odd
// (truncate, then set the last bit if anything was dropped), which keeps the information the
// second rounding needs (FP32 carries 16 more significand bits than BF16), so the result is the
// correctly rounded BF16 of the FP64 value, not a double rounding through the nearest FP32
__device__ inline uint16_t d_to_bf16(double v) {
    const float t = __double2float_rz(v);
    uint32_t u;
    memcpy(&u, &t, 4);
    if ((double)t != v) u |= 1u;
    float f;
    memcpy(&f, &u, 4);
    return f_to_bf16(f);
}

template <class T>
struct Store;
template <>
struct Store<double> {
    using acc = double;
    static constexpr int prec = CPG_FP64;
    __device__ static void put(double *p, size_t e, double v, float) { p[e] = v; }
    __device__ static double get(const double *p, size_t e) { return p[e]; }
};
// Reduced-precision factors are only a storage format: the sweeps convert each element on
// load and accumulate, and keep their intermediate vectors, in FP64.  FP32 accumulation
// through the dependent colour levels of an ill-conditioned DILU costs 2-9x more Krylov
// iterations at CFL 20, while the extra bytes of FP64 intermediates are ~1% of the traffic.
template <>
struct Store<float> {
    using acc = double;
    static constexpr int prec = CPG_FP32;
    __device__ static void put(float *p, size_t e, double v, float) { p[e] = (float)v; }
    __device__ static float get(const float *p, size_t e) { return p[e]; }
};
template <>
struct Store<bf16> {
    using acc = double;
    static constexpr int prec = CPG_BF16;
    __device__ static void put(bf16 *p, size_t e, double v, float inv_scale) {
        p[e].v = d_to_bf16(v * (double)inv_scale);
    }
    __device__ static float get(const bf16 *p, size_t e) { return bf16_to_f(p[e].v); }
};

// ---- workgroup helpers ---------------------------------------------------------------
constexpr int WG = 256;

__device__ inline double wg_max_abs(double v, double *scratch) {
    // block-wide max of |v| (WG threads, 4 wavefronts)
    double a = fabs(v);
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) scratch[w] = a;
    __syncthreads();
    double r = scratch[0];
    for (int k = 1; k < WG / 64; k++) r = fmax(r, scratch[k]);
    __syncthreads();
    return r;
}

// power-of-two scale s with max|v|/s in (0.5, 1]; 1 for a zero block.  The exponent is kept in
// FP32's normal range, so s and 1/s are finite and exact; a block whose max lies outside it (beyond
// ~1e38 or below ~1e-38) is stored relative to the nearest end of the range instead
__device__ inline float pow2_scale(double mx) {
    if (!(mx > 0.0) || !isfinite(mx)) return 1.0f;
    int e;
    frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
    return ldexpf(1.0f, min(max(e, -125), 127));
}

// C(entries of this thread) = P * Q, all bl x bl column-major in LDS
template <int NE>
__device__ inline void gemm_lds(int bl, const double *P, const double *Q, double (&c)[NE]) {
    const int bb = bl * bl;
#pragma unroll
    for (int j = 0; j < NE; j++) {
        const int e = threadIdx.x + WG * j;
        double s = 0.0;
        if (e < bb) {
            const int r = e % bl, col = e / bl;
            const double *q = Q + col * bl;
            for (int k = 0; k < bl; k++) s = fma(P[k * bl + r], q[k], s);
        }
        c[j] = s;
    }
}

template <int NE>
__device__ inline void load_block(int bl, double *dst, const double *src) {
    const int bb = bl * bl;
#pragma unroll
    for (int j = 0; j < NE; j++) {
        const int e = threadIdx.x + WG * j;
        if (e < bb) dst[e] = src[e];
    }
}

// store a bl x bl block held in registers (entries threadIdx.x + WG*j) in precision T
template <class T, int NE>
__device__ inline void store_block(int bl, T *dst, float *scale_out, const double (&c)[NE], double *scratch) {
    const int bb = bl * bl;
    float inv = 1.0f;
    if constexpr (Store<T>::prec == CPG_BF16) {
        double mx = 0.0;
#pragma unroll
        for (int j = 0; j < NE; j++)
            if (threadIdx.x + WG * j < bb) mx = fmax(mx, fabs(c[j]));
        mx = wg_max_abs(mx, scratch);
        const float s = pow2_scale(mx);
        inv = 1.0f / s;
        if (threadIdx.x == 0) *scale_out = s;
    }
#pragma unroll
    for (int j = 0; j < NE; j++) {
        const int e = threadIdx.x + WG * j;
        if (e < bb) Store<T>::put(dst, e, c[j], inv);
    }
}

// Gauss-Jordan inversion with partial pivoting (first maximal |pivot|) of the bl x bl
// left half of AUG (column-major, bl rows, 2*bl columns); the right half must hold I.
// Returns false (uniformly) if a zero pivot was met.
__device__ bool gauss_jordan(int bl, double *AUG, double *fac, int *piv_row) {
    __shared__ double s_invd;
    bool ok = true;
    const int t = threadIdx.x;
    const int w2 = 2 * bl;
    for (int p = 0; p < bl; p++) {
        if (t < 64) {
            // wavefront 0 finds the pivot row among rows p..bl-1 (bl <= 64 handled per lane; larger in a loop)
            double best = -1.0;
            int bi = bl;
            for (int r = p + t; r < bl; r += 64) {
                const double a = fabs(AUG[p * bl + r]);
                if (a > best) { best = a; bi = r; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (t == 0) *piv_row = bi;
        }
        __syncthreads();
        const int pr = *piv_row;
        if (pr != p) {
            for (int c = t; c < w2; c += WG) {
                const double tmp = AUG[c * bl + p];
                AUG[c * bl + p] = AUG[c * bl + pr];
                AUG[c * bl + pr] = tmp;
            }
        }
        __syncthreads();
        // one thread reads the pivot; the barrier keeps every other thread from seeing it
        // after the row scaling below has overwritten it with 1
        if (t == 0) {
            const double d = AUG[p * bl + p];
            s_invd = d == 0.0 ? 0.0 : 1.0 / d;
        }
        __syncthreads();
        const double invd = s_invd;
        if (invd == 0.0) ok = false;
        for (int c = t; c < w2; c += WG) AUG[c * bl + p] *= (invd == 0.0 ? 1.0 : invd);
        for (int r = t; r < bl; r += WG) fac[r] = (r == p) ? 0.0 : AUG[p * bl + r];
        __syncthreads();
        const int ncell = bl * w2;
        for (int e = t; e < ncell; e += WG) {
            const int c = e / bl, r = e - c * bl;
            if (r != p) AUG[c * bl + r] = fma(-fac[r], AUG[c * bl + p], AUG[c * bl + r]);
        }
        __syncthreads();
    }
    return ok;
}

// ---- factor kernels ----------------------------------------------------------------
struct RowLists {
    const int *pos;       // packed block of each slot of A
    const int *rows;      // rows of this level
    const int *diag;      // diagonal slot per row
    const int *lptr, *lslot, *lcol, *ltrans;  // lower slots, their columns, transposed slot
    const int *uptr, *uslot, *ucol;           // upper slots and columns
};

// LDS: AUG (2 bl^2) + P (bl^2) + Q (bl^2) + fac (bl) + scratch
template <class T, int NE>
__global__ __launch_bounds__(WG) void k_dilu_factor(int bl, RowLists L, const double *__restrict__ A,
                                                    T *__restrict__ F, float *__restrict__ scale,
                                                    double *__restrict__ DD64, int *__restrict__ status) {
    extern __shared__ double lds[];
    const int bb = bl * bl;
    double *AUG = lds;
    double *P = AUG + 2 * bb;
    double *Q = P + bb;
    double *fac = Q + bb;
    double *scratch = fac + bl;
    __shared__ int piv_row;
    const int i = L.rows[blockIdx.x];
    const int d = L.diag[i];
    const int t = threadIdx.x;
    // AUG = [A_ii | I]
    for (int e = t; e < bb; e += WG) {
        AUG[e] = A[(size_t)d * bb + e];
        const int r = e % bl, c = e / bl;
        AUG[bb + e] = (r == c) ? 1.0 : 0.0;
    }
    double c[NE];
    for (int q = L.lptr[i]; q < L.lptr[i + 1]; q++) {
        const int s = L.lslot[q], k = L.lcol[q], st = L.ltrans[q];
        __syncthreads();
        load_block<NE>(bl, P, A + (size_t)s * bb);
        load_block<NE>(bl, Q, DD64 + (size_t)k * bb);
        __syncthreads();
        gemm_lds<NE>(bl, P, Q, c);  // L_ik = A_ik D_k^-1
        store_block<T, NE>(bl, F + (size_t)L.pos[s] * bb, scale ? scale + L.pos[s] : nullptr, c, scratch);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NE; j++) {
            const int e = t + WG * j;
            if (e < bb) P[e] = c[j];
        }
        load_block<NE>(bl, Q, A + (size_t)st * bb);
        __syncthreads();
        gemm_lds<NE>(bl, P, Q, c);  // L_ik A_ki
#pragma unroll
        for (int j = 0; j < NE; j++) {
            const int e = t + WG * j;
            if (e < bb) AUG[e] -= c[j];
        }
    }
    __syncthreads();
    const bool ok = gauss_jordan(bl, AUG, fac, &piv_row);
    if (!ok && t == 0) atomicOr(status, 1);
    // D_i^-1 = right half of AUG
    double *Dinv = AUG + bb;
#pragma unroll
    for (int j = 0; j < NE; j++) {
        const int e = t + WG * j;
        c[j] = e < bb ? Dinv[e] : 0.0;
        if (e < bb) DD64[(size_t)i * bb + e] = c[j];
    }
    store_block<T, NE>(bl, F + (size_t)L.pos[d] * bb, scale ? scale + L.pos[d] : nullptr, c, scratch);
    for (int q = L.uptr[i]; q < L.uptr[i + 1]; q++) {
        const int s = L.uslot[q];
        __syncthreads();
        load_block<NE>(bl, Q, A + (size_t)s * bb);
        __syncthreads();
        gemm_lds<NE>(bl, Dinv, Q, c);  // DU_ij = D_i^-1 A_ij
        store_block<T, NE>(bl, F + (size_t)L.pos[s] * bb, scale ? scale + L.pos[s] : nullptr, c, scratch);
    }
}

// block Jacobi: inverse of each diagonal block
template <class T, int NE>
__global__ __launch_bounds__(WG) void k_bjacobi_factor(int bl, const int *__restrict__ diag,
                                                       const double *__restrict__ A, T *__restrict__ F,
                                                       float *__restrict__ scale, int *__restrict__ status) {
    extern __shared__ double lds[];
    const int bb = bl * bl;
    double *AUG = lds;
    double *fac = AUG + 2 * bb;
    double *scratch = fac + bl;
    __shared__ int piv_row;
    const int i = blockIdx.x;
    const int d = diag[i];
    for (int e = threadIdx.x; e < bb; e += WG) {
        AUG[e] = A[(size_t)d * bb + e];
        AUG[bb + e] = ((e % bl) == (e / bl)) ? 1.0 : 0.0;
    }
    __syncthreads();
    const bool ok = gauss_jordan(bl, AUG, fac, &piv_row);
    if (!ok && threadIdx.x == 0) atomicOr(status, 1);
    double c[NE];
#pragma unroll
    for (int j = 0; j < NE; j++) {
        const int e = threadIdx.x + WG * j;
        c[j] = e < bb ? AUG[bb + e] : 0.0;
    }
    store_block<T, NE>(bl, F + (size_t)i * bb, scale ? scale + i : nullptr, c, scratch);
}

// ---- wavefront-per-row factor: FP64 MFMA block products, in-register Gauss-Jordan -------
//
// Per level, one 64-lane wavefront per block row forms the L blocks and the Schur
// complement, a second launch inverts the complements, and the DU blocks follow in one
// launch over all upper blocks.  Block products C = P Q (bl x bl, FP64) run on
// v_mfma_f64_16x16x4f64 over ceil(bl/16)^2 output tiles with operands loaded from L2
// straight into the MFMA registers (or from one LDS buffer holding the previous product);
// results pass through LDS so that they are written to HBM coalesced.  The diagonal block
// is inverted by Gauss-Jordan with partial pivoting with one matrix row per lane: the pivot
// search is a cross-lane argmax, the pivot row is broadcast with readlane, and rows are
// never swapped (the row/column order is restored when the inverse is written out).
typedef double dx4 __attribute__((ext_vector_type(4)));

template <int BL>
struct WF {
    static constexpr int LD = BL | 1;                                   // odd: conflict-free LDS columns
    static constexpr int TM = (BL + 15) / 16;                           // 16x16 tiles per dimension
    static constexpr int KS = (BL + 3) / 4;                             // k-steps of the 16x16x4 MFMA
    static constexpr int BB = BL * BL;
    static constexpr int VEC = BL % 4 == 0 ? 4 : BL % 2 == 0 ? 2 : 1;   // consecutive elements of one column
    static constexpr int NCH = (BB + 64 * VEC - 1) / (64 * VEC);        // copy chunks per lane
};

__device__ inline double readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// one DPP step of a 64-bit max: v = max(v, v of the lane selected by CTRL) on the rows in RM
template <int CTRL, int RM>
__device__ inline unsigned long long dpp_max_u64(unsigned long long v) {
    const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
    const int mlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, RM, 0xF, false);
    const int mhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, RM, 0xF, false);
    const unsigned long long m = ((unsigned long long)(unsigned)mhi << 32) | (unsigned)mlo;
    return m > v ? m : v;
}

// maximum over the 64 lanes (uniform result): quad swaps and half-row/row mirrors reduce each
// row of 16 lanes, row_bcast:15 and row_bcast:31 carry the row maxima into lane 63
__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
    v = dpp_max_u64<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v = dpp_max_u64<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v = dpp_max_u64<0x141, 0xF>(v);   // row_half_mirror
    v = dpp_max_u64<0x140, 0xF>(v);   // row_mirror
    v = dpp_max_u64<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v = dpp_max_u64<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// copy chunks: a block in HBM (column-major, contiguous) <-> registers <-> LDS (column stride LD)
template <int BL>
struct Chunks {
    double v[WF<BL>::NCH][WF<BL>::VEC];

    __device__ void load(const double *__restrict__ g) {
        using W = WF<BL>;
#pragma unroll
        for (int j = 0; j < W::NCH; j++) {
            const int e = (threadIdx.x + 64 * j) * W::VEC;
            if (e < W::BB) {
                if constexpr (W::VEC == 4) {
                    const double2 a = *reinterpret_cast<const double2 *>(g + e);
                    const double2 b = *reinterpret_cast<const double2 *>(g + e + 2);
                    v[j][0] = a.x; v[j][1] = a.y; v[j][2] = b.x; v[j][3] = b.y;
                } else if constexpr (W::VEC == 2) {
                    const double2 a = *reinterpret_cast<const double2 *>(g + e);
                    v[j][0] = a.x; v[j][1] = a.y;
                } else {
                    v[j][0] = g[e];
                }
            }
        }
    }
    __device__ void to_lds(double *X) const {
        using W = WF<BL>;
#pragma unroll
        for (int j = 0; j < W::NCH; j++) {
            const int e = (threadIdx.x + 64 * j) * W::VEC;
            if (e < W::BB) {
                const int c = e / BL, r = e - c * BL;
#pragma unroll
                for (int q = 0; q < W::VEC; q++) X[c * W::LD + r + q] = v[j][q];
            }
        }
    }
    __device__ void from_lds(const double *X) {
        using W = WF<BL>;
#pragma unroll
        for (int j = 0; j < W::NCH; j++) {
            const int e = (threadIdx.x + 64 * j) * W::VEC;
            if (e < W::BB) {
                const int c = e / BL, r = e - c * BL;
#pragma unroll
                for (int q = 0; q < W::VEC; q++) v[j][q] = X[c * W::LD + r + q];
            } else {
#pragma unroll
                for (int q = 0; q < W::VEC; q++) v[j][q] = 0.0;
            }
        }
    }
    // store in precision T (BF16: one power-of-two scale for the block, written to *scale_out)
    template <class T>
    __device__ void store(T *__restrict__ g, float *scale_out) const {
        using W = WF<BL>;
        float inv = 1.0f;
        if constexpr (Store<T>::prec == CPG_BF16) {
            double mx = 0.0;
#pragma unroll
            for (int j = 0; j < W::NCH; j++)
#pragma unroll
                for (int q = 0; q < W::VEC; q++) mx = fmax(mx, fabs(v[j][q]));
            const float s = pow2_scale(wave_max(mx));
            inv = 1.0f / s;
            if (threadIdx.x == 0) *scale_out = s;
        }
#pragma unroll
        for (int j = 0; j < W::NCH; j++) {
            const int e = (threadIdx.x + 64 * j) * W::VEC;
            if (e < W::BB) {
#pragma unroll
                for (int q = 0; q < W::VEC; q++) Store<T>::put(g, e + q, v[j][q], inv);
            }
        }
    }
};

template <int BL>
__device__ inline void tiles_zero(dx4 (&t)[WF<BL>::TM][WF<BL>::TM]) {
#pragma unroll
    for (int a = 0; a < WF<BL>::TM; a++)
#pragma unroll
        for (int b = 0; b < WF<BL>::TM; b++) t[a][b] = dx4{0.0, 0.0, 0.0, 0.0};
}

// tile element (tm, tn, i) of lane l sits at row 16 tm + l / 16 + 4 i, column 16 tn + l % 16
// (the v_mfma_f64_16x16x4f64 result layout)
template <int BL>
__device__ inline void tiles_load(dx4 (&t)[WF<BL>::TM][WF<BL>::TM], const double *__restrict__ g) {
    using W = WF<BL>;
    const int r0 = threadIdx.x >> 4, c0 = threadIdx.x & 15;
#pragma unroll
    for (int tm = 0; tm < W::TM; tm++)
#pragma unroll
        for (int tn = 0; tn < W::TM; tn++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 16 * tm + r0 + 4 * i, col = 16 * tn + c0;
                t[tm][tn][i] = (row < BL && col < BL) ? g[col * BL + row] : 0.0;
            }
}

template <int BL>
__device__ inline void tiles_to_lds(double *X, const dx4 (&t)[WF<BL>::TM][WF<BL>::TM]) {
    using W = WF<BL>;
    const int r0 = threadIdx.x >> 4, c0 = threadIdx.x & 15;
#pragma unroll
    for (int tm = 0; tm < W::TM; tm++)
#pragma unroll
        for (int tn = 0; tn < W::TM; tn++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 16 * tm + r0 + 4 * i, col = 16 * tn + c0;
                if (row < BL && col < BL) X[col * W::LD + row] = t[tm][tn][i];
            }
}

// Gauss-Jordan inverse of a bl x bl block held one row per lane (w[c] = entry (lane, c);
// lanes past bl hold zeros), partial pivoting on the first maximal |pivot| among rows not
// yet used.  Rows are never swapped and the pivot row is not rescaled during the
// elimination: a row keeps its stored values and a scale 1/pivot from the step it was
// chosen, so every lane applies the same update w -= (w_p / pivot) * w_pivotrow (zero for
// the pivot lane itself).  On return the inverse is A^-1[my_step][rho_q] = my_scale * w[q],
// with rho_q = my_rho of lane q.  Returns false (uniformly) on a zero pivot.
template <int BL>
__device__ inline bool wave_gj_core(double (&w)[BL], int &my_step, int &my_rho, double &my_scale) {
    const int lane = threadIdx.x & 63;
    bool used = lane >= BL;
    my_step = 0;     // step at which this lane's row was the pivot row
    my_rho = 0;      // lane p: pivot row of step p
    my_scale = 1.0;  // 1 / pivot of that step
    bool ok = true;
    // w is rotated by one column per step: at step p, w[0] holds column p and the finished
    // column goes to the end, so after bl steps the columns are back in their order and no
    // register is ever indexed by the loop variable
#pragma unroll 1
    for (int p = 0; p < BL; p++) {
        const double v = w[0];
        // first maximal |W[r][p]| over unused rows r; a NaN entry ranks below every number but
        // above the used rows (and the lanes past bl), so the pivot is always an unused row
        // key: used rows < NaN < numbers; among numbers larger |v| first (|v| compared to 46
        // mantissa bits), then the lower lane
        unsigned long long key = (unsigned long long)(63 - lane);
        if (!used) {
            if (v != v) key += 64;
            else key |= ((unsigned long long)__double_as_longlong(fabs(v)) & ~63ull) + 128;
        }
        const int rho = 63 - (int)(wave_max_u64(key) & 63);
        const double piv = readlane_d(v, rho);
        ok = ok && piv != 0.0;
        const double inv = piv != 0.0 ? 1.0 / piv : 1.0;
        const bool me = lane == rho;
        if (me) { used = true; my_step = p; my_scale = inv; }
        if (lane == p) my_rho = rho;
        const double f = me ? 0.0 : v * inv;
#pragma unroll
        for (int c = 1; c < BL; c++) w[c - 1] = fma(-f, readlane_d(w[c], rho), w[c]);
        w[BL - 1] = me ? 1.0 : -f;
    }
    return ok;
}

// In-place Gauss-Jordan inverse of the bl x bl block in X (column stride LD)
template <int BL>
__device__ inline bool wave_gauss_jordan(double *X) {
    using W = WF<BL>;
    const int lane = threadIdx.x;
    const bool row_lane = lane < BL;
    double w[BL];
#pragma unroll
    for (int c = 0; c < BL; c++) w[c] = row_lane ? X[c * W::LD + lane] : 0.0;
    int my_step, my_rho;
    double my_scale;
    const bool ok = wave_gj_core<BL>(w, my_step, my_rho, my_scale);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BL; q++) {
        const int col = __builtin_amdgcn_readlane(my_rho, q);
        if (row_lane) X[col * W::LD + my_step] = w[q] * my_scale;
    }
    __syncthreads();
    return ok;
}

// ---- blocked Gauss-Jordan inverse on FP64 MFMA (bl = 120) ------------------------------
//
// Gauss-Jordan in exchange form with partial pivoting: step k takes the pivot row p among the
// rows not yet pivotal (largest |X[p][k]|, the lower row on ties) and sets, with v = X[p][k],
//     X[p][k] = 1/v,  X[p][c] = -X[p][c]/v,  X[r][k] = X[r][k]/v,  X[r][c] -= X[r][k] X[p][c]/v.
// Rows are never swapped.  Here the steps are organised by panels of GJP columns.  The block
// lives in MFMA accumulator tiles (the v_mfma_f64_16x16x4f64 result layout) over the padded
// NP x NP matrix: wavefront w of the workgroup holds the tile columns w and w + 4.  Per panel:
//  1. the wavefront holding the panel's columns copies them to LDS and runs the GJP exchange
//     steps on them alone, two matrix rows per lane (rows l and l + 64): the pivot search is a
//     wave-wide max, the pivot row comes by readlane, no barrier inside;
//  2. the steps' effect on every other column c is
//         X[:, c] <- X[:, c] - E_P X[P, c] - Q X[P, c],
//     with P the panel's pivot rows, E_P their identity columns and Q the panel's columns after
//     its steps: the rows P are cleared and a rank-GJP product is subtracted, which all four
//     wavefronts do on MFMA with -Q and the previous rows P staged in LDS.
// The arithmetic is that of the unblocked elimination except for the order in which each
// element's rank-1 updates are summed.  The inverse is A^-1[s(r)][p(c)] = X[r][c] (s(r) the step
// at which row r was the pivot row, p(c) the pivot row of step c), written out through LDS in
// chunks of OC destination columns so that the global stores are contiguous.
constexpr int GJP = 8;

template <int BL>
struct MG {
    static constexpr int NP = (BL + 15) / 16 * 16;   // padded size
    static constexpr int TM = NP / 16;               // tile rows and tile columns
    static constexpr int PLD = GJP + 1;              // panel row stride
    static constexpr int BLD = NP + 16;              // pivot-row stride: the four k rows of one
                                                     // operand read fall on both bank halves
    static constexpr int OC = 32;                    // destination columns per output chunk
    static constexpr int OLD = BL + 1;
    static constexpr int PB = NP * PLD + GJP * BLD;
    static constexpr int STAGE = PB > OC * OLD ? PB : OC * OLD;
};

// pivot key: used rows < NaN < numbers; among numbers the larger |v| compared to its exponent and
// leading 13 mantissa bits (a candidate within a factor 1 - 2^-13 of the largest may be chosen
// over it), then the lower row (r < 128)
__device__ inline unsigned gj_key(double v, bool used, int r) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v) & 0x7fffffffffffffffull;
    const unsigned tag = 127u - (unsigned)r;
    const unsigned num = ((unsigned)(bits >> 39) + 2u) << 7;   // |v| < 2^1024: bits >> 39 < 2^24
    const unsigned k = v != v ? 128u : num;
    return (used ? 0u : k) | tag;
}

// one DPP step of a 32-bit max over the rows in RM
template <int CTRL, int RM>
__device__ inline unsigned dpp_max_u32(unsigned v) {
    const unsigned m = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, RM, 0xF, false);
    return m > v ? m : v;
}

// maximum over the 64 lanes (uniform result), same lane pattern as wave_max_u64
__device__ inline unsigned wave_max_u32(unsigned v) {
    v = dpp_max_u32<0xB1, 0xF>(v);
    v = dpp_max_u32<0x4E, 0xF>(v);
    v = dpp_max_u32<0x141, 0xF>(v);
    v = dpp_max_u32<0x140, 0xF>(v);
    v = dpp_max_u32<0x142, 0xA>(v);
    v = dpp_max_u32<0x143, 0xC>(v);
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// inverse of the block at a (column-major, FP64) to outT in precision T (BF16: its scale to
// *scale_out) and, with OUT64, in FP64 to out64 (which may equal a: the block is read into
// registers before the first barrier); whole workgroup of 4 wavefronts
template <class T, int BL, bool OUT64>
__device__ inline void mfma_gauss_jordan(const double *a, double *out64, T *__restrict__ outT,
                                         float *scale_out, int *__restrict__ status) {
    using G = MG<BL>;
    constexpr int NP = G::NP, TM = G::TM, PLD = G::PLD, BLD = G::BLD, OC = G::OC, OLD = G::OLD;
    static_assert(TM <= 8 && NP <= 128 && BL % GJP == 0 && GJP % 4 == 0, "blocked inverse tiling");
    __shared__ double stage[G::STAGE];
    __shared__ int rowstep[NP];   // step at which row r was the pivot row, -1 before
    __shared__ int pr[NP];        // pivot row of step k
    __shared__ double red[WG / 64];
    double *Pan = stage, *Brow = stage + NP * PLD;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    dx4 acc[TM][2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int col = 16 * (wv + 4 * h) + lr;
#pragma unroll
        for (int tm = 0; tm < TM; tm++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 16 * tm + lq + 4 * i;
                acc[tm][h][i] = (row < BL && col < BL) ? a[(size_t)col * BL + row] : 0.0;
            }
    }
    if (tid < NP) rowstep[tid] = -1;
    bool ok = true;   // false in the wavefront that met a zero or non-finite pivot
#pragma unroll 1
    for (int k0 = 0; k0 < BL; k0 += GJP) {
        const int t = k0 >> 4, hf = (k0 >> 3) & 1;   // tile column of the panel, and its half
        const int owner = t & 3, oh = t >> 2;
        __syncthreads();   // the previous update is done with Pan and Brow
        if (wv == owner && (lr >> 3) == hf) {
            const int c = lr & 7;
            if (oh) {
#pragma unroll
                for (int tm = 0; tm < TM; tm++)
#pragma unroll
                    for (int i = 0; i < 4; i++) Pan[(16 * tm + lq + 4 * i) * PLD + c] = acc[tm][1][i];
            } else {
#pragma unroll
                for (int tm = 0; tm < TM; tm++)
#pragma unroll
                    for (int i = 0; i < 4; i++) Pan[(16 * tm + lq + 4 * i) * PLD + c] = acc[tm][0][i];
            }
        }
        if (wv == owner) {
            // the owner's own writes above, complete before its reads below
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int ra = lane, rb = lane + 64;
            double xa[GJP], xb[GJP];
#pragma unroll
            for (int c = 0; c < GJP; c++) {
                xa[c] = Pan[ra * PLD + c];
                xb[c] = rb < NP ? Pan[rb * PLD + c] : 0.0;
            }
            bool ua = ra >= BL || rowstep[ra] >= 0;
            bool ub = rb >= BL || rowstep[rb] >= 0;
#pragma unroll
            for (int j = 0; j < GJP; j++) {
                const unsigned ka = gj_key(xa[j], ua, ra), kb = gj_key(xb[j], ub, rb);
                const unsigned K = wave_max_u32(ka > kb ? ka : kb);
                const int rho = 127 - (int)(K & 127u);
                const bool hi = rho >= 64;
                const int rl = rho & 63;
                double q[GJP];
                if (hi) {
#pragma unroll
                    for (int c = 0; c < GJP; c++) q[c] = readlane_d(xb[c], rl);
                } else {
#pragma unroll
                    for (int c = 0; c < GJP; c++) q[c] = readlane_d(xa[c], rl);
                }
                const double v = q[j];
                const bool good = K >= 256u && v != 0.0 && isfinite(v);
                ok = ok && good;
                const double inv = good ? 1.0 / v : 1.0;
#pragma unroll
                for (int c = 0; c < GJP; c++) q[c] *= inv;
                const double fa = xa[j], fb = xb[j];
#pragma unroll
                for (int c = 0; c < GJP; c++) {
                    xa[c] = fma(-fa, q[c], xa[c]);
                    xb[c] = fma(-fb, q[c], xb[c]);
                }
                xa[j] = fa * inv;
                xb[j] = fb * inv;
                // the pivot row: -q, and 1 / v in column j
                if (lane == rl) {
                    if (hi) {
#pragma unroll
                        for (int c = 0; c < GJP; c++) xb[c] = c == j ? inv : -q[c];
                        ub = true;
                    } else {
#pragma unroll
                        for (int c = 0; c < GJP; c++) xa[c] = c == j ? inv : -q[c];
                        ua = true;
                    }
                }
                if (lane == 0) {
                    rowstep[rho] = k0 + j;
                    pr[k0 + j] = rho;
                }
            }
#pragma unroll
            for (int c = 0; c < GJP; c++) {
                Pan[ra * PLD + c] = xa[c];
                if (rb < NP) Pan[rb * PLD + c] = xb[c];
            }
        }
        __syncthreads();
        // panel columns take Q; elsewhere the rows P go to Brow (in step order) and are cleared;
        // Brow is zero in the panel columns, so the product below leaves them as they are.
        // Row 16 tm + lq + 4 i of a lane is its slot 4 tm + i; the slots holding pivot rows of
        // this panel are found from the (uniform) pivot rows, not by a lookup per slot.
        {
            unsigned smask = 0;                   // slots of the panel's pivot rows (uniform)
            unsigned mine = 0;                    // ... held by this lane (lq matches)
            unsigned tj[4] = {0u, 0u, 0u, 0u};    // step of each of my slots, 4 bits a slot
#pragma unroll
            for (int j = 0; j < GJP; j++) {
                const int r = __builtin_amdgcn_readfirstlane(pr[k0 + j]);
                const int sl = (r >> 4) * 4 + ((r >> 2) & 3);
                smask |= 1u << sl;
                const bool me = (r & 3) == lq;
                mine |= me ? 1u << sl : 0u;
#pragma unroll
                for (int w = 0; w < 4; w++) tj[w] |= (me && (sl >> 3) == w) ? (unsigned)j << (4 * (sl & 7)) : 0u;
            }
            smask = __builtin_amdgcn_readfirstlane(smask);
#pragma unroll
            for (int tm = 0; tm < TM; tm++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int sl = 4 * tm + i;
                    if ((smask >> sl) & 1u) {
                        if ((mine >> sl) & 1u) {
                            const int js = (int)((tj[sl >> 3] >> (4 * (sl & 7))) & 15u);
#pragma unroll
                            for (int h = 0; h < 2; h++) {
                                const int col = 16 * (wv + 4 * h) + lr;
                                const bool pc = (col >> 3) == (k0 >> 3);
                                Brow[js * BLD + col] = pc ? 0.0 : acc[tm][h][i];
                                acc[tm][h][i] = 0.0;
                            }
                        }
                    }
                }
            if (wv == owner) {
                const int c = lr & 7;
                if ((lr >> 3) == hf) {
                    if (oh) {
#pragma unroll
                        for (int tm = 0; tm < TM; tm++)
#pragma unroll
                            for (int i = 0; i < 4; i++) acc[tm][1][i] = Pan[(16 * tm + lq + 4 * i) * PLD + c];
                    } else {
#pragma unroll
                        for (int tm = 0; tm < TM; tm++)
#pragma unroll
                            for (int i = 0; i < 4; i++) acc[tm][0][i] = Pan[(16 * tm + lq + 4 * i) * PLD + c];
                    }
                }
            }
        }
        __syncthreads();
        // X -= Q Brow
#pragma unroll
        for (int ks = 0; ks < GJP / 4; ks++) {
            const int k = 4 * ks + lq;
            const double b0 = Brow[k * BLD + 16 * wv + lr];
            const double b1 = Brow[k * BLD + 16 * (wv + 4) + lr];
#pragma unroll
            for (int tm = 0; tm < TM; tm++) {
                const double am = -Pan[(16 * tm + lr) * PLD + k];
                acc[tm][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(am, b0, acc[tm][0], 0, 0, 0);
                acc[tm][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(am, b1, acc[tm][1], 0, 0, 0);
            }
        }
    }
    __syncthreads();
    float inv = 1.0f;
    if constexpr (Store<T>::prec == CPG_BF16) {
        double mx = 0.0;
#pragma unroll
        for (int tm = 0; tm < TM; tm++)
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int i = 0; i < 4; i++) mx = fmax(mx, fabs(acc[tm][h][i]));
        mx = wg_max_abs(mx, red);
        const float sc = pow2_scale(mx);
        inv = 1.0f / sc;
        if (tid == 0) *scale_out = sc;
    }
    double *Out = stage;
#pragma unroll 1
    for (int d0 = 0; d0 < BL; d0 += OC) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int col = 16 * (wv + 4 * h) + lr;
            const int d = col < BL ? pr[col] - d0 : -1;
            if (d >= 0 && d < OC) {
#pragma unroll
                for (int tm = 0; tm < TM; tm++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int row = 16 * tm + lq + 4 * i;
                        if (row < BL) Out[d * OLD + rowstep[row]] = acc[tm][h][i];
                    }
            }
        }
        __syncthreads();
        const int ne = min(OC, BL - d0) * BL;
        for (int e = tid; e < ne; e += WG) {
            const int dc = e / BL, r = e - dc * BL;
            const double v = Out[dc * OLD + r];
            const size_t g = (size_t)d0 * BL + e;
            Store<T>::put(outT, g, v, inv);
            if constexpr (OUT64) out64[g] = v;
        }
        __syncthreads();
    }
    if (!ok && lane == 0) atomicOr(status, 1);
}

template <class T, int BL>
__global__ __launch_bounds__(WG, 2) void k_bjacobi_factor_rt(const int *__restrict__ diag, const double *__restrict__ A,
                                                             T *__restrict__ F, float *__restrict__ scale,
                                                             int *__restrict__ status) {
    constexpr size_t bb = (size_t)BL * BL;
    const int i = blockIdx.x;
    mfma_gauss_jordan<T, BL, false>(A + (size_t)diag[i] * bb, nullptr, F + (size_t)i * bb,
                                    scale ? scale + i : nullptr, status);
}

// ---- DILU factor, one level: two kernels per level, then one launch for all DU blocks.
// Schur kernel (one workgroup of four wavefronts per row i of the level):
//   L_ik = A_ik D_k^-1 (stored in the factor precision),  S_i = A_ii - sum_k L_ik A_ki,
// with S_i written in FP64 to DD64[i].  Inverse kernel: D_i^-1 = S_i^-1 by the in-register
// Gauss-Jordan, written over DD64[i] (FP64, used by the later levels and the DU blocks) and
// to the factor array.  Splitting keeps each kernel's registers low enough for several
// wavefronts per SIMD, so the latency-bound elimination of one row overlaps the MFMA work
// and memory traffic of others.
// workgroup of 4 wavefronts: the ceil(bl/16)^2 output tiles of a block product are dealt
// round-robin to the waves; blocks are staged in LDS by all 256 threads with 16-byte loads
constexpr int FWG = 256;

template <int BL>
struct FT {
    static constexpr int NT = WF<BL>::TM * WF<BL>::TM;   // output tiles
    static constexpr int NJ = (NT + 3) / 4;                // tiles per wave
    static constexpr int V = BL % 2 == 0 ? 2 : 1;          // doubles per load
};

// use(e, load(e)) for every e < N of this thread (e = threadIdx.x + FWG j, all threads of the
// workgroup), the loads issued B at a time ahead of their uses so that B global loads per thread
// are in flight together (a loop that loads and then uses one value waits out each load's
// latency in turn); the rounds stay a loop so that no more than B loaded values are live; loads
// past the end of the last round read index N - 1 and are not used
template <int N, int B = 16, class Load, class Use>
__device__ inline void wg_batched(Load &&load, Use &&use) {
    constexpr int IT = (N + FWG - 1) / FWG;
    constexpr int BB = IT < B ? IT : B;
    using V = decltype(load(0));
    const int t = threadIdx.x;
#pragma unroll 1
    for (int j0 = 0; j0 < IT; j0 += BB) {
        V v[BB];
#pragma unroll
        for (int j = 0; j < BB; j++) v[j] = load(min(t + FWG * (j0 + j), N - 1));
#pragma unroll
        for (int j = 0; j < BB; j++) {
            const int e = t + FWG * (j0 + j);
            if (e < N) use(e, v[j]);
        }
    }
}

// X (column stride LD) <- the bl x bl block at g, all threads of the workgroup
template <int BL>
__device__ inline void wg_block_to_lds(double *X, const double *__restrict__ g) {
    using W = WF<BL>;
    if constexpr (FT<BL>::V == 2) {
        wg_batched<W::BB / 2>([&](int u) { return *reinterpret_cast<const double2 *>(g + 2 * u); },
                              [&](int u, double2 a) {
                                  const int c = (2 * u) / BL, r = 2 * u - c * BL;
                                  X[c * W::LD + r] = a.x;
                                  X[c * W::LD + r + 1] = a.y;
                              });
    } else {
        wg_batched<W::BB>([&](int e) { return g[e]; },
                          [&](int e, double a) {
                              const int c = e / BL, r = e - c * BL;
                              X[c * W::LD + r] = a;
                          });
    }
}

// the block in X (FP64, or FP32 for the coupled factor's FP32 products) to global memory in
// precision T, all threads (BF16: one power-of-two scale for the block from a workgroup max,
// written to *scale_out)
template <class T, int BL, class XT>
__device__ inline void wg_lds_block_store(const XT *X, T *__restrict__ g, float *scale_out, double *red) {
    using W = WF<BL>;
    float inv = 1.0f;
    if constexpr (Store<T>::prec == CPG_BF16) {
        double mx = 0.0;
        for (int e = threadIdx.x; e < W::BB; e += FWG) {
            const int c = e / BL, r = e - c * BL;
            mx = fmax(mx, fabs((double)X[c * W::LD + r]));
        }
        mx = wave_max(mx);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
        __syncthreads();
        mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        const float sc = pow2_scale(mx);
        inv = 1.0f / sc;
        if (threadIdx.x == 0) *scale_out = sc;
    }
    for (int e = threadIdx.x; e < W::BB; e += FWG) {
        const int c = e / BL, r = e - c * BL;
        Store<T>::put(g, e, (double)X[c * W::LD + r], inv);
    }
}

// acc (+)= (NEG ? -P : P) Q for output tile (tm, tn); P, Q in LDS (column stride LD)
template <int BL, bool NEG>
__device__ inline void mma_tile_lds(const double *P, const double *Q, int tm, int tn, dx4 &acc) {
    using W = WF<BL>;
    const int lr = threadIdx.x & 15, lk = (threadIdx.x & 63) >> 4;
    const int m = 16 * tm + lr, n = 16 * tn + lr;
#pragma unroll
    for (int ks = 0; ks < W::KS; ks++) {
        const int k = 4 * ks + lk;
        const bool ka = k < BL && m < BL, kb = k < BL && n < BL;
        const double pa = P[ka ? k * W::LD + m : 0];
        const double a = ka ? (NEG ? -pa : pa) : 0.0;
        const double b = kb ? Q[n * W::LD + k] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
}

// tile (tm, tn) of this wave's accumulator to X (column stride LD)
template <int BL>
__device__ inline void tile_to_lds(double *X, int tm, int tn, const dx4 &t) {
    using W = WF<BL>;
    const int l = threadIdx.x & 63;
    const int r0 = l >> 4, c0 = l & 15;
    const int col = 16 * tn + c0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int row = 16 * tm + r0 + 4 * e;
        if (row < BL && col < BL) X[col * W::LD + row] = t[e];
    }
}

template <int BL>
__device__ inline void tile_load(dx4 &t, int tm, int tn, const double *__restrict__ g) {
    const int l = threadIdx.x & 63;
    const int r0 = l >> 4, c0 = l & 15;
    const int col = 16 * tn + c0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int row = 16 * tm + r0 + 4 * e;
        t[e] = (row < BL && col < BL) ? g[col * BL + row] : 0.0;
    }
}

template <class T, int BL>
__global__ __launch_bounds__(FWG) void k_dilu_schur(RowLists L, const double *__restrict__ A, T *__restrict__ F,
                                                    float *__restrict__ scale, double *__restrict__ DD64) {
    using W = WF<BL>;
    using U = FT<BL>;
    __shared__ double X[BL * W::LD];
    __shared__ double Y[BL * W::LD];
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6;
    const int i = L.rows[blockIdx.x];
    const int d = L.diag[i];
    constexpr size_t bb = (size_t)W::BB;
    dx4 acc[U::NJ];
#pragma unroll
    for (int j = 0; j < U::NJ; j++) {
        const int tt = wv + 4 * j;
        if (tt < U::NT) tile_load<BL>(acc[j], tt / W::TM, tt % W::TM, A + (size_t)d * bb);
        else acc[j] = dx4{0.0, 0.0, 0.0, 0.0};
    }
    for (int q = L.lptr[i]; q < L.lptr[i + 1]; q++) {
        const int s = L.lslot[q], k = L.lcol[q], st = L.ltrans[q];
        wg_block_to_lds<BL>(X, A + (size_t)s * bb);
        wg_block_to_lds<BL>(Y, DD64 + (size_t)k * bb);
        __syncthreads();
        dx4 t[U::NJ];
#pragma unroll
        for (int j = 0; j < U::NJ; j++) {
            t[j] = dx4{0.0, 0.0, 0.0, 0.0};
            const int tt = wv + 4 * j;
            if (tt < U::NT) mma_tile_lds<BL, false>(X, Y, tt / W::TM, tt % W::TM, t[j]);   // L_ik = A_ik D_k^-1
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < U::NJ; j++) {
            const int tt = wv + 4 * j;
            if (tt < U::NT) tile_to_lds<BL>(X, tt / W::TM, tt % W::TM, t[j]);
        }
        wg_block_to_lds<BL>(Y, A + (size_t)st * bb);
        __syncthreads();
        wg_lds_block_store<T, BL>(X, F + (size_t)L.pos[s] * bb, scale ? scale + L.pos[s] : nullptr, red);
#pragma unroll
        for (int j = 0; j < U::NJ; j++) {
            const int tt = wv + 4 * j;
            if (tt < U::NT) mma_tile_lds<BL, true>(X, Y, tt / W::TM, tt % W::TM, acc[j]);   // S_i -= L_ik A_ki
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < U::NJ; j++) {
        const int tt = wv + 4 * j;
        if (tt < U::NT) tile_to_lds<BL>(X, tt / W::TM, tt % W::TM, acc[j]);
    }
    __syncthreads();
    wg_lds_block_store<double, BL>(X, DD64 + (size_t)i * bb, nullptr, red);
}

template <class T, int BL>
__global__ __launch_bounds__(64) void k_dilu_inv_wf(const int *__restrict__ rows, const int *__restrict__ diag,
                                                    const int *__restrict__ pos, double *__restrict__ DD64,
                                                    T *__restrict__ F, float *__restrict__ scale,
                                                    int *__restrict__ status) {
    constexpr size_t bb = (size_t)BL * BL;
    const int lane = threadIdx.x;
    const int i = rows[blockIdx.x];
    const int pd = pos[diag[i]];
    double *S = DD64 + (size_t)i * bb;
    double w[BL];
#pragma unroll
    for (int c = 0; c < BL; c++) w[c] = lane < BL ? S[c * BL + lane] : 0.0;
    int my_step, my_rho;
    double my_scale;
    const bool ok = wave_gj_core<BL>(w, my_step, my_rho, my_scale);
    if (!ok && lane == 0) atomicOr(status, 1);
    T *Fd = F + (size_t)pd * bb;
    float inv = 1.0f;
    if constexpr (Store<T>::prec == CPG_BF16) {
        double mx = 0.0;
#pragma unroll
        for (int q = 0; q < BL; q++) mx = fmax(mx, lane < BL ? fabs(w[q] * my_scale) : 0.0);
        const float sc = pow2_scale(wave_max(mx));
        inv = 1.0f / sc;
        if (lane == 0) scale[pd] = sc;
    }
    // D^-1[my_step][rho_q] = my_scale * w[q]; every lane has read its row of S before any write
#pragma unroll
    for (int q = 0; q < BL; q++) {
        const int col = __builtin_amdgcn_readlane(my_rho, q);
        if (lane < BL) {
            const double v = w[q] * my_scale;
            const size_t e = (size_t)col * BL + my_step;
            S[e] = v;
            Store<T>::put(Fd, e, v, inv);
        }
    }
}

// DU_ij = D_i^-1 A_ij for the upper blocks of row urows[w], one workgroup per row, after all
// levels (D_i^-1 staged once per row)
template <class T, int BL>
__global__ __launch_bounds__(FWG) void k_dilu_upper(const int *__restrict__ urows, const int *__restrict__ uptr,
                                                    const int *__restrict__ uslot, const int *__restrict__ pos,
                                                    const double *__restrict__ A, const double *__restrict__ DD64,
                                                    T *__restrict__ F, float *__restrict__ scale) {
    using W = WF<BL>;
    using U = FT<BL>;
    __shared__ double X[BL * W::LD];
    __shared__ double Y[BL * W::LD];
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6;
    const int i = urows[blockIdx.x];
    constexpr size_t bb = (size_t)W::BB;
    wg_block_to_lds<BL>(X, DD64 + (size_t)i * bb);
    for (int q = uptr[i]; q < uptr[i + 1]; q++) {
        const int s = uslot[q];
        wg_block_to_lds<BL>(Y, A + (size_t)s * bb);
        __syncthreads();
        dx4 t[U::NJ];
#pragma unroll
        for (int j = 0; j < U::NJ; j++) {
            t[j] = dx4{0.0, 0.0, 0.0, 0.0};
            const int tt = wv + 4 * j;
            if (tt < U::NT) mma_tile_lds<BL, false>(X, Y, tt / W::TM, tt % W::TM, t[j]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < U::NJ; j++) {
            const int tt = wv + 4 * j;
            if (tt < U::NT) tile_to_lds<BL>(Y, tt / W::TM, tt % W::TM, t[j]);
        }
        __syncthreads();
        wg_lds_block_store<T, BL>(Y, F + (size_t)pos[s] * bb, scale ? scale + pos[s] : nullptr, red);
        __syncthreads();
    }
}

template <class T, int BL>
__global__ __launch_bounds__(64) void k_bjacobi_factor_wf(const int *__restrict__ diag, const double *__restrict__ A,
                                                          T *__restrict__ F, float *__restrict__ scale,
                                                          int *__restrict__ status) {
    using W = WF<BL>;
    __shared__ double X[BL * W::LD];
    const int i = blockIdx.x;
    const int d = diag[i];
    constexpr size_t bb = (size_t)W::BB;
    Chunks<BL> c;
    c.load(A + (size_t)d * bb);
    c.to_lds(X);
    __syncthreads();
    const bool ok = wave_gauss_jordan<BL>(X);
    if (!ok && threadIdx.x == 0) atomicOr(status, 1);
    c.from_lds(X);
    c.template store<T>(F + (size_t)i * bb, scale ? scale + i : nullptr);
}

// factors to FP64 in A's slot order (for download): out block s = packed block pos[s]
template <class T>
__global__ void k_to_fp64(const T *__restrict__ F, const float *__restrict__ scale, const int *__restrict__ pos,
                          size_t bb, size_t n, double *__restrict__ out) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const size_t s = e / bb;
    const size_t p = pos ? (size_t)pos[s] : s;
    double v = (double)Store<T>::get(F, p * bb + (e - s * bb));
    if (scale) v *= (double)scale[p];
    out[e] = v;
}

// ---- inner FP32 GMRES (variable preconditioner of the FP64 flexible GMRES) -------------
// Vectors and the matrix copy are FP32; dot products and norms accumulate in FP64; the
// Hessenberg matrix, Givens rotations and the small triangular solve live on the device
// in FP64, so an inner cycle is a fixed launch sequence with no host synchronisation.
constexpr int IWG = 256;
constexpr int IPT = 8;
constexpr int ICH = IWG * IPT;

__global__ void k_d2f(const double *__restrict__ a, float *__restrict__ b, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = (float)a[e];
}

__global__ void k_f2d(const float *__restrict__ a, double *__restrict__ b, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = (double)a[e];
}

// partial[b*nv + j] = sum over chunk b of Q_j . w   (FP64 accumulation)
__global__ __launch_bounds__(IWG) void k_f32_mdot(const float *__restrict__ Q, size_t N, int nv,
                                                  const float *__restrict__ w, double *__restrict__ partial) {
    __shared__ double red[IWG / 64][64];
    const size_t base = (size_t)blockIdx.x * ICH + threadIdx.x;
    float wv[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const size_t e = base + (size_t)k * IWG;
        wv[k] = e < N ? w[e] : 0.f;
    }
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j0 = 0; j0 < nv; j0 += 64) {
        const int jn = min(64, nv - j0);
        for (int jj = 0; jj < jn; jj++) {
            const float *q = Q + (size_t)(j0 + jj) * N;
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < IPT; k++) {
                const size_t e = base + (size_t)k * IWG;
                if (e < N) s = fma((double)q[e], (double)wv[k], s);
            }
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) red[wid][jj] = s;
        }
        __syncthreads();
        if (threadIdx.x < jn) {
            double s = 0.0;
            for (int q = 0; q < IWG / 64; q++) s += red[q][threadIdx.x];
            partial[(size_t)blockIdx.x * nv + j0 + threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// out[j] = sum_b partial[b*nv + j] (sqrt if requested)
__global__ __launch_bounds__(IWG) void k_f32_reduce(const double *__restrict__ partial, int nb, int nv,
                                                    double *__restrict__ out, int sqrt_out) {
    __shared__ double red[IWG / 64];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += IWG) s += partial[(size_t)b * nv + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < IWG / 64; q++) t += red[q];
        out[j] = sqrt_out ? sqrt(t) : t;
    }
}

// w -= sum_j h[j] Q_j ; partial[b] = sum over chunk b of w^2
__global__ __launch_bounds__(IWG) void k_f32_orth(const float *__restrict__ Q, size_t N, int nv,
                                                  const double *__restrict__ h, float *__restrict__ w,
                                                  double *__restrict__ partial) {
    __shared__ double red[IWG / 64];
    const size_t base = (size_t)blockIdx.x * ICH + threadIdx.x;
    double wv[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const size_t e = base + (size_t)k * IWG;
        wv[k] = e < N ? (double)w[e] : 0.0;
    }
    for (int j = 0; j < nv; j++) {
        const double hj = h[j];
        const float *q = Q + (size_t)j * N;
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            const size_t e = base + (size_t)k * IWG;
            if (e < N) wv[k] = fma(-hj, (double)q[e], wv[k]);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const size_t e = base + (size_t)k * IWG;
        if (e < N) {
            const float f = (float)wv[k];
            w[e] = f;
            s = fma((double)f, (double)f, s);
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// q = float(x64) and partial sums of q^2
__global__ __launch_bounds__(IWG) void k_f32_load_norm(const double *__restrict__ x, float *__restrict__ q, size_t N,
                                                       double *__restrict__ partial) {
    __shared__ double red[IWG / 64];
    const size_t base = (size_t)blockIdx.x * ICH + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const size_t e = base + (size_t)k * IWG;
        if (e < N) {
            const float f = (float)x[e];
            q[e] = f;
            s = fma((double)f, (double)f, s);
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// q = w / norm (0 if norm is 0)
__global__ void k_f32_scale(const float *__restrict__ w, float *__restrict__ q, size_t N,
                            const double *__restrict__ norm) {
    const double nv = *norm;
    const float inv = nv > 0.0 ? (float)(1.0 / nv) : 0.f;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N) q[e] = w[e] * inv;
}

// g = (beta, 0, ..., 0)
__global__ void k_inner_init(double *__restrict__ g, int k, const double *__restrict__ beta) {
    for (int j = threadIdx.x; j <= k; j += blockDim.x) g[j] = j == 0 ? *beta : 0.0;
}

// Column i of the Hessenberg matrix (ld = k+1, column-major; H(i+1,i) = hnew): apply the
// previous rotations, form rotation i (c = rd/dd, s = -hnew/dd as the reference) and
// rotate the right-hand side g.
__global__ void k_inner_givens(double *__restrict__ H, int ld, int i, const double *__restrict__ hnew,
                               double *__restrict__ cs, double *__restrict__ sn, double *__restrict__ g) {
    if (threadIdx.x != 0) return;
    double *col = H + (size_t)i * ld;
    for (int k = 0; k < i; k++) {
        const double tmp = cs[k] * col[k] - sn[k] * col[k + 1];
        col[k + 1] = sn[k] * col[k] + cs[k] * col[k + 1];
        col[k] = tmp;
    }
    const double rd = col[i], hn = *hnew;
    const double dd = sqrt(rd * rd + hn * hn);
    if (dd > 0.0) {
        cs[i] = rd / dd;
        sn[i] = -hn / dd;
    } else {
        cs[i] = 1.0;
        sn[i] = 0.0;
    }
    col[i] = dd;
    g[i + 1] = sn[i] * g[i];
    g[i] = cs[i] * g[i];
}

// y = R^-1 g for the leading k x k upper triangle R of H (zero where R_jj vanishes)
__global__ void k_inner_trisolve(const double *__restrict__ H, int ld, int k, const double *__restrict__ g,
                                 double *__restrict__ y) {
    if (threadIdx.x != 0) return;
    for (int c = k - 1; c >= 0; c--) {
        double s = g[c];
        for (int j = c + 1; j < k; j++) s -= H[(size_t)j * ld + c] * y[j];
        const double d = H[(size_t)c * ld + c];
        y[c] = d != 0.0 ? s / d : 0.0;
    }
}

// x = sum_j y_j Q_j
__global__ __launch_bounds__(IWG) void k_f32_combine(const float *__restrict__ Q, size_t N, int nk,
                                                     const double *__restrict__ y, float *__restrict__ x) {
    const size_t base = (size_t)blockIdx.x * ICH + threadIdx.x;
    double xv[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) xv[k] = 0.0;
    for (int j = 0; j < nk; j++) {
        const double yj = y[j];
        const float *q = Q + (size_t)j * N;
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            const size_t e = base + (size_t)k * IWG;
            if (e < N) xv[k] = fma(yj, (double)q[e], xv[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const size_t e = base + (size_t)k * IWG;
        if (e < N) x[e] = (float)xv[k];
    }
}

// ---- block DILU of the stage-coupled Radau IIA system (cell-interleaved 120 x 120 blocks) ---
// The coupled block of slot q (block row i, block column j) is never stored:
//   C_q[(s,r),(t,c)] = coef[s][t] J_s[q](r,c)  (+ m_i(r mod bs, c mod bs) when i = j, s = t and
//   r, c lie in the same component),
// built from the three stage Jacobians (40 x 40 column-major blocks of one pattern) and the
// per-cell mass blocks.  The factor is the DILU above on these blocks, on 16x16x4 MFMA (8 x 8
// tiles of 16 over the padded 128 x 128 block; wavefront w of the workgroup owns tile columns w
// and w + 4, 16 tiles): FP64 products for FP64 factors, FP32 products for FP32 and BF16 factors
// (k_cdilu_schur32 / k_cdilu_upper32 below), the inverses in FP64 for every precision:
//   Schur (one workgroup per row i of a level): per lower neighbour k: X <- (coef x I) D_k^-1,
//     T = BD(J_ik) X (registers), X <- T (stored as L_ik), accY += X BD(J_ki) (a second set of
//     accumulators); then S_i = C_ii - accY (coef x I) to DD64[i] element by element;
//   inverse: D_i^-1 = S_i^-1 by the blocked Gauss-Jordan on MFMA, to DD64[i] and the factors;
//   upper (after all levels): DU_ij = (D_i^-1 BD(J_ij)) (coef x I).
// The block-diagonal products take 180 (BD X) and at most 320 (X BD) MFMAs per wavefront
// against 480 for a dense 120 x 120 product.
// LDS: X (120 x 121 doubles) and the three stage blocks of one slot (3 x 40 x 41): 152 KB.
struct CoupledSrc {
    const double *J0, *J1, *J2;
    const double *mass;
    double coef[9];   // coef[3 s + t]
    int bs;
};
constexpr int CBL = 120, CSB = 40, CLD = CSB + 1, CTM = 8, CXL = WF<CBL>::LD;
constexpr size_t CBB = (size_t)CBL * CBL;
constexpr int CSTAGE = CSB * CLD;   // LDS doubles of one staged stage block
constexpr size_t CLDS = ((size_t)CBL * CXL + 3 * CSTAGE) * sizeof(double);

__device__ inline double coupled_coef(const CoupledSrc &src, int s, int t) {
    const int idx = 3 * s + t;
    double v = src.coef[0];
#pragma unroll
    for (int j = 1; j < 9; j++) v = idx == j ? src.coef[j] : v;
    return v;
}

// the three stage blocks of slot q into registers (CSU double2 per thread), and from there to Jl
// (column stride CLD): the loads of a stage set are issued before a product phase and stored after it
constexpr int CSU = (3 * CSB * CSB / 2 + FWG - 1) / FWG;
__device__ inline void coupled_stage_load(const CoupledSrc &src, int q, double2 (&v)[CSU]) {
    constexpr int SB = CSB * CSB, NU = 3 * SB / 2;
#pragma unroll
    for (int j = 0; j < CSU; j++) {
        const int u = min((int)threadIdx.x + FWG * j, NU - 1);
        const int st = (2 * u) / SB, rem = 2 * u - st * SB;
        const double *J = st == 0 ? src.J0 : st == 1 ? src.J1 : src.J2;
        v[j] = *reinterpret_cast<const double2 *>(J + (size_t)q * SB + rem);
    }
}
__device__ inline void coupled_stage_store(double *Jl, const double2 (&v)[CSU]) {
    constexpr int SB = CSB * CSB, NU = 3 * SB / 2;
#pragma unroll
    for (int j = 0; j < CSU; j++) {
        const int u = (int)threadIdx.x + FWG * j;
        if (u < NU) {
            const int st = (2 * u) / SB, rem = 2 * u - st * SB, c = rem / CSB, r = rem - c * CSB;
            Jl[st * CSTAGE + c * CLD + r] = v[j].x;
            Jl[st * CSTAGE + c * CLD + r + 1] = v[j].y;
        }
    }
}

// The coupled block factors as C_q = BD(J_q) (coef x I): BD the block diagonal of the three
// stage blocks (row stripe s of C_q holds J_s), (coef x I) the 3 x 3 stage mixing.  Products with
// BD need only the k-steps of a tile's own stage block(s): 8 x 8 output tiles of 16 over the
// padded 128 x 128 block, wavefront w owns tile columns w and w + 4; tile row or column index j
// covers rows 16 j .. 16 j + 15, i.e. stages (16 j) / 40 .. min((16 j + 15) / 40, 2).

// X <- (coef x I) X in place: rows (s, r) of every column mixed over s
__device__ inline void coupled_mix_rows(double *X, const CoupledSrc &src) {
    for (int e = threadIdx.x; e < CSB * CBL; e += FWG) {
        const int n = e / CSB, r = e - n * CSB;
        double *x = X + (size_t)n * CXL + r;
        const double x0 = x[0], x1 = x[CSB], x2 = x[2 * CSB];
#pragma unroll
        for (int st = 0; st < 3; st++)
            x[st * CSB] = fma(src.coef[3 * st], x0, fma(src.coef[3 * st + 1], x1, src.coef[3 * st + 2] * x2));
    }
}

// acc[2 tm + h] = tile (tm, w + 4 h) of BD(J) X (J from the staged Jl, X in LDS, stride CXL)
__device__ inline void bd_times_x(const double *Jl, const double *X, dx4 (&acc)[16]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = dx4{0.0, 0.0, 0.0, 0.0};
    const int n0 = 16 * wv + lr, n1 = 16 * (wv + 4) + lr;
    const bool v1 = n1 < CBL;
#pragma unroll
    for (int t = 0; t < 3; t++) {
#pragma unroll
        for (int kk = 0; kk < CSB / 4; kk++) {
            const int c = 4 * kk + lk, k = CSB * t + c;
            const double b0 = X[n0 * CXL + k];
            const double b1 = v1 ? X[n1 * CXL + k] : 0.0;
#pragma unroll
            for (int tm = 0; tm < CTM; tm++) {
                // tile rows of stage t only (compile-time per tm, t)
                if ((16 * tm) / CSB > t || (16 * tm + 15) / CSB < t) continue;
                const int m = 16 * tm + lr;
                const double a = (m < CBL && m / CSB == t) ? Jl[t * CSTAGE + c * CLD + (m - t * CSB)] : 0.0;
                acc[2 * tm] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[2 * tm], 0, 0, 0);
                acc[2 * tm + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[2 * tm + 1], 0, 0, 0);
            }
        }
    }
}

// stage s's k-steps of X BD(J) for the tile columns enabled by H0, H1
template <bool H0, bool H1>
__device__ inline void x_bd_stage(const double *X, const double *Jl, int s, int c0, int t0, int c1, int t1, int lr,
                                  int lk, dx4 (&acc)[16]) {
#pragma unroll
    for (int kk = 0; kk < CSB / 4; kk++) {
        const int r = 4 * kk + lk, k = CSB * s + r;
        const double b0 = H0 && t0 == s ? Jl[s * CSTAGE + c0 * CLD + r] : 0.0;
        const double b1 = H1 && t1 == s ? Jl[s * CSTAGE + c1 * CLD + r] : 0.0;
#pragma unroll
        for (int tm = 0; tm < CTM; tm++) {
            const int m = 16 * tm + lr;
            const double a = m < CBL ? X[k * CXL + m] : 0.0;
            if (H0) acc[2 * tm] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[2 * tm], 0, 0, 0);
            if (H1) acc[2 * tm + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[2 * tm + 1], 0, 0, 0);
        }
    }
}

// acc[2 tm + h] += tile (tm, w + 4 h) of X BD(J)
__device__ inline void x_times_bd(const double *X, const double *Jl, dx4 (&acc)[16]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int n0 = 16 * wv + lr, n1 = 16 * (wv + 4) + lr;
    const bool v1 = n1 < CBL;
    const int t0 = n0 / CSB, c0 = n0 - t0 * CSB;
    const int t1 = v1 ? n1 / CSB : 3, c1 = v1 ? n1 - t1 * CSB : 0;
    for (int s = 0; s < 3; s++) {
        // stages of this wavefront's two tile columns (uniform per wavefront)
        const bool h0 = (16 * wv) / CSB <= s && s <= (16 * wv + 15) / CSB;
        const bool h1 = (16 * (wv + 4)) / CSB <= s && s <= (16 * (wv + 4) + 15) / CSB;
        if (h0 && h1) x_bd_stage<true, true>(X, Jl, s, c0, t0, c1, t1, lr, lk, acc);
        else if (h0) x_bd_stage<true, false>(X, Jl, s, c0, t0, c1, t1, lr, lk, acc);
        else if (h1) x_bd_stage<false, true>(X, Jl, s, c0, t0, c1, t1, lr, lk, acc);
    }
}

// Y <- Y (coef x I) in place: columns (s, c) of every row mixed over s
__device__ inline void coupled_mix_cols(double *Y, const CoupledSrc &src) {
    for (int e = threadIdx.x; e < CSB * CBL; e += FWG) {
        const int c = e / CBL, m = e - c * CBL;
        double *y = Y + (size_t)c * CXL + m;
        const double y0 = y[0], y1 = y[(size_t)CSB * CXL], y2 = y[(size_t)2 * CSB * CXL];
#pragma unroll
        for (int t = 0; t < 3; t++)
            y[(size_t)t * CSB * CXL] = fma(src.coef[t], y0, fma(src.coef[3 + t], y1, src.coef[6 + t] * y2));
    }
}

// element e (column-major) of Y (coef x I) for Y in LDS (stride CXL)
__device__ inline double coupled_mix_col(const double *Y, const CoupledSrc &src, int e) {
    const int col = e / CBL, m = e - col * CBL, t = col / CSB, c = col - t * CSB;
    const double *y = Y + (size_t)c * CXL + m;
    return fma(coupled_coef(src, 0, t), y[0],
               fma(coupled_coef(src, 1, t), y[(size_t)CSB * CXL], coupled_coef(src, 2, t) * y[(size_t)2 * CSB * CXL]));
}

template <class T>
__global__ __launch_bounds__(FWG) void k_cdilu_schur(RowLists L, CoupledSrc src, T *__restrict__ F,
                                                     float *__restrict__ scale, double *__restrict__ DD64) {
    extern __shared__ double lds_c[];
    double *X = lds_c;
    double *Jl = X + CBL * CXL;
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6;
    const int i = L.rows[blockIdx.x];
    // accY = sum over lower neighbours k of L_ik BD(J_ki), kept in the accumulators; then
    // S_i = C_ii - accY (coef x I) in one pass
    dx4 accY[16];
#pragma unroll
    for (int j = 0; j < 16; j++) accY[j] = dx4{0.0, 0.0, 0.0, 0.0};
    // the stage sets pass through registers: each is loaded one product phase before it is stored
    double2 nxt[CSU];
    const int q0 = L.lptr[i], q1 = L.lptr[i + 1];
    if (q0 < q1) coupled_stage_load(src, L.lslot[q0], nxt);
    for (int q = q0; q < q1; q++) {
        const int k = L.lcol[q], st = L.ltrans[q];
        __syncthreads();   // X and Jl free
        coupled_stage_store(Jl, nxt);   // J_ik's stage blocks
        wg_block_to_lds<CBL>(X, DD64 + (size_t)k * CBB);
        __syncthreads();
        coupled_mix_rows(X, src);   // W = (coef x I) D_k^-1
        __syncthreads();
        coupled_stage_load(src, st, nxt);   // J_ki's, in flight during the product
        dx4 acc[16];
        bd_times_x(Jl, X, acc);     // L_ik = C_ik D_k^-1 = BD(J_ik) W
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; j++) tile_to_lds<CBL>(X, j >> 1, wv + 4 * (j & 1), acc[j]);
        coupled_stage_store(Jl, nxt);
        __syncthreads();
        const int sl = L.lslot[q];
        wg_lds_block_store<T, CBL>(X, F + (size_t)L.pos[sl] * CBB, scale ? scale + L.pos[sl] : nullptr, red);
        if (q + 1 < q1) coupled_stage_load(src, L.lslot[q + 1], nxt);   // the next J_ik's
        x_times_bd(X, Jl, accY);    // accY += L_ik BD(J_ki)
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) tile_to_lds<CBL>(X, j >> 1, wv + 4 * (j & 1), accY[j]);
    __syncthreads();
    // S_i = C_ii - accY (coef x I), mass term included
    const int d = L.diag[i];
    const int bs = src.bs;
    double *S = DD64 + (size_t)i * CBB;
    wg_batched<(int)CBB>(
        [&](int e) {
            const int col = e / CBL, m = e - col * CBL;
            const int st = m / CSB, r = m - st * CSB, c = col % CSB;
            const double *J = st == 0 ? src.J0 : st == 1 ? src.J1 : src.J2;
            return double2{J[(size_t)d * CSB * CSB + c * CSB + r], src.mass[(size_t)i * bs * bs + (c % bs) * bs + (r % bs)]};
        },
        [&](int e, double2 a) {
            const int col = e / CBL, m = e - col * CBL;
            const int st = m / CSB, r = m - st * CSB, t = col / CSB, c = col - t * CSB;
            S[e] = fma(coupled_coef(src, st, t), a.x, st == t && r / bs == c / bs ? a.y : 0.0) - coupled_mix_col(X, src, e);
        });
}

template <class T>
__global__ __launch_bounds__(WG, 2) void k_cdilu_inv(const int *__restrict__ rows, const int *__restrict__ diag,
                                                     const int *__restrict__ pos, double *DD64, T *__restrict__ F,
                                                     float *__restrict__ scale, int *__restrict__ status) {
    const int i = rows[blockIdx.x];
    const int pd = pos[diag[i]];
    double *S = DD64 + (size_t)i * CBB;
    mfma_gauss_jordan<T, CBL, true>(S, S, F + (size_t)pd * CBB, scale ? scale + pd : nullptr, status);
}

template <class T>
__global__ __launch_bounds__(FWG) void k_cdilu_upper(const int *__restrict__ urows, const int *__restrict__ uptr,
                                                     const int *__restrict__ uslot, const int *__restrict__ pos,
                                                     CoupledSrc src, const double *__restrict__ DD64,
                                                     T *__restrict__ F, float *__restrict__ scale) {
    extern __shared__ double lds_c[];
    double *X = lds_c;
    double *Jl = X + CBL * CXL;
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6;
    const int i = urows[blockIdx.x];
    // each slot's stage set passes through registers, loaded during the previous slot's product
    double2 nxt[CSU];
    const int q0 = uptr[i], q1 = uptr[i + 1];
    if (q0 < q1) coupled_stage_load(src, uslot[q0], nxt);
    for (int q = q0; q < q1; q++) {
        const int sl = uslot[q];
        __syncthreads();   // X, Jl and red free
        coupled_stage_store(Jl, nxt);
        wg_block_to_lds<CBL>(X, DD64 + (size_t)i * CBB);
        __syncthreads();
        if (q + 1 < q1) coupled_stage_load(src, uslot[q + 1], nxt);
        dx4 acc[16];
#pragma unroll
        for (int j = 0; j < 16; j++) acc[j] = dx4{0.0, 0.0, 0.0, 0.0};
        x_times_bd(X, Jl, acc);   // Y = D_i^-1 BD(J_ij)
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; j++) tile_to_lds<CBL>(X, j >> 1, wv + 4 * (j & 1), acc[j]);
        __syncthreads();
        coupled_mix_cols(X, src);   // DU_ij = Y (coef x I)
        __syncthreads();
        wg_lds_block_store<T, CBL>(X, F + (size_t)pos[sl] * CBB, scale ? scale + pos[sl] : nullptr, red);
    }
}

// ---- the coupled factor with FP32 products, for FP32 and BF16 factors ----------------------
// Factors stored in FP32 or BF16 take their Schur products and DU blocks on
// v_mfma_f32_16x16x4f32, at twice the FP64 MFMA rate, from FP32 copies of D_k^-1 and of the stage
// blocks staged in LDS: 76 KB per workgroup, so two workgroups share a CU where the FP64 kernels
// hold one.  The stage mixing, S_i = C_ii - accY (coef x I) and the Gauss-Jordan inverse stay in
// FP64; the FP32 rounding of the products (about 1e-7 relative to |L_ik| |BD(J_ki)|) is below the
// FP32 storage of the factor it feeds.  The FP32 16x16x4 MFMA takes its inputs like the FP64 one
// (lane l: A row l % 16, B column l % 16, k-step l / 16) but returns row 4 (l / 16) + e of column
// l % 16 in register e (the FP64 one: row l / 16 + 4 e).
typedef float fx4 __attribute__((ext_vector_type(4)));
constexpr size_t CLDS32 = ((size_t)CBL * CXL + 3 * CSTAGE) * sizeof(float);

__device__ inline void coupled_stage_store32(float *Jl, const double2 (&v)[CSU]) {
    constexpr int SB = CSB * CSB, NU = 3 * SB / 2;
#pragma unroll
    for (int j = 0; j < CSU; j++) {
        const int u = (int)threadIdx.x + FWG * j;
        if (u < NU) {
            const int st = (2 * u) / SB, rem = 2 * u - st * SB, c = rem / CSB, r = rem - c * CSB;
            Jl[st * CSTAGE + c * CLD + r] = (float)v[j].x;
            Jl[st * CSTAGE + c * CLD + r + 1] = (float)v[j].y;
        }
    }
}

// the three stage blocks of slot q straight to Jl (FP32), all threads, loads batched
__device__ inline void coupled_stage_to_lds32(float *Jl, const CoupledSrc &src, int q) {
    constexpr int SB = CSB * CSB, NU = 3 * SB / 2;
    wg_batched<NU>(
        [&](int u) {
            const int st = (2 * u) / SB, rem = 2 * u - st * SB;
            const double *J = st == 0 ? src.J0 : st == 1 ? src.J1 : src.J2;
            return *reinterpret_cast<const double2 *>(J + (size_t)q * SB + rem);
        },
        [&](int u, double2 v) {
            const int st = (2 * u) / SB, rem = 2 * u - st * SB, c = rem / CSB, r = rem - c * CSB;
            Jl[st * CSTAGE + c * CLD + r] = (float)v.x;
            Jl[st * CSTAGE + c * CLD + r + 1] = (float)v.y;
        });
}

// X (FP32, column stride CXL) <- the FP64 120 x 120 block at g
__device__ inline void coupled_block_to_lds32(float *X, const double *__restrict__ g) {
    wg_batched<(int)CBB / 2>([&](int u) { return *reinterpret_cast<const double2 *>(g + 2 * u); },
                             [&](int u, double2 a) {
                                 const int c = (2 * u) / CBL, r = 2 * u - c * CBL;
                                 X[c * CXL + r] = (float)a.x;
                                 X[c * CXL + r + 1] = (float)a.y;
                             });
}

// X <- D^-1 of the row whose diagonal slot is pd (FP32): FP32 factors hold it rounded to FP32 in
// that slot, the same values as rounding DD64's, at half the bytes; BF16 factors round DD64's
template <class T>
__device__ inline void coupled_dinv_to_lds32(float *X, const double *__restrict__ DD64, const T *F, int row, int pd) {
    if constexpr (std::is_same<T, float>::value) {
        const float *g = F + (size_t)pd * CBB;
        wg_batched<(int)CBB / 4>([&](int u) { return *reinterpret_cast<const float4 *>(g + 4 * u); },
                                 [&](int u, float4 a) {
                                     const int c = (4 * u) / CBL, r = 4 * u - c * CBL;
                                     float *x = X + c * CXL + r;
                                     x[0] = a.x;
                                     x[1] = a.y;
                                     x[2] = a.z;
                                     x[3] = a.w;
                                 });
    } else {
        coupled_block_to_lds32(X, DD64 + (size_t)row * CBB);
    }
}

__device__ inline void coupled_mix_rows32(float *X, const CoupledSrc &src) {
    for (int e = threadIdx.x; e < CSB * CBL; e += FWG) {
        const int n = e / CSB, r = e - n * CSB;
        float *x = X + (size_t)n * CXL + r;
        const double x0 = x[0], x1 = x[CSB], x2 = x[2 * CSB];
#pragma unroll
        for (int st = 0; st < 3; st++)
            x[st * CSB] = (float)fma(src.coef[3 * st], x0, fma(src.coef[3 * st + 1], x1, src.coef[3 * st + 2] * x2));
    }
}

// Y <- Y (coef x I); returns this thread's max of |mixed value|
__device__ inline double coupled_mix_cols32(float *Y, const CoupledSrc &src) {
    double pm = 0.0;
    for (int e = threadIdx.x; e < CSB * CBL; e += FWG) {
        const int c = e / CBL, m = e - c * CBL;
        float *y = Y + (size_t)c * CXL + m;
        const double y0 = y[0], y1 = y[(size_t)CSB * CXL], y2 = y[(size_t)2 * CSB * CXL];
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const float v = (float)fma(src.coef[t], y0, fma(src.coef[3 + t], y1, src.coef[6 + t] * y2));
            y[(size_t)t * CSB * CXL] = v;
            pm = fmax(pm, fabs((double)v));
        }
    }
    return pm;
}

__device__ inline double coupled_mix_col32(const float *Y, const CoupledSrc &src, int e) {
    const int col = e / CBL, m = e - col * CBL, t = col / CSB, c = col - t * CSB;
    const float *y = Y + (size_t)c * CXL + m;
    return fma(coupled_coef(src, 0, t), (double)y[0],
               fma(coupled_coef(src, 1, t), (double)y[(size_t)CSB * CXL],
                   coupled_coef(src, 2, t) * (double)y[(size_t)2 * CSB * CXL]));
}

// wg_lds_block_store with the BF16 block max given: pm is this thread's max of |stored value| over
// its share of the block (taken where the values were formed, which saves a pass over X)
template <class T>
__device__ inline void coupled_store32(const float *X, T *__restrict__ g, float *scale_out, double *red, double pm) {
    using W = WF<CBL>;
    float inv = 1.0f;
    if constexpr (Store<T>::prec == CPG_BF16) {
        pm = wave_max(pm);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pm;
        __syncthreads();
        const double mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        const float sc = pow2_scale(mx);
        inv = 1.0f / sc;
        if (threadIdx.x == 0) *scale_out = sc;
    }
    for (int e = threadIdx.x; e < W::BB; e += FWG) {
        const int c = e / CBL, r = e - c * CBL;
        Store<T>::put(g, e, (double)X[c * W::LD + r], inv);
    }
}

// tile (tm, tn) of an FP32 16x16x4 accumulator to X
__device__ inline void tile_to_lds32(float *X, int tm, int tn, const fx4 &t) {
    const int l = threadIdx.x & 63;
    const int col = 16 * tn + (l & 15);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int row = 16 * tm + 4 * (l >> 4) + e;
        if (row < CBL && col < CBL) X[col * CXL + row] = t[e];
    }
}

// acc[2 tm + h] = tile (tm, w + 4 h) of BD(J) X (bd_times_x on FP32)
__device__ inline void bd_times_x32(const float *Jl, const float *X, fx4 (&acc)[16]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = fx4{0.0f, 0.0f, 0.0f, 0.0f};
    const int n0 = 16 * wv + lr, n1 = 16 * (wv + 4) + lr;
    const bool v1 = n1 < CBL;
#pragma unroll
    for (int t = 0; t < 3; t++) {
#pragma unroll
        for (int kk = 0; kk < CSB / 4; kk++) {
            const int c = 4 * kk + lk, k = CSB * t + c;
            const float b0 = X[n0 * CXL + k];
            const float b1 = v1 ? X[n1 * CXL + k] : 0.0f;
#pragma unroll
            for (int tm = 0; tm < CTM; tm++) {
                if ((16 * tm) / CSB > t || (16 * tm + 15) / CSB < t) continue;
                const int m = 16 * tm + lr;
                const float a = (m < CBL && m / CSB == t) ? Jl[t * CSTAGE + c * CLD + (m - t * CSB)] : 0.0f;
                acc[2 * tm] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b0, acc[2 * tm], 0, 0, 0);
                acc[2 * tm + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b1, acc[2 * tm + 1], 0, 0, 0);
            }
        }
    }
}

template <bool H0, bool H1>
__device__ inline void x_bd_stage32(const float *X, const float *Jl, int s, int c0, int t0, int c1, int t1, int lr,
                                    int lk, fx4 (&acc)[16]) {
#pragma unroll
    for (int kk = 0; kk < CSB / 4; kk++) {
        const int r = 4 * kk + lk, k = CSB * s + r;
        const float b0 = H0 && t0 == s ? Jl[s * CSTAGE + c0 * CLD + r] : 0.0f;
        const float b1 = H1 && t1 == s ? Jl[s * CSTAGE + c1 * CLD + r] : 0.0f;
#pragma unroll
        for (int tm = 0; tm < CTM; tm++) {
            const int m = 16 * tm + lr;
            const float a = m < CBL ? X[k * CXL + m] : 0.0f;
            if (H0) acc[2 * tm] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b0, acc[2 * tm], 0, 0, 0);
            if (H1) acc[2 * tm + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b1, acc[2 * tm + 1], 0, 0, 0);
        }
    }
}

// acc[2 tm + h] += tile (tm, w + 4 h) of X BD(J) (x_times_bd on FP32)
__device__ inline void x_times_bd32(const float *X, const float *Jl, fx4 (&acc)[16]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int n0 = 16 * wv + lr, n1 = 16 * (wv + 4) + lr;
    const bool v1 = n1 < CBL;
    const int t0 = n0 / CSB, c0 = n0 - t0 * CSB;
    const int t1 = v1 ? n1 / CSB : 3, c1 = v1 ? n1 - t1 * CSB : 0;
    for (int s = 0; s < 3; s++) {
        const bool h0 = (16 * wv) / CSB <= s && s <= (16 * wv + 15) / CSB;
        const bool h1 = (16 * (wv + 4)) / CSB <= s && s <= (16 * (wv + 4) + 15) / CSB;
        if (h0 && h1) x_bd_stage32<true, true>(X, Jl, s, c0, t0, c1, t1, lr, lk, acc);
        else if (h0) x_bd_stage32<true, false>(X, Jl, s, c0, t0, c1, t1, lr, lk, acc);
        else if (h1) x_bd_stage32<false, true>(X, Jl, s, c0, t0, c1, t1, lr, lk, acc);
    }
}

// k_cdilu_schur with FP32 products (T = float or bf16)
template <class T>
__global__ __launch_bounds__(FWG, 2) void k_cdilu_schur32(RowLists L, CoupledSrc src, T *__restrict__ F,
                                                         float *__restrict__ scale, double *__restrict__ DD64) {
    extern __shared__ float lds_f[];
    float *X = lds_f;
    float *Jl = X + CBL * CXL;
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6;
    const int i = L.rows[blockIdx.x];
    fx4 accY[16];
#pragma unroll
    for (int j = 0; j < 16; j++) accY[j] = fx4{0.0f, 0.0f, 0.0f, 0.0f};
    const int q0 = L.lptr[i], q1 = L.lptr[i + 1];
    for (int q = q0; q < q1; q++) {
        const int k = L.lcol[q], st = L.ltrans[q];
        __syncthreads();   // X and Jl free
        coupled_stage_to_lds32(Jl, src, L.lslot[q]);   // J_ik's stage blocks
        coupled_dinv_to_lds32<T>(X, DD64, F, k, L.pos[L.diag[k]]);
        __syncthreads();
        coupled_mix_rows32(X, src);   // W = (coef x I) D_k^-1
        __syncthreads();
        fx4 acc[16];
        bd_times_x32(Jl, X, acc);     // L_ik = BD(J_ik) W
        double pm = 0.0;   // BF16 scale: the block max from the accumulators (padding holds zeros)
        if constexpr (Store<T>::prec == CPG_BF16) {
#pragma unroll
            for (int j = 0; j < 16; j++)
#pragma unroll
                for (int e = 0; e < 4; e++) pm = fmax(pm, fabs((double)acc[j][e]));
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; j++) tile_to_lds32(X, j >> 1, wv + 4 * (j & 1), acc[j]);
        coupled_stage_to_lds32(Jl, src, st);   // J_ki's
        __syncthreads();
        const int sl = L.lslot[q];
        coupled_store32<T>(X, F + (size_t)L.pos[sl] * CBB, scale ? scale + L.pos[sl] : nullptr, red, pm);
        x_times_bd32(X, Jl, accY);    // accY += L_ik BD(J_ki)
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) tile_to_lds32(X, j >> 1, wv + 4 * (j & 1), accY[j]);
    __syncthreads();
    const int d = L.diag[i];
    const int bs = src.bs;
    double *S = DD64 + (size_t)i * CBB;
    wg_batched<(int)CBB>(
        [&](int e) {
            const int col = e / CBL, m = e - col * CBL;
            const int st = m / CSB, r = m - st * CSB, c = col % CSB;
            const double *J = st == 0 ? src.J0 : st == 1 ? src.J1 : src.J2;
            return double2{J[(size_t)d * CSB * CSB + c * CSB + r], src.mass[(size_t)i * bs * bs + (c % bs) * bs + (r % bs)]};
        },
        [&](int e, double2 a) {
            const int col = e / CBL, m = e - col * CBL;
            const int st = m / CSB, r = m - st * CSB, t = col / CSB, c = col - t * CSB;
            S[e] = fma(coupled_coef(src, st, t), a.x, st == t && r / bs == c / bs ? a.y : 0.0) - coupled_mix_col32(X, src, e);
        });
}

// k_cdilu_upper with FP32 products (T = float or bf16)
template <class T>
__global__ __launch_bounds__(FWG, 2) void k_cdilu_upper32(const int *__restrict__ urows, const int *__restrict__ uptr,
                                                         const int *__restrict__ uslot, const int *__restrict__ pos,
                                                         const int *__restrict__ diag, CoupledSrc src,
                                                         const double *__restrict__ DD64, T *__restrict__ F,
                                                         float *__restrict__ scale) {
    extern __shared__ float lds_f[];
    float *X = lds_f;
    float *Jl = X + CBL * CXL;
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6;
    const int i = urows[blockIdx.x];
    double2 nxt[CSU];
    const int q0 = uptr[i], q1 = uptr[i + 1];
    if (q0 < q1) coupled_stage_load(src, uslot[q0], nxt);
    for (int q = q0; q < q1; q++) {
        const int sl = uslot[q];
        __syncthreads();   // X, Jl and red free
        coupled_stage_store32(Jl, nxt);
        coupled_dinv_to_lds32<T>(X, DD64, F, i, pos[diag[i]]);
        __syncthreads();
        if (q + 1 < q1) coupled_stage_load(src, uslot[q + 1], nxt);
        fx4 acc[16];
#pragma unroll
        for (int j = 0; j < 16; j++) acc[j] = fx4{0.0f, 0.0f, 0.0f, 0.0f};
        x_times_bd32(X, Jl, acc);   // Y = D_i^-1 BD(J_ij)
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; j++) tile_to_lds32(X, j >> 1, wv + 4 * (j & 1), acc[j]);
        __syncthreads();
        const double pm = coupled_mix_cols32(X, src);   // DU_ij = Y (coef x I)
        __syncthreads();
        coupled_store32<T>(X, F + (size_t)pos[sl] * CBB, scale ? scale + pos[sl] : nullptr, red, pm);
    }
}

}  // namespace cpg

namespace cpg {

// ---- block DILU apply in one launch (FP64 vectors) ------------------------------------------------
// The level launches of apply_t order the sweep's rows by launch boundaries, so every level pays a
// ramp and a tail.  Here one workgroup per row task takes the tasks in level order by workgroup index
// (forward rows of levels 1 on, then backward rows): a task issues its row's first block loads, waits
// until every row it reads has finished in this apply, runs the row-stream body and then marks its own
// row finished, so a level's tail overlaps the next level's start.  Completion counters, one per row
// and sweep direction, rise by one per apply; a task reads its own row's counter g (the applies so
// far) and waits for g + 1 on the rows it reads.  The vectors written inside the launch (y, x) are
// stored and gathered at device scope, and a row's stores complete before its counter rises, so no
// cache writeback or invalidation is needed.  Workgroups of one launch are dispatched to each XCD in
// index order and a task waits only on tasks of lower index, so every wait ends; the waits are still
// bounded (spin_max polls): one that runs out sets the abort word and gives up, the task goes on (the
// counters stay consistent) and the host reports the apply as failed.
//
// Task t's record (rs ints at rec + t rs, built with the preconditioner): row, first packed block,
// block count, the fused D^-1 block (the last forward level, whose rows have no upper blocks, or -1),
// then the source tags of its blocks (stream_source), padded with tags of the input vector.  The
// record's header comes through scalar loads and its tags through one vector load per lane, neither
// depending on other data, so a task's dependency counters are one round trip behind its start.
constexpr int kFlowMaxNb = 32;   // blocks per row

struct FlowTable {
    const int *rec = nullptr;
    int rs = 0;                                    // ints per record: 4 + the padded tag count
    int bwd0 = 0;                                  // first backward task
    unsigned *fdone = nullptr, *bdone = nullptr;   // completions of each row's forward / backward step
    unsigned *abort = nullptr;                     // host-mapped word, set by a wait that ran out
    int spin_max = 0;
    const float *DF = nullptr;                     // CPG_BF16_D32: the FP32 inverses (row order)
    double *hbuf = nullptr;                        // and the products h = D^-1 y
};

// MODE 0: a task's row stream.  For factors with FP32 diagonal inverses (CPG_BF16_D32) a backward
// task runs two: MODE 1 forms h_i = D_i^-1 (y_i or in_i) from the FP32 inverse (the record's fourth
// entry is that product's source tag) after waiting for every source of the task, MODE 2 then
// streams the BF16 DU blocks onto h_i and raises the task's counter.
template <int MODE>
struct FlowHooks {
    static constexpr bool coh_src = true, coh_out = MODE != 1;
    const int *rec;         // this task's record
    int *stag;              // LDS: the row's source tags
    int ntag;               // padded tag count
    unsigned *own;          // completion counters of this task's step
    const unsigned *dep1;   // counters of the src1 rows (y forward, x backward)
    const unsigned *dep0;   // counters of the src0 rows (y, backward)
    unsigned *abort;
    int spin_max;
    unsigned g = 0;         // this task's completions before this apply
    int jt = 0;             // the thread's source tag (threads < ntag; MODE 1: thread ntag the D^-1 source)
    const unsigned *fc = nullptr;   // and the counter it waits on (null: none)
    unsigned fv = 0;        // that counter's first poll
    __device__ void row(const RowStream &, int, int &i, int &b0, int &nb) const {
        i = rec[0];
        b0 = MODE == 1 ? i : rec[1];   // FP32 inverses in row order
        nb = MODE == 1 ? 1 : rec[2];
    }
    __device__ int tag(const RowStream &, int, int b, int) const { return MODE == 1 ? rec[3] : stag[b]; }
    __device__ bool has_d(const RowStream &) const { return MODE == 0 && rec[3] >= 0; }
    __device__ int dblock(const RowStream &, int) const { return rec[3]; }
    // the thread's tag (one per thread: ntag <= kFlowMaxNb < the workgroup size), its counter's
    // first poll and the task's own counter, issued before the row's first block loads, so that
    // waiting for them leaves those in flight
    __device__ void prefetch(int i, int tid) {
        if constexpr (MODE == 2) return;
        jt = tid < ntag ? rec[4 + tid] : MODE == 1 && tid == ntag ? rec[3] : -1 - kSrc2Tag;
        g = __hip_atomic_load(own + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int u = -1 - jt;
        fc = jt >= 0 ? dep1 + jt : u < kSrc2Tag ? dep0 + u : nullptr;   // src2 (the input) needs none
        if (fc) fv = __hip_atomic_load(fc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ void wait(int, int, int, const int *, int tid, int) {
        if constexpr (MODE == 2) return;   // MODE 1 waited for this task's sources
        if (tid <= ntag) {
            if (tid < ntag) stag[tid] = jt;
            for (int k = 0; fc && fv != g + 1u; k++) {
                if (k >= spin_max ||
                    ((k & 255) == 255 && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))) {
                    __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                fv = __hip_atomic_load(fc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
    }
    __device__ void publish(int i, int tid) {
        if constexpr (MODE == 1) {
            __syncthreads();   // h_i (this workgroup's plain stores) before the DU stream reads it
            return;
        }
        // every wave's result stores (device scope) complete before the counter is raised
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(own + i, g + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// x = M^-1 in through y (forward result); in may equal x (the pattern is structurally symmetric, so
// the backward step of a row that forward rows read as input comes after theirs).  Tasks [t0, t1):
// every task, or for factors with FP32 diagonal inverses the forward ones (k_dilu_flow_d32b runs
// the backward sweep)
template <class T, int BL, int WPR>
__global__ __launch_bounds__(64 * WPR) void k_dilu_flow(FlowTable A, const int *__restrict__ active,
                                                        const T *__restrict__ F, const float *__restrict__ scale,
                                                        const double *in, double *y, double *x) {
    __shared__ int stag[kFlowMaxNb];
    if (active && !*active) return;
    const int t = blockIdx.x;
    const int *rec = A.rec + (size_t)t * A.rs;
    const bool bwd = t >= A.bwd0, fin = rec[3] >= 0;   // backward; forward finished by D^-1
    FlowHooks<0> fl{rec, stag, A.rs - 4, bwd || fin ? A.bdone : A.fdone, bwd ? A.bdone : A.fdone,
                    bwd ? A.fdone : nullptr, A.abort, A.spin_max};
    const RowStream S{nullptr, nullptr, nullptr};
    // forward: y_i = in_i - sum L_ij y_j (level-0 sources read in); backward: x_i = D_i^-1 y_i - sum DU_ij x_j
    rowstream_row<T, BL, WPR, double>(S, 0, F, scale, bwd ? y : nullptr, bwd ? x : y, -1.0, in, bwd ? 1.0 : -1.0,
                                      bwd ? nullptr : in, nullptr, bwd || fin ? x : y, fl);
}

// The backward sweep of factors with FP32 diagonal inverses in one launch (task A.bwd0 + blockIdx.x):
// h_i = D_i^-1 y_i (in_i where row i has no lower blocks) from the FP32 inverse, then
// x_i = h_i - sum DU_ij x_j from the BF16 stream.  A launch of its own: with these two row-stream
// bodies inlined beside the forward task's in one kernel, the compiler's CFG simplification crashes,
// and as a called function the kernel spilled to a stack at half the occupancy.
template <class T, int BL, int WPR>
__global__ __launch_bounds__(64 * WPR) void k_dilu_flow_d32b(FlowTable A, const int *__restrict__ active,
                                                             const T *__restrict__ F, const float *__restrict__ scale,
                                                             const double *in, double *y, double *x) {
    __shared__ int stag[kFlowMaxNb];
    if (active && !*active) return;
    const int *rec = A.rec + (size_t)(A.bwd0 + (int)blockIdx.x) * A.rs;
    const RowStream S{nullptr, nullptr, nullptr};
    FlowHooks<1> f1{rec, stag, A.rs - 4, A.bdone, A.bdone, A.fdone, A.abort, A.spin_max};
    rowstream_row<float, BL, WPR, double>(S, 0, A.DF, nullptr, y, nullptr, 0.0, in, 1.0, nullptr, nullptr, A.hbuf,
                                          f1);
    FlowHooks<2> f2{rec, stag, A.rs - 4, A.bdone, A.bdone, A.fdone, A.abort, A.spin_max};
    f2.g = f1.g;
    rowstream_row<T, BL, WPR, double>(S, 0, F, scale, nullptr, x, -1.0, nullptr, 0.0, A.hbuf, nullptr, x, f2);
}

inline bool flow_compiled(int bl) {
    return bl == 3 || bl == 6 || bl == 10 || bl == 12 || bl == 15 || bl == 24 || bl == 40 || bl == 60;
}

// a host word the device can write
struct MappedWord {
    unsigned *h = nullptr, *d = nullptr;
    MappedWord() = default;
    MappedWord(const MappedWord &) = delete;
    MappedWord &operator=(const MappedWord &) = delete;
    ~MappedWord() {
        if (h) (void)hipHostFree(h);
    }
    void alloc() {
        if (h) return;
        CPG_HIP(hipHostMalloc((void **)&h, sizeof(unsigned), hipHostMallocMapped));
        *h = 0;
        CPG_HIP(hipHostGetDevicePointer((void **)&d, h, 0));
    }
};

}  // namespace cpg

using namespace cpg;

// ---- handle --------------------------------------------------------------------------
struct cpg_prec_s {
    uint64_t gen = cpg::next_generation();   // renewed by cpg_prec_set_inner, which reallocates
    int kind = CPG_PC_NONE, precision = CPG_FP64;
    int n = 0, bl = 0, nnzb = 0;
    // pattern-derived lists
    DevBuf<int> diag, lptr, lslot, lcol, ltrans, uptr, uslot, ucol;
    DevBuf<int> urow;             // row of each upper slot (flat, in uslot order)
    DevBuf<int> urows;            // rows with at least one upper block
    int nup = 0, nurows = 0;
    DevBuf<int> frows, brows;
    std::vector<int> flev, blev;  // level pointers into frows / brows
    std::vector<int> frow0, brow0;  // first row of each level whose rows are contiguous, else -1
    // packed factor streams: pos[slot] = packed block; forward stream (L blocks) and backward
    // stream (D^-1, DU blocks) pointers per position in frows / brows; scol = source of each
    // packed block in the tags of stream_source: the column block j (src1 = y forward, x
    // backward), -1-i for D_i^-1 y_i (src0 = y), or the input vector (src2) where the source
    // row has no lower blocks, so that y = x there and forward level 0 needs no launch
    DevBuf<int> pos, fptr, bptr, scol;
    std::vector<int> h_frows, h_brows, h_fptr, h_bptr, h_scol, h_hcol;   // host copies (persistent batched
                                                                         // solves, single-launch records)
    int nbmax_f = 1, nbmax_b = 1;
    // storage
    DevBuf<unsigned char> F;      // factor blocks in `precision`
    DevBuf<float> scale;          // per-block scale (BF16)
    DevBuf<double> DD64;          // FP64 D^-1 (DILU)
    DevBuf<double> ybuf;          // forward-sweep result y
    DevBuf<int> status;
    // CPG_BF16_D32: L and DU blocks in BF16, the diagonal inverses D_i^-1 in FP32 (DF, row order).
    // The backward stream then packs only the DU blocks; D_i^-1 (y_i or x_i) goes to hbuf first.
    bool diag32 = false;
    DevBuf<float> DF;
    DevBuf<double> hbuf;
    DevBuf<int> hcol;             // source tag of each row's D^-1 product (stream_source)
    std::vector<int> h_diag;      // diagonal slot of each row (host copy)
    size_t elem = 8;
    // inner FP32 GMRES (kind CPG_PC_INNER_GMRES)
    int inner_k = 20;
    size_t N = 0;
    int nchunk = 0;
    const cpg_bsr_s *A = nullptr;             // matrix of the last factor (pattern for the FP32 matvec)
    // FP32 copy of A (a BF16 copy is too coarse for this solve: with per-block scales its
    // ~2^-9 relative error, amplified by the condition number of M - dt a J at large CFL,
    // leaves the inner solve a poor approximation of A^-1 and multiplies the outer
    // iterations), basis (k+1) x N, work vectors
    DevBuf<float> A32, Q, iw, it;
    DevBuf<double> ipart, H, g, cs, sn, y, nrm;
    // block DILU apply in one launch (k_dilu_flow): level table, completion counters, abort word
    FlowTable flow;
    int flow_tasks = 0;
    bool flow_ok = false;        // the table applies to this preconditioner
    bool flow_enabled = true;    // cpg_prec_set_flow
    int flow_spin = 1 << 22;
    DevBuf<unsigned> fdone, bdone;
    DevBuf<int> flow_rec;        // task records
    MappedWord flow_abort;

    bool dilu() const { return kind == CPG_PC_DILU || kind == CPG_PC_INNER_GMRES; }

    RowLists lists(bool fwd) const {
        RowLists L;
        L.pos = pos.p;
        L.rows = fwd ? frows.p : brows.p;
        L.diag = diag.p;
        L.lptr = lptr.p; L.lslot = lslot.p; L.lcol = lcol.p; L.ltrans = ltrans.p;
        L.uptr = uptr.p; L.uslot = uslot.p; L.ucol = ucol.p;
        return L;
    }
};

namespace cpg {
uint64_t prec_generation(cpg_prec P) { return P ? P->gen : 0; }

PrecTables prec_tables(cpg_prec P) {
    PrecTables t;
    t.kind = P->kind;
    t.precision = P->precision;
    t.n = P->n;
    t.bl = P->bl;
    t.diag32 = P->diag32;
    t.F = P->F.p;
    t.scale = P->scale.p;
    t.scol = P->scol.p;
    t.flev = &P->flev;
    t.frows = &P->h_frows;
    t.fptr = &P->h_fptr;
    t.blev = &P->blev;
    t.brows = &P->h_brows;
    t.bptr = &P->h_bptr;
    return t;
}
}  // namespace cpg

namespace {

size_t elem_size(int prec) { return prec == CPG_FP64 ? 8 : prec == CPG_FP32 ? 4 : 2; }

template <class T>
T *Fp(cpg_prec_s &P) { return reinterpret_cast<T *>(P.F.p); }

template <class K>
void set_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        CPG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)bytes));
}

template <class T, int NE>
void factor_dilu_t(cpg_prec_s &P, const cpg_bsr_s &A, hipStream_t s) {
    const int bl = P.bl;
    const size_t lds = (size_t)(4 * bl * bl + bl + 8) * sizeof(double);
    CPG_CHECK(lds <= 160 * 1024, CPG_ERR_UNSUPPORTED, "block size %d too large for the LDS factor kernel", bl);
    auto kern = k_dilu_factor<T, NE>;
    set_lds(kern, lds);
    RowLists L = P.lists(true);
    for (size_t lv = 0; lv + 1 < P.flev.size(); lv++) {
        const int nr = P.flev[lv + 1] - P.flev[lv];
        if (!nr) continue;
        RowLists Ll = L;
        Ll.rows = P.frows.p + P.flev[lv];
        hipLaunchKernelGGL(kern, dim3(nr), dim3(WG), lds, s, bl, Ll, A.blocks.p, Fp<T>(P),
                           P.precision == CPG_BF16 ? P.scale.p : nullptr, P.DD64.p, P.status.p);
        CPG_LAUNCH_CHECK();
    }
}

template <class T, int BL>
void factor_bj_rt_bl(cpg_prec_s &P, const cpg_bsr_s &A, hipStream_t s) {
    hipLaunchKernelGGL((k_bjacobi_factor_rt<T, BL>), dim3(P.n), dim3(WG), 0, s, P.diag.p, A.blocks.p, Fp<T>(P),
                       P.precision == CPG_BF16 ? P.scale.p : nullptr, P.status.p);
    CPG_LAUNCH_CHECK();
}

// block sizes beyond the augmented kernel: the coupled Radau IIA blocks (ns stages x 4 x bs),
// 120 for 3 stages at p = 3 and 2 stages at p = 4
template <class T>
void factor_bj_big_t(cpg_prec_s &P, const cpg_bsr_s &A, hipStream_t s) {
    if (P.bl == 120) {
        factor_bj_rt_bl<T, 120>(P, A, s);
        return;
    }
    CPG_CHECK(false, CPG_ERR_UNSUPPORTED, "block size %d too large for block Jacobi", P.bl);
}

template <class T, int NE>
void factor_bj_t(cpg_prec_s &P, const cpg_bsr_s &A, hipStream_t s) {
    const int bl = P.bl;
    const size_t lds = (size_t)(2 * bl * bl + bl + 8) * sizeof(double);
    if (lds > 160 * 1024) {
        factor_bj_big_t<T>(P, A, s);
        return;
    }
    CPG_CHECK(lds <= 160 * 1024, CPG_ERR_UNSUPPORTED, "block size %d too large for the LDS factor kernel", bl);
    auto kern = k_bjacobi_factor<T, NE>;
    set_lds(kern, lds);
    hipLaunchKernelGGL(kern, dim3(P.n), dim3(WG), lds, s, bl, P.diag.p, A.blocks.p, Fp<T>(P),
                       P.precision == CPG_BF16 ? P.scale.p : nullptr, P.status.p);
    CPG_LAUNCH_CHECK();
}

template <class T, int BL>
void factor_wf_t(cpg_prec_s &P, const cpg_bsr_s &A, hipStream_t s) {
    float *sc = P.precision == CPG_BF16 ? P.scale.p : nullptr;
    if (!P.dilu()) {
        hipLaunchKernelGGL((k_bjacobi_factor_wf<T, BL>), dim3(P.n), dim3(64), 0, s, P.diag.p, A.blocks.p, Fp<T>(P), sc,
                           P.status.p);
        CPG_LAUNCH_CHECK();
        return;
    }
    RowLists L = P.lists(true);
    for (size_t lv = 0; lv + 1 < P.flev.size(); lv++) {
        const int nr = P.flev[lv + 1] - P.flev[lv];
        if (!nr) continue;
        RowLists Ll = L;
        Ll.rows = P.frows.p + P.flev[lv];
        hipLaunchKernelGGL((k_dilu_schur<T, BL>), dim3(nr), dim3(FWG), 0, s, Ll, A.blocks.p, Fp<T>(P), sc,
                           P.DD64.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_dilu_inv_wf<T, BL>), dim3(nr), dim3(64), 0, s, Ll.rows, P.diag.p, P.pos.p, P.DD64.p,
                           Fp<T>(P), sc, P.status.p);
        CPG_LAUNCH_CHECK();
    }
    if (P.nup > 0) {
        hipLaunchKernelGGL((k_dilu_upper<T, BL>), dim3(P.nurows), dim3(FWG), 0, s, P.urows.p, P.uptr.p, P.uslot.p,
                           P.pos.p, A.blocks.p, P.DD64.p, Fp<T>(P), sc);
        CPG_LAUNCH_CHECK();
    }
}

// block DILU of the stage-coupled system: per level the Schur and inverse kernels, then the DU
// blocks of every row
// FP32 and BF16 factors take the FP32-product kernels unless CPOLYG_COUPLED_FP64_PRODUCTS=1 (read
// at every factor call)
bool coupled_fp64_products() {
    const char *e = std::getenv("CPOLYG_COUPLED_FP64_PRODUCTS");
    return e && e[0] == '1';
}

template <class T>
void factor_coupled_t(cpg_prec_s &P, const CoupledSrc &src, hipStream_t s) {
    float *sc = P.precision == CPG_BF16 ? P.scale.p : nullptr;
    bool f32 = false;
    if constexpr (!std::is_same<T, double>::value) f32 = !coupled_fp64_products();
    if constexpr (!std::is_same<T, double>::value) {
        set_lds(k_cdilu_schur32<T>, CLDS32);
        set_lds(k_cdilu_upper32<T>, CLDS32);
    }
    set_lds(k_cdilu_schur<T>, CLDS);
    set_lds(k_cdilu_upper<T>, CLDS);
    RowLists L = P.lists(true);
    for (size_t lv = 0; lv + 1 < P.flev.size(); lv++) {
        const int nr = P.flev[lv + 1] - P.flev[lv];
        if (!nr) continue;
        RowLists Ll = L;
        Ll.rows = P.frows.p + P.flev[lv];
        if constexpr (!std::is_same<T, double>::value) {
            if (f32) hipLaunchKernelGGL(k_cdilu_schur32<T>, dim3(nr), dim3(FWG), CLDS32, s, Ll, src, Fp<T>(P), sc, P.DD64.p);
        }
        if (!f32) hipLaunchKernelGGL(k_cdilu_schur<T>, dim3(nr), dim3(FWG), CLDS, s, Ll, src, Fp<T>(P), sc, P.DD64.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_cdilu_inv<T>, dim3(nr), dim3(WG), 0, s, Ll.rows, P.diag.p, P.pos.p, P.DD64.p, Fp<T>(P), sc,
                           P.status.p);
        CPG_LAUNCH_CHECK();
    }
    if (P.nup > 0) {
        if constexpr (!std::is_same<T, double>::value) {
            if (f32)
                hipLaunchKernelGGL(k_cdilu_upper32<T>, dim3(P.nurows), dim3(FWG), CLDS32, s, P.urows.p, P.uptr.p,
                                   P.uslot.p, P.pos.p, P.diag.p, src, P.DD64.p, Fp<T>(P), sc);
        }
        if (!f32)
            hipLaunchKernelGGL(k_cdilu_upper<T>, dim3(P.nurows), dim3(FWG), CLDS, s, P.urows.p, P.uptr.p, P.uslot.p,
                               P.pos.p, src, P.DD64.p, Fp<T>(P), sc);
        CPG_LAUNCH_CHECK();
    }
}

// Block sizes of the DG systems (4 bs for Euler, bs for scalar advection, bs = (p+1)(p+2)/2)
// take the wavefront MFMA kernels; CPOLYG_FACTOR_LDS=1 selects the workgroup LDS kernels.
bool use_lds_factor() {
    static const bool v = [] {
        const char *e = std::getenv("CPOLYG_FACTOR_LDS");
        return e && e[0] == '1';
    }();
    return v;
}

template <class T>
void factor_t(cpg_prec_s &P, const cpg_bsr_s &A, hipStream_t s) {
    if (!use_lds_factor()) {
        switch (P.bl) {
        case 3: factor_wf_t<T, 3>(P, A, s); return;
        case 6: factor_wf_t<T, 6>(P, A, s); return;
        case 10: factor_wf_t<T, 10>(P, A, s); return;
        case 12: factor_wf_t<T, 12>(P, A, s); return;
        case 15: factor_wf_t<T, 15>(P, A, s); return;
        case 24: factor_wf_t<T, 24>(P, A, s); return;
        case 40: factor_wf_t<T, 40>(P, A, s); return;
        case 60: factor_wf_t<T, 60>(P, A, s); return;
        default: break;
        }
    }
    const int ne = ceil_div((long long)P.bl * P.bl, WG);
#define CPG_NE_CASE(NEV)                                                    \
    if (ne <= NEV) {                                                        \
        if (P.dilu()) factor_dilu_t<T, NEV>(P, A, s);                       \
        else factor_bj_t<T, NEV>(P, A, s);                                  \
        return;                                                             \
    }
    CPG_NE_CASE(1) CPG_NE_CASE(2) CPG_NE_CASE(4) CPG_NE_CASE(7) CPG_NE_CASE(15) CPG_NE_CASE(29) CPG_NE_CASE(57)
#undef CPG_NE_CASE
    CPG_CHECK(false, CPG_ERR_UNSUPPORTED, "block size %d too large", P.bl);
}

// The single-launch apply's level table (FP64 vectors; factors without FP32 diagonal inverses, a block
// size with a compiled k_dilu_flow): the forward levels from 1 (level 0 has y = x), the last one
// finishing its rows with D^-1 where the first backward level is the same rows, as apply_t fuses them,
// then the backward levels.  More levels than the table holds keep the level launches.
void flow_setup(cpg_prec_s &P) {
    P.flow_ok = false;
    P.flow_tasks = 0;
    if (P.kind != CPG_PC_DILU || !flow_compiled(P.bl)) return;
    const int nbm = std::max(P.nbmax_f, P.nbmax_b);
    if (nbm > kFlowMaxNb) return;
    const int nlf = (int)P.flev.size() - 1;
    const bool fuse = !P.diag32 && nlf >= 2 && P.frow0[nlf - 1] >= 0 && P.brow0.size() > 0 &&
                      P.brow0[0] == P.frow0[nlf - 1] && P.blev[1] - P.blev[0] == P.flev[nlf] - P.flev[nlf - 1];
    const int rs = 4 + ((nbm + 3) & ~3);   // 16-byte records
    std::vector<int> rec;
    // the fourth entry: the fused D^-1 block of a finishing forward row (else -1), or for a backward row
    // of factors with FP32 inverses the source tag of its product D_i^-1 y_i
    auto add = [&](const std::vector<int> &rows, const std::vector<int> &ptr, int p0, int p1, bool fin, bool bwd) {
        for (int p = p0; p < p1; p++) {
            const int i = rows[p], b0 = ptr[p], nb = ptr[p + 1] - b0;
            rec.push_back(i);
            rec.push_back(b0);
            rec.push_back(nb);
            rec.push_back(fin ? P.h_bptr[P.blev[0] + (p - p0)] : bwd && P.diag32 ? P.h_hcol[i] : -1);
            for (int b = 0; b < rs - 4; b++) rec.push_back(b < nb ? P.h_scol[b0 + b] : -1 - (kSrc2Tag + i));
        }
    };
    for (int lv = 1; lv < nlf; lv++)
        add(P.h_frows, P.h_fptr, P.flev[lv], P.flev[lv + 1], fuse && lv == nlf - 1, false);
    const int bwd0 = (int)(rec.size() / rs);
    for (int lv = fuse ? 1 : 0; lv + 1 < (int)P.blev.size(); lv++)
        add(P.h_brows, P.h_bptr, P.blev[lv], P.blev[lv + 1], false, true);
    const int tasks = (int)(rec.size() / rs);
    if (!tasks) return;
    P.flow_rec.upload(rec);
    if (P.fdone.n != (size_t)P.n) {
        P.fdone.alloc(P.n);
        P.bdone.alloc(P.n);
        CPG_HIP(hipMemset(P.fdone.p, 0, (size_t)P.n * sizeof(unsigned)));
        CPG_HIP(hipMemset(P.bdone.p, 0, (size_t)P.n * sizeof(unsigned)));
    }
    P.flow_abort.alloc();
    FlowTable T;
    T.rec = P.flow_rec.p;
    T.rs = rs;
    T.bwd0 = bwd0;
    T.fdone = P.fdone.p;
    T.bdone = P.bdone.p;
    T.abort = P.flow_abort.d;
    T.DF = P.diag32 ? P.DF.p : nullptr;
    T.hbuf = P.diag32 ? P.hbuf.p : nullptr;
    P.flow = T;
    P.flow_tasks = tasks;
    P.flow_ok = true;
}

// A single-launch apply whose wait ran out gave a wrong result: report it, and take the level
// launches from then on (graphs that captured the single launch are captured again)
void flow_check(cpg_prec_s &P) {
    if (!P.flow_abort.h || !*(volatile unsigned *)P.flow_abort.h) return;
    *(volatile unsigned *)P.flow_abort.h = 0;
    P.flow_enabled = false;
    P.gen = next_generation();
    CPG_CHECK(false, CPG_ERR_HIP,
              "block DILU apply: a row waited more than %d polls for the rows it reads; that apply's result is "
              "invalid, and the preconditioner now applies one launch per level",
              P.flow_spin);
}

template <class T, int BL, int WPR>
void launch_flow_w(cpg_prec_s &P, const double *in, double *out, hipStream_t s, const int *active) {
    const size_t lds = rowstream_lds<T, BL, WPR>(std::max(P.nbmax_f, P.nbmax_b));
    FlowTable A = P.flow;
    A.spin_max = P.flow_spin;
    const float *sc = P.precision == CPG_BF16 ? P.scale.p : nullptr;
    if constexpr (std::is_same<T, bf16>::value) {
        if (P.diag32) {
            // forward sweep, then backward sweep: one launch each
            const int nb = P.flow_tasks - A.bwd0;
            if (A.bwd0 > 0) {
                set_lds(k_dilu_flow<T, BL, WPR>, lds);
                hipLaunchKernelGGL((k_dilu_flow<T, BL, WPR>), dim3(A.bwd0), dim3(64 * WPR), lds, s, A, active,
                                   (const T *)Fp<T>(P), sc, in, P.ybuf.p, out);
                CPG_LAUNCH_CHECK();
            }
            const size_t lb = std::max(lds, rowstream_lds<float, BL, WPR>(1));
            set_lds(k_dilu_flow_d32b<T, BL, WPR>, lb);
            hipLaunchKernelGGL((k_dilu_flow_d32b<T, BL, WPR>), dim3(nb), dim3(64 * WPR), lb, s, A, active,
                               (const T *)Fp<T>(P), sc, in, P.ybuf.p, out);
            CPG_LAUNCH_CHECK();
            return;
        }
    }
    set_lds(k_dilu_flow<T, BL, WPR>, lds);
    launch_timed(k_dilu_flow<T, BL, WPR>, dim3(P.flow_tasks), dim3(64 * WPR), lds, s, A, active,
                 (const T *)Fp<T>(P), sc, in, P.ybuf.p, out);
    CPG_LAUNCH_CHECK();
}

template <class T, int BL>
void launch_flow_bl(cpg_prec_s &P, const double *in, double *out, hipStream_t s, const int *active) {
    switch (stream_wpr((int)sizeof(T), BL)) {
    case 1: launch_flow_w<T, BL, 1>(P, in, out, s, active); break;
    case 2: launch_flow_w<T, BL, 2>(P, in, out, s, active); break;
    default: launch_flow_w<T, BL, 4>(P, in, out, s, active); break;
    }
}

template <class T>
void launch_flow(cpg_prec_s &P, const double *in, double *out, hipStream_t s, const int *active) {
    switch (P.bl) {
    case 3: launch_flow_bl<T, 3>(P, in, out, s, active); return;
    case 6: launch_flow_bl<T, 6>(P, in, out, s, active); return;
    case 10: launch_flow_bl<T, 10>(P, in, out, s, active); return;
    case 12: launch_flow_bl<T, 12>(P, in, out, s, active); return;
    case 15: launch_flow_bl<T, 15>(P, in, out, s, active); return;
    case 24: launch_flow_bl<T, 24>(P, in, out, s, active); return;
    case 40: launch_flow_bl<T, 40>(P, in, out, s, active); return;
    default: launch_flow_bl<T, 60>(P, in, out, s, active); return;
    }
}

// out = M^-1 in (in may equal out)
template <class T, class XT>
void apply_t(cpg_prec_s &P, const XT *in, XT *out, hipStream_t s, const int *active) {
    const int bl = P.bl;
    const float *sc = P.precision == CPG_BF16 ? P.scale.p : nullptr;
    const T *F = Fp<T>(P);
    if (P.kind == CPG_PC_BJACOBI) {
        StreamArgs<XT> a;
        a.S = RowStream{nullptr, nullptr, nullptr, active};
        a.items = P.n;
        a.src1 = in;
        a.out = out;
        launch_rowstream<T, XT>(bl, a, F, sc, s);
        return;
    }
    constexpr bool x64 = std::is_same<XT, double>::value;
    if constexpr (x64) {
        flow_check(P);
        if (P.flow_ok && P.flow_enabled) {
            launch_flow<T>(P, in, out, s, active);
            return;
        }
    }
    double *y = P.ybuf.p;
    auto stream_of = [](const DevBuf<int> &rows, const std::vector<int> &row0, const DevBuf<int> &ptr, int lv, int r0,
                        const int *scol, const int *active) {
        RowStream S{row0[lv] >= 0 ? nullptr : rows.p + r0, ptr.p + r0, scol, active};
        S.row0 = row0[lv] >= 0 ? row0[lv] : 0;
        return S;
    };
    // Level 0 rows have no lower blocks, y_i = x_i; the packed sources of y there name x
    // (src2), so the level needs no launch for FP64 vectors.  FP32 vectors (the inner solver's)
    // run it as a copy into y, which their forward sources then read.
    // The last forward level has no upper blocks, so its backward step is x_i = D_i^-1 y_i; with
    // contiguous levels the first backward level is the same rows in the same order, and the
    // forward launch of the last level finishes them (FP64 vectors).
    const int nlf = (int)P.flev.size() - 1;
    const bool fuse = x64 && !P.diag32 && rowstream_compiled(bl) && nlf >= 2 && P.frow0[nlf - 1] >= 0 &&
                      P.brow0.size() > 0 &&
                      P.brow0[0] == P.frow0[nlf - 1] && P.blev[1] - P.blev[0] == P.flev[nlf] - P.flev[nlf - 1];
    for (size_t lv = x64 ? 1 : 0; lv + 1 < P.flev.size(); lv++) {
        const int r0 = P.flev[lv], nr = P.flev[lv + 1] - r0;
        if (!nr) continue;
        // y_i = x_i - sum_k L_ik y_k
        StreamArgs<double> a;
        a.S = stream_of(P.frows, P.frow0, P.fptr, (int)lv, r0, P.scol.p, active);
        a.items = nr;
        a.nbmax = P.nbmax_f;
        a.src1 = y;
        a.c1 = -1.0;
        a.c2 = -1.0;
        if constexpr (x64) {
            a.base64 = in;
            a.src2 = in;
        } else {
            a.base32 = in;
            a.src2 = y;
        }
        a.out = y;
        if constexpr (x64) {
            if (fuse && (int)lv == nlf - 1) {
                a.S.dptr = P.bptr.p + P.blev[0];
                a.out = out;
            }
        }
        launch_rowstream<T, double>(bl, a, F, sc, s);
    }
    if constexpr (x64 && std::is_same<T, bf16>::value) {
        if (P.diag32) {
            // per backward level: h_i = D_i^-1 y_i from the FP32 inverses (y_i = x_i where row i has
            // no lower blocks), then x_i = h_i - sum_j DU_ij x_j from the BF16 stream of DU blocks
            for (size_t lv = 0; lv + 1 < P.blev.size(); lv++) {
                const int r0 = P.blev[lv], nr = P.blev[lv + 1] - r0;
                if (!nr) continue;
                StreamArgs<double> h;
                h.S = RowStream{nullptr, nullptr, P.hcol.p, active};
                h.S.row0 = P.brow0[lv];   // contiguous levels: item w is row row0 + w, DF block row0 + w
                h.items = nr;
                h.src0 = y;
                h.src2 = in;
                h.c2 = 1.0;
                h.out = P.hbuf.p;
                launch_rowstream<float, double>(bl, h, P.DF.p, nullptr, s);
                StreamArgs<double> a;
                a.S = stream_of(P.brows, P.brow0, P.bptr, (int)lv, r0, P.scol.p, active);
                a.items = nr;
                a.nbmax = P.nbmax_b;
                a.src1 = out;
                a.c1 = -1.0;
                a.base64 = P.hbuf.p;
                a.out = out;
                launch_rowstream<T, double>(bl, a, F, sc, s);
            }
            return;
        }
    }
    for (size_t lv = fuse ? 1 : 0; lv + 1 < P.blev.size(); lv++) {
        const int r0 = P.blev[lv], nr = P.blev[lv + 1] - r0;
        if (!nr) continue;
        // x_i = D_i^-1 y_i - sum_j DU_ij x_j
        StreamArgs<XT> a;
        a.S = stream_of(P.brows, P.brow0, P.bptr, (int)lv, r0, P.scol.p, active);
        a.items = nr;
        a.nbmax = P.nbmax_b;
        a.src0 = y;
        a.src1 = out;
        a.c1 = -1.0;
        a.src2 = in;   // in_i = out_i until row i itself writes it (in-place applies)
        a.c2 = 1.0;
        a.out = out;
        launch_rowstream<T, XT>(bl, a, F, sc, s);
    }
}

// DILU / block-Jacobi sweep on vectors of type XT with the factors in their stored precision
template <class XT>
void apply_factors(cpg_prec_s &P, const XT *in, XT *out, hipStream_t s, const int *active = nullptr) {
    switch (P.precision) {
    case CPG_FP64: apply_t<double, XT>(P, in, out, s, active); break;
    case CPG_FP32: apply_t<float, XT>(P, in, out, s, active); break;
    default: apply_t<bf16, XT>(P, in, out, s, active); break;
    }
}

void matvec_f32(cpg_prec_s &P, const float *x, float *y, hipStream_t s) {
    StreamArgs<float> a;
    a.S = RowStream{nullptr, P.A->rowptr.p, P.A->colind.p};
    a.items = P.n;
    a.nbmax = P.A->nbmax;
    a.src1 = x;
    a.out = y;
    launch_rowstream<float, float>(P.bl, a, P.A32.p, nullptr, s);
}

// out <- inner GMRES(k) approximation of A^-1 in: FP32 Arnoldi on the FP32 copy of A, right-preconditioned by the DILU factors, zero initial guess.  A fixed sequence
// of launches on `s` with every scalar on the device.
void apply_inner(cpg_prec_s &P, const double *x, double *xout, hipStream_t s) {
    const size_t N = P.N;
    const int k = P.inner_k, ld = k + 1, nb = P.nchunk;
    float *Q = P.Q.p, *w = P.iw.p, *t = P.it.p;
    const int g256 = ceil_div(N, 256);
    hipLaunchKernelGGL(k_f32_load_norm, dim3(nb), dim3(IWG), 0, s, x, w, N, P.ipart.p);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_f32_reduce, dim3(1), dim3(IWG), 0, s, P.ipart.p, nb, 1, P.nrm.p, 1);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_f32_scale, dim3(g256), dim3(256), 0, s, w, Q, N, P.nrm.p);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_inner_init, dim3(1), dim3(64), 0, s, P.g.p, k, P.nrm.p);
    CPG_LAUNCH_CHECK();
    for (int i = 0; i < k; i++) {
        float *qi = Q + N * i, *qn = Q + N * (i + 1);
        apply_factors<float>(P, qi, t, s);                 // t = M^-1 q_i
        matvec_f32(P, t, w, s);                             // w = A t
        hipLaunchKernelGGL(k_f32_mdot, dim3(nb), dim3(IWG), 0, s, Q, N, i + 1, w, P.ipart.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_f32_reduce, dim3(i + 1), dim3(IWG), 0, s, P.ipart.p, nb, i + 1, P.H.p + (size_t)i * ld, 0);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_f32_orth, dim3(nb), dim3(IWG), 0, s, Q, N, i + 1, P.H.p + (size_t)i * ld, w, P.ipart.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_f32_reduce, dim3(1), dim3(IWG), 0, s, P.ipart.p, nb, 1, P.nrm.p, 1);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_f32_scale, dim3(g256), dim3(256), 0, s, w, qn, N, P.nrm.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_inner_givens, dim3(1), dim3(64), 0, s, P.H.p, ld, i, P.nrm.p, P.cs.p, P.sn.p, P.g.p);
        CPG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_inner_trisolve, dim3(1), dim3(64), 0, s, P.H.p, ld, k, P.g.p, P.y.p);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_f32_combine, dim3(nb), dim3(IWG), 0, s, Q, N, k, P.y.p, t);
    CPG_LAUNCH_CHECK();
    apply_factors<float>(P, t, t, s);                       // x = M^-1 Q y
    hipLaunchKernelGGL(k_f2d, dim3(g256), dim3(256), 0, s, t, xout, N);
    CPG_LAUNCH_CHECK();
}

// Contiguous levels: row i starts a new level when one of its lower neighbours lies in the
// current one.  For a colour-ordered matrix every level ends no earlier than the colour it
// starts in, so there are at most as many levels as colours, each a contiguous row range.
// Returns the level pointers (row indices).
std::vector<int> segments(int n, const std::vector<int> &lptr, const std::vector<int> &lcol) {
    std::vector<int> ptr{0};
    int start = 0;
    for (int i = 0; i < n; i++) {
        for (int q = lptr[i]; q < lptr[i + 1]; q++)
            if (lcol[q] >= start) {
                ptr.push_back(i);
                start = i;
                break;
            }
    }
    if (n > 0) ptr.push_back(n);
    return ptr;
}

// level schedule from the pattern: level(i) = 1 + max level(j) over dependencies
void levels(int n, const std::vector<int> &ptr, const std::vector<int> &cols, bool forward, std::vector<int> &lev_ptr,
            std::vector<int> &rows) {
    std::vector<int> lev(n, 0);
    int nl = 0;
    if (forward) {
        for (int i = 0; i < n; i++) {
            int l = 0;
            for (int q = ptr[i]; q < ptr[i + 1]; q++) l = std::max(l, lev[cols[q]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    } else {
        for (int i = n - 1; i >= 0; i--) {
            int l = 0;
            for (int q = ptr[i]; q < ptr[i + 1]; q++) l = std::max(l, lev[cols[q]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    }
    lev_ptr.assign(nl + 1, 0);
    for (int i = 0; i < n; i++) lev_ptr[lev[i] + 1]++;
    for (int l = 0; l < nl; l++) lev_ptr[l + 1] += lev_ptr[l];
    rows.assign(n, 0);
    std::vector<int> fill(lev_ptr.begin(), lev_ptr.end() - 1);
    for (int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
}

}  // namespace

extern "C" {

int cpg_prec_create(cpg_bsr A, int kind, int precision, cpg_prec *out) {
    return guard([&] {
        CPG_CHECK(A && out, CPG_ERR_ARG, "null argument");
        CPG_CHECK(kind == CPG_PC_NONE || kind == CPG_PC_BJACOBI || kind == CPG_PC_DILU || kind == CPG_PC_INNER_GMRES,
                  CPG_ERR_ARG, "bad kind %d", kind);
        CPG_CHECK(precision >= CPG_FP64 && precision <= CPG_BF16_D32, CPG_ERR_ARG, "bad precision %d", precision);
        CPG_CHECK(precision != CPG_BF16_D32 || kind == CPG_PC_DILU, CPG_ERR_ARG,
                  "BF16 factors with FP32 diagonal inverses are a block-DILU storage");
        auto *P = new cpg_prec_s;
        try {
            P->kind = kind;
            P->diag32 = precision == CPG_BF16_D32;
            if (P->diag32) precision = CPG_BF16;   // the packed stream's storage
            P->precision = precision;
            P->n = A->n;
            P->bl = A->bl;
            P->nnzb = A->nnzb;
            P->elem = elem_size(precision);
            const int n = A->n, bl = A->bl;
            const size_t bb = (size_t)bl * bl;
            const auto &rp = A->h_rowptr;
            const auto &ci = A->h_colind;
            std::vector<int> diag(n, -1), lptr(n + 1, 0), uptr(n + 1, 0), lslot, lcol, ltrans, uslot, ucol, urow;
            for (int i = 0; i < n; i++) {
                for (int s = rp[i]; s < rp[i + 1]; s++) {
                    const int j = ci[s];
                    if (j == i) {
                        if (diag[i] < 0) diag[i] = s;
                    } else if (j < i) {
                        lslot.push_back(s);
                        lcol.push_back(j);
                        int st = -1;
                        for (int s2 = rp[j]; s2 < rp[j + 1]; s2++)
                            if (ci[s2] == i) { st = s2; break; }
                        CPG_CHECK(!(kind == CPG_PC_DILU || kind == CPG_PC_INNER_GMRES) || st >= 0, CPG_ERR_ARG,
                                  "pattern is not structurally symmetric at (%d,%d)", i, j);
                        ltrans.push_back(st);
                    } else {
                        uslot.push_back(s);
                        ucol.push_back(j);
                        urow.push_back(i);
                    }
                }
                CPG_CHECK(kind == CPG_PC_NONE || diag[i] >= 0, CPG_ERR_ARG, "row %d has no diagonal block", i);
                lptr[i + 1] = (int)lslot.size();
                uptr[i + 1] = (int)uslot.size();
            }
            P->diag.upload(diag);
            P->lptr.upload(lptr);
            P->uptr.upload(uptr);
            P->lslot.upload(lslot);
            P->lcol.upload(lcol);
            P->ltrans.upload(ltrans);
            P->uslot.upload(uslot);
            P->ucol.upload(ucol);
            P->urow.upload(urow);
            P->nup = (int)urow.size();
            {
                std::vector<int> ur;
                for (int i = 0; i < n; i++)
                    if (uptr[i + 1] > uptr[i]) ur.push_back(i);
                P->nurows = (int)ur.size();
                if (!ur.empty()) P->urows.upload(ur);
            }
            std::vector<int> frows, brows;
            levels(n, lptr, lcol, true, P->flev, frows);
            const std::vector<int> seg = segments(n, lptr, lcol);
            if (P->dilu() && seg.size() <= P->flev.size()) {
                // contiguous levels; the backward sweep takes them from the last (valid for the
                // structurally symmetric pattern checked above: an upper neighbour of a row lies
                // in a later level)
                const int nl = (int)seg.size() - 1;
                P->flev = seg;
                for (int i = 0; i < n; i++) frows[i] = i;
                P->blev.assign(1, 0);
                brows.clear();
                for (int l = nl - 1; l >= 0; l--) {
                    for (int i = seg[l]; i < seg[l + 1]; i++) brows.push_back(i);
                    P->blev.push_back((int)brows.size());
                }
            } else {
                levels(n, uptr, ucol, false, P->blev, brows);
            }
            auto first_rows = [](const std::vector<int> &lev, const std::vector<int> &rows) {
                std::vector<int> r0;
                for (size_t l = 0; l + 1 < lev.size(); l++) {
                    bool cont = lev[l + 1] > lev[l];
                    for (int p = lev[l]; cont && p < lev[l + 1]; p++) cont = rows[p] == rows[lev[l]] + (p - lev[l]);
                    r0.push_back(cont ? rows[lev[l]] : -1);
                }
                return r0;
            };
            P->frow0 = first_rows(P->flev, frows);
            P->brow0 = first_rows(P->blev, brows);
            P->frows.upload(frows);
            P->brows.upload(brows);
            if (P->dilu()) {
                // packed streams in sweep order
                CPG_CHECK(n < kSrc2Tag - 1, CPG_ERR_UNSUPPORTED, "%d block rows: too many for the packed sources", n);
                std::vector<int> pos(A->nnzb, -1), fptr(n + 1, 0), bptr(n + 1, 0), scol;
                // a lower neighbour k with no lower blocks of its own has y_k = x_k: read x
                auto src_of = [&](int k) { return lptr[k + 1] > lptr[k] ? k : -1 - (kSrc2Tag + k); };
                int cur = 0;
                for (int p = 0; p < n; p++) {
                    const int i = frows[p];
                    fptr[p] = cur;
                    for (int q = lptr[i]; q < lptr[i + 1]; q++) {
                        pos[lslot[q]] = cur++;
                        scol.push_back(src_of(lcol[q]));
                    }
                    P->nbmax_f = std::max(P->nbmax_f, lptr[i + 1] - lptr[i]);
                }
                fptr[n] = cur;
                for (int p = 0; p < n; p++) {
                    const int i = brows[p];
                    bptr[p] = cur;
                    if (!P->diag32) {
                        pos[diag[i]] = cur++;
                        scol.push_back(lptr[i + 1] > lptr[i] ? -1 - i : -1 - (kSrc2Tag + i));
                    }
                    for (int q = uptr[i]; q < uptr[i + 1]; q++) {
                        pos[uslot[q]] = cur++;
                        scol.push_back(ucol[q]);
                    }
                    P->nbmax_b = std::max(P->nbmax_b, (P->diag32 ? 0 : 1) + uptr[i + 1] - uptr[i]);
                }
                bptr[n] = cur;
                if (P->diag32) {
                    // the BF16 copies of D^-1 (written by the factor, not read by the apply) after the
                    // streams; the apply's D^-1 products read DF in row order, with these sources
                    std::vector<int> hcol(n);
                    for (int i = 0; i < n; i++) {
                        pos[diag[i]] = cur++;
                        hcol[i] = lptr[i + 1] > lptr[i] ? -1 - i : -1 - (kSrc2Tag + i);
                    }
                    P->hcol.upload(hcol);
                    P->h_hcol = hcol;
                    bool contiguous = true;
                    for (int r : P->brow0) contiguous = contiguous && r >= 0;
                    for (int r : P->frow0) contiguous = contiguous && r >= 0;
                    CPG_CHECK(contiguous, CPG_ERR_UNSUPPORTED,
                              "BF16 factors with FP32 diagonal inverses need contiguous colour levels "
                              "(a colour-ordered matrix)");
                }
                // every slot is the diagonal, a lower or an upper block of its row
                for (int q = 0; q < A->nnzb; q++)
                    CPG_CHECK(pos[q] >= 0, CPG_ERR_ARG, "slot %d is a second diagonal block of its row", q);
                P->pos.upload(pos);
                P->fptr.upload(fptr);
                P->bptr.upload(bptr);
                P->scol.upload(scol);
                P->h_fptr = fptr;
                P->h_bptr = bptr;
                P->h_scol = scol;
            }
            P->h_frows = frows;
            P->h_brows = brows;
            P->status.alloc(1);
            if (P->dilu()) {
                P->F.alloc((size_t)A->nnzb * bb * P->elem);
                if (precision == CPG_BF16) P->scale.alloc(A->nnzb);
                P->DD64.alloc((size_t)n * bb);
                P->ybuf.alloc((size_t)n * bl);   // FP64 intermediate for every factor precision
                if (P->diag32) {
                    P->DF.alloc((size_t)n * bb);
                    P->hbuf.alloc((size_t)n * bl);
                }
            }
            P->h_diag = diag;
            if (kind == CPG_PC_INNER_GMRES) {
                CPG_CHECK(precision != CPG_FP64, CPG_ERR_ARG, "the inner FP32 GMRES needs FP32 or BF16 factors");
                P->N = (size_t)n * bl;
                P->nchunk = std::max(1, ceil_div(P->N, ICH));
                P->A32.alloc((size_t)A->nnzb * bb);
                P->iw.alloc(P->N);
                P->it.alloc(P->N);
                P->nrm.alloc(2);
                // basis, Hessenberg and rotations are sized by cpg_prec_set_inner
            } else if (kind == CPG_PC_BJACOBI) {
                P->F.alloc((size_t)n * bb * P->elem);
                if (precision == CPG_BF16) P->scale.alloc(n);
            }
            flow_setup(*P);
        } catch (...) {
            delete P;
            throw;
        }
        *out = P;
    });
}

int cpg_prec_destroy(cpg_prec P) {
    return guard([&] { delete P; });
}

int cpg_prec_apply_to(cpg_prec P, const double *in, double *out, void *stream) {
    return guard([&] {
        CPG_CHECK(P, CPG_ERR_ARG, "null argument");
        if (P->n == 0) return;
        CPG_CHECK(in && out && in != out, CPG_ERR_ARG, "in and out must be distinct device vectors");
        prec_apply_to(P, in, out, S(stream));
    });
}

int cpg_prec_set_flow(cpg_prec P, int enable, int spin_limit) {
    return guard([&] {
        CPG_CHECK(P, CPG_ERR_ARG, "null argument");
        if (enable >= 0 && (enable != 0) != P->flow_enabled) {
            P->flow_enabled = enable != 0;
            P->gen = next_generation();   // graphs that captured the other form are captured again
        }
        if (spin_limit >= 0) P->flow_spin = spin_limit;
    });
}

int cpg_prec_flow_state(cpg_prec P, int *in_use) {
    return guard([&] {
        CPG_CHECK(P && in_use, CPG_ERR_ARG, "null argument");
        *in_use = P->flow_ok && P->flow_enabled ? 1 : 0;
    });
}

int cpg_prec_levels(cpg_prec P, int *n_levels) {
    return guard([&] {
        CPG_CHECK(P && n_levels, CPG_ERR_ARG, "null argument");
        *n_levels = (int)P->flev.size() - 1;
    });
}

int cpg_prec_factor(cpg_prec P, cpg_bsr A, void *stream) {
    return guard([&] {
        CPG_CHECK(P && A, CPG_ERR_ARG, "null argument");
        CPG_CHECK(A->n == P->n && A->bl == P->bl && A->nnzb == P->nnzb, CPG_ERR_ARG, "matrix does not match");
        if (P->kind == CPG_PC_NONE || P->n == 0) return;
        hipStream_t s = S(stream);
        CPG_HIP(hipMemsetAsync(P->status.p, 0, sizeof(int), s));
        switch (P->precision) {
        case CPG_FP64: factor_t<double>(*P, *A, s); break;
        case CPG_FP32: factor_t<float>(*P, *A, s); break;
        default: factor_t<bf16>(*P, *A, s); break;
        }
        if (P->diag32) {   // the FP64 D^-1 of every row, rounded once to FP32
            hipLaunchKernelGGL(k_d2f, dim3(ceil_div(P->DF.n, 256)), dim3(256), 0, s, P->DD64.p, P->DF.p, P->DF.n);
            CPG_LAUNCH_CHECK();
        }
        if (P->kind == CPG_PC_INNER_GMRES) {
            if (!P->Q.p) {
                int st2 = cpg_prec_set_inner(P, P->inner_k);
                if (st2) throw Error{st2};
            }
            P->A = A;
            hipLaunchKernelGGL(k_d2f, dim3(ceil_div(A->blocks.n, 256)), dim3(256), 0, s, A->blocks.p, P->A32.p,
                               A->blocks.n);
            CPG_LAUNCH_CHECK();
        }
        int st = 0;
        CPG_HIP(hipMemcpyAsync(&st, P->status.p, sizeof(int), hipMemcpyDeviceToHost, s));
        CPG_HIP(hipStreamSynchronize(s));
        CPG_CHECK(st == 0, CPG_ERR_SINGULAR, "zero pivot in a diagonal block inverse");
    });
}

int cpg_prec_apply(cpg_prec P, double *x, void *stream) {
    return guard([&] {
        CPG_CHECK(P, CPG_ERR_ARG, "null argument");
        if (P->kind == CPG_PC_NONE || P->n == 0) return;   // zero-length vectors may be null
        CPG_CHECK(x, CPG_ERR_ARG, "null argument");
        hipStream_t s = S(stream);
        if (P->kind == CPG_PC_INNER_GMRES) {
            CPG_CHECK(P->A, CPG_ERR_ARG, "inner GMRES preconditioner applied before its first factor");
            apply_inner(*P, x, x, s);
            return;
        }
        apply_factors<double>(*P, x, x, s);
    });
}

}  // extern "C"

namespace cpg {
void prec_flow_check(cpg_prec P) {
    if (P) flow_check(*P);
}

// out = M^-1 in, in and out distinct (used by flexible GMRES for z_i = M^-1 v_i)
void prec_apply_to(cpg_prec P, const double *in, double *out, hipStream_t s, const int *active) {
    if (P->kind == CPG_PC_NONE || P->n == 0) {
        CPG_HIP(hipMemcpyAsync(out, in, (size_t)P->n * P->bl * sizeof(double), hipMemcpyDeviceToDevice, s));
        return;
    }
    if (P->kind == CPG_PC_INNER_GMRES) {
        CPG_CHECK(P->A, CPG_ERR_ARG, "inner GMRES preconditioner applied before its first factor");
        apply_inner(*P, in, out, s);
        return;
    }
    apply_factors<double>(*P, in, out, s, active);
}
}  // namespace cpg

extern "C" {

int cpg_prec_set_inner(cpg_prec P, int iterations) {
    return guard([&] {
        CPG_CHECK(P && iterations > 0 && iterations <= 200, CPG_ERR_ARG, "bad inner iteration count");
        CPG_CHECK(P->kind == CPG_PC_INNER_GMRES, CPG_ERR_ARG, "not an inner-GMRES preconditioner");
        const int k = iterations;
        P->inner_k = k;
        P->gen = next_generation();
        P->Q.alloc(P->N * (k + 1));
        P->ipart.alloc((size_t)P->nchunk * (k + 1));
        P->H.alloc((size_t)(k + 1) * k);
        P->g.alloc(k + 1);
        P->cs.alloc(k);
        P->sn.alloc(k);
        P->y.alloc(k);
    });
}

int cpg_prec_download(cpg_prec P, double *host_blocks) {
    return guard([&] {
        CPG_CHECK(P && host_blocks, CPG_ERR_ARG, "null argument");
        if (P->kind == CPG_PC_NONE) return;
        const size_t bb = (size_t)P->bl * P->bl;
        const size_t nblk = P->dilu() ? P->nnzb : P->n;
        const size_t cnt = nblk * bb;
        DevBuf<double> tmp(cnt);
        const float *sc = P->precision == CPG_BF16 ? P->scale.p : nullptr;
        const int *pos = P->dilu() ? P->pos.p : nullptr;
        const int g = ceil_div(cnt, 256);
        switch (P->precision) {
        case CPG_FP64: hipLaunchKernelGGL(k_to_fp64<double>, dim3(g), dim3(256), 0, nullptr, Fp<double>(*P), sc, pos, bb, cnt, tmp.p); break;
        case CPG_FP32: hipLaunchKernelGGL(k_to_fp64<float>, dim3(g), dim3(256), 0, nullptr, Fp<float>(*P), sc, pos, bb, cnt, tmp.p); break;
        default: hipLaunchKernelGGL(k_to_fp64<bf16>, dim3(g), dim3(256), 0, nullptr, Fp<bf16>(*P), sc, pos, bb, cnt, tmp.p); break;
        }
        CPG_LAUNCH_CHECK();
        CPG_HIP(hipMemcpy(host_blocks, tmp.p, cnt * sizeof(double), hipMemcpyDeviceToHost));
        if (P->diag32) {   // the diagonal slots hold the FP32 inverses the apply uses
            std::vector<float> df(P->DF.n);
            CPG_HIP(hipMemcpy(df.data(), P->DF.p, df.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (int i = 0; i < P->n; i++)
                for (size_t e = 0; e < bb; e++) host_blocks[(size_t)P->h_diag[i] * bb + e] = df[(size_t)i * bb + e];
        }
    });
}

}  // extern "C"

namespace cpg {

void prec_factor_coupled(cpg_prec P, const CoupledBlocks &C, hipStream_t s) {
    CPG_CHECK(!P->diag32, CPG_ERR_UNSUPPORTED, "the coupled factor stores every block in one precision");
    CPG_CHECK(P && P->kind == CPG_PC_DILU && P->bl == CBL && C.bs > 0 && CSB % C.bs == 0, CPG_ERR_UNSUPPORTED,
              "the coupled DILU takes 3 stages of 40 x 40 blocks");
    CPG_CHECK(P->n == C.n && P->nnzb == C.nnzb, CPG_ERR_ARG, "preconditioner pattern (%d rows, %d slots) does not "
              "match the stage Jacobians (%d, %d)", P->n, P->nnzb, C.n, C.nnzb);
    if (P->n == 0) return;
    CoupledSrc src{C.J[0], C.J[1], C.J[2], C.mass, {}, C.bs};
    for (int j = 0; j < 9; j++) src.coef[j] = C.coef[j];
    CPG_HIP(hipMemsetAsync(P->status.p, 0, sizeof(int), s));
    switch (P->precision) {
    case CPG_FP64: factor_coupled_t<double>(*P, src, s); break;
    case CPG_FP32: factor_coupled_t<float>(*P, src, s); break;
    default: factor_coupled_t<bf16>(*P, src, s); break;
    }
    int st = 0;
    CPG_HIP(hipMemcpyAsync(&st, P->status.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CPG_HIP(hipStreamSynchronize(s));
    CPG_CHECK(st == 0, CPG_ERR_SINGULAR, "zero pivot in a coupled diagonal block inverse");
}

}  // namespace cpg
