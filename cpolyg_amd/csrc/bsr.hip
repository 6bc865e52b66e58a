// Block sparse row matrix in HBM and its matvec.
//
// Replaces BlockMatrix storage and BlockMatrix::matvec (src/BlockMatrix.cpp:10-49),
// which does one BLAS dgemv per block on a heap copy of x.  Here one workgroup
// owns one block row: the row's blocks are contiguous (nnzb_row * bl * bl
// doubles) and column-major, so the row is read as a stream of bl-long columns.
// A group of bl/VW lanes covers one column with VW-wide loads; G = 256/(bl/VW)
// groups walk the row's columns with stride G; the G partial vectors are summed
// through LDS.  x is gathered one double per lane (bl/VW lanes share it).
#include "cpg_internal.hpp"

#include <cstdarg>
#include <cstring>

namespace cpg {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

// ------------------------------------------------------------------------------------
template <int BL, int VW>
__global__ __launch_bounds__(256) void k_bsr_matvec(int bl_rt, const int *__restrict__ rowptr,
                                                    const int *__restrict__ colind,
                                                    const double *__restrict__ A, const double *__restrict__ x,
                                                    double *__restrict__ y) {
    extern __shared__ double red[];
    const int bl = BL > 0 ? BL : bl_rt;
    const int tpc = bl / VW;
    const int G = 256 / tpc;
    const int t = threadIdx.x;
    const int g = t / tpc;
    const int rr = (t - g * tpc) * VW;
    const int i = blockIdx.x;
    const int s0 = rowptr[i], s1 = rowptr[i + 1];
    const int ncols = (s1 - s0) * bl;
    double acc[VW];
#pragma unroll
    for (int v = 0; v < VW; v++) acc[v] = 0.0;
    if (g < G) {
        const double *base = A + (size_t)s0 * bl * bl + rr;
#pragma unroll 4
        for (int cg = g; cg < ncols; cg += G) {
            const int blk = cg / bl;
            const int col = cg - blk * bl;
            const double xv = x[(size_t)colind[s0 + blk] * bl + col];
            if constexpr (VW == 2) {
                const double2 a = *reinterpret_cast<const double2 *>(base + (size_t)cg * bl);
                acc[0] = fma(a.x, xv, acc[0]);
                acc[1] = fma(a.y, xv, acc[1]);
            } else {
                acc[0] = fma(base[(size_t)cg * bl], xv, acc[0]);
            }
        }
#pragma unroll
        for (int v = 0; v < VW; v++) red[g * bl + rr + v] = acc[v];
    }
    __syncthreads();
    if (t < bl) {
        double s = 0.0;
        for (int gg = 0; gg < G; gg++) s += red[gg * bl + t];
        y[(size_t)i * bl + t] = s;
    }
}

template <int BL, int VW>
static void launch_matvec(const cpg_bsr_s &A, const double *x, double *y, hipStream_t s) {
    const int bl = A.bl;
    const int tpc = bl / VW;
    const int G = 256 / tpc;
    size_t lds = (size_t)G * bl * sizeof(double);
    hipLaunchKernelGGL((k_bsr_matvec<BL, VW>), dim3(A.n), dim3(256), lds, s, bl, A.rowptr.p, A.colind.p,
                       A.blocks.p, x, y);
    CPG_LAUNCH_CHECK();
}

void bsr_matvec(const cpg_bsr_s &A, const double *x, double *y, hipStream_t s) {
    if (A.n == 0) return;
    switch (A.bl) {
    case 40: launch_matvec<40, 2>(A, x, y, s); break;
    case 60: launch_matvec<60, 2>(A, x, y, s); break;
    case 120: launch_matvec<120, 2>(A, x, y, s); break;
    case 10: launch_matvec<10, 2>(A, x, y, s); break;
    default:
        CPG_CHECK(A.bl <= 256, CPG_ERR_UNSUPPORTED, "block size %d > 256", A.bl);
        if (A.bl % 2 == 0) launch_matvec<0, 2>(A, x, y, s);
        else launch_matvec<0, 1>(A, x, y, s);
    }
}

// ------------------------------------------------------------------------------------
__global__ void k_fill(double *x, double v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_scale(double *x, double v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= v;
}

void vec_fill(double *x, double v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill, dim3(ceil_div(n, 256)), dim3(256), 0, s, x, v, n);
    CPG_LAUNCH_CHECK();
}

}  // namespace cpg

using namespace cpg;

// ---- C ABI -------------------------------------------------------------------------
extern "C" {

const char *cpg_last_error(void) { return cpg::g_err; }

const char *cpg_status_string(int st) {
    switch (st) {
    case CPG_OK: return "ok";
    case CPG_ERR_ARG: return "bad argument";
    case CPG_ERR_HIP: return "HIP runtime error";
    case CPG_ERR_OOM: return "out of device memory";
    case CPG_ERR_SINGULAR: return "singular block";
    case CPG_ERR_NONFINITE: return "non-finite value";
    case CPG_ERR_NOT_CONVERGED: return "not converged";
    case CPG_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
    }
}

int cpg_version(void) { return 100; }

int cpg_device_alloc(void **ptr, unsigned long long bytes) {
    return guard([&] {
        CPG_CHECK(ptr, CPG_ERR_ARG, "null ptr");
        *ptr = nullptr;
        if (bytes) CPG_HIP(hipMalloc(ptr, bytes));
    });
}

int cpg_device_free(void *ptr) {
    return guard([&] { if (ptr) CPG_HIP(hipFree(ptr)); });
}

int cpg_memcpy_h2d(void *dst, const void *src, unsigned long long bytes, void *stream) {
    return guard([&] { CPG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream))); });
}

int cpg_memcpy_d2h(void *dst, const void *src, unsigned long long bytes, void *stream) {
    return guard([&] {
        CPG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
        CPG_HIP(hipStreamSynchronize(S(stream)));
    });
}

int cpg_memcpy_d2d(void *dst, const void *src, unsigned long long bytes, void *stream) {
    return guard([&] { CPG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream))); });
}

int cpg_stream_sync(void *stream) {
    return guard([&] { CPG_HIP(hipStreamSynchronize(S(stream))); });
}

int cpg_bsr_create(int n_rows, int bl, int nnzb, const int *rowptr, const int *colind, cpg_bsr *out) {
    return guard([&] {
        CPG_CHECK(out && rowptr && (colind || nnzb == 0), CPG_ERR_ARG, "null argument");
        CPG_CHECK(n_rows >= 0 && bl > 0 && nnzb >= 0, CPG_ERR_ARG, "bad sizes n=%d bl=%d nnzb=%d", n_rows, bl, nnzb);
        CPG_CHECK(rowptr[0] == 0 && rowptr[n_rows] == nnzb, CPG_ERR_ARG, "rowptr does not span nnzb");
        for (int i = 0; i < n_rows; i++)
            CPG_CHECK(rowptr[i + 1] >= rowptr[i], CPG_ERR_ARG, "rowptr not monotone at %d", i);
        for (int s = 0; s < nnzb; s++)
            CPG_CHECK(colind[s] >= 0 && colind[s] < n_rows, CPG_ERR_ARG, "colind[%d]=%d out of range", s, colind[s]);
        auto *A = new cpg_bsr_s;
        try {
            A->n = n_rows;
            A->bl = bl;
            A->nnzb = nnzb;
            A->h_rowptr.assign(rowptr, rowptr + n_rows + 1);
            A->h_colind.assign(colind, colind + nnzb);
            A->rowptr.upload(A->h_rowptr);
            A->colind.upload(A->h_colind);
            A->blocks.alloc((size_t)nnzb * bl * bl);
            CPG_HIP(hipMemset(A->blocks.p, 0, A->blocks.n * sizeof(double)));
        } catch (...) {
            delete A;
            throw;
        }
        *out = A;
    });
}

int cpg_bsr_destroy(cpg_bsr A) {
    return guard([&] { delete A; });
}

int cpg_bsr_info(cpg_bsr A, int *n_rows, int *bl, int *nnzb) {
    return guard([&] {
        CPG_CHECK(A, CPG_ERR_ARG, "null matrix");
        if (n_rows) *n_rows = A->n;
        if (bl) *bl = A->bl;
        if (nnzb) *nnzb = A->nnzb;
    });
}

int cpg_bsr_upload(cpg_bsr A, const double *host_blocks, void *stream) {
    return guard([&] {
        CPG_CHECK(A && host_blocks, CPG_ERR_ARG, "null argument");
        A->blocks.upload(host_blocks, A->blocks.n, S(stream));
    });
}

int cpg_bsr_download(cpg_bsr A, double *host_blocks, void *stream) {
    return guard([&] {
        CPG_CHECK(A && host_blocks, CPG_ERR_ARG, "null argument");
        CPG_HIP(hipMemcpyAsync(host_blocks, A->blocks.p, A->blocks.n * sizeof(double), hipMemcpyDeviceToHost,
                               S(stream)));
        CPG_HIP(hipStreamSynchronize(S(stream)));
    });
}

int cpg_bsr_device_blocks(cpg_bsr A, double **dev_blocks) {
    return guard([&] {
        CPG_CHECK(A && dev_blocks, CPG_ERR_ARG, "null argument");
        *dev_blocks = A->blocks.p;
    });
}

int cpg_bsr_scale(cpg_bsr A, double s, void *stream) {
    return guard([&] {
        CPG_CHECK(A, CPG_ERR_ARG, "null matrix");
        if (!A->blocks.n) return;
        hipLaunchKernelGGL(k_scale, dim3(ceil_div(A->blocks.n, 256)), dim3(256), 0, S(stream), A->blocks.p, s,
                           A->blocks.n);
        CPG_LAUNCH_CHECK();
    });
}

int cpg_bsr_matvec(cpg_bsr A, const double *x, double *y, void *stream) {
    return guard([&] {
        CPG_CHECK(A && x && y, CPG_ERR_ARG, "null argument");
        CPG_CHECK(x != y, CPG_ERR_ARG, "x and y alias");
        bsr_matvec(*A, x, y, S(stream));
    });
}

}  // extern "C"
