// Block sparse row matrix in HBM and its matvec.
//
// Replaces BlockMatrix storage and BlockMatrix::matvec (src/BlockMatrix.cpp:10-49),
// which does one BLAS dgemv per block on a heap copy of x.  A row's blocks are
// contiguous (nnzb_row * bl * bl doubles, column-major), so the matvec is the row-stream
// kernel of rowstream.hpp with the BSR arrays as its stream: one workgroup per block row,
// x gathered once into LDS, 16-byte column loads, fixed-order LDS reduction.
#include "cpg_internal.hpp"
#include "rowstream.hpp"

#include <algorithm>
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

// ---- row-stream launch shape --------------------------------------------------------
// Waves per row of the compiled row-stream kernels, per storage element size (8, 4, 2) and
// block size.  Defaults come from timing sweeps on MI355X (cpg_stream_autotune re-measures
// them): four waves, two for BF16 blocks (the BF16 DILU apply at 65,536 rows of 40x40 blocks:
// 304 us against 316 us with four).  Results do not depend on the choice.
namespace {
struct WprEntry {
    int elem, bl, wpr;
};
std::vector<WprEntry> &wpr_table() {
    static std::vector<WprEntry> t;
    return t;
}
}  // namespace

int stream_wpr(int elem_bytes, int bl) {
    for (const auto &e : wpr_table())
        if (e.elem == elem_bytes && e.bl == bl) return e.wpr;
    return elem_bytes == 2 ? 2 : 4;
}

void set_stream_wpr(int elem_bytes, int bl, int wpr) {
    for (auto &e : wpr_table())
        if (e.elem == elem_bytes && e.bl == bl) {
            e.wpr = wpr;
            return;
        }
    wpr_table().push_back({elem_bytes, bl, wpr});
}

void bsr_matvec(const cpg_bsr_s &A, const double *x, double *y, hipStream_t s, const int *active) {
    if (A.n == 0) return;
    StreamArgs<double> a;
    a.S = RowStream{nullptr, A.rowptr.p, A.colind.p, active};
    a.items = A.n;
    a.nbmax = A.nbmax;
    a.src1 = x;
    a.out = y;
    launch_rowstream<double, double>(A.bl, a, A.blocks.p, nullptr, s);
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

// out = a x + b y (y null: out = a x); out may alias x or y (each element is read before it is written)
__global__ void k_axpby(size_t n, double a, const double *x, double b, const double *y, double *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = y ? a * x[i] + b * y[i] : a * x[i];
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
            for (int i = 0; i < n_rows; i++) A->nbmax = std::max(A->nbmax, rowptr[i + 1] - rowptr[i]);
            A->rowptr.upload(A->h_rowptr);
            A->colind.upload(A->h_colind);
            A->blocks.alloc((size_t)nnzb * bl * bl);
            if (A->blocks.n) CPG_HIP(hipMemset(A->blocks.p, 0, A->blocks.n * sizeof(double)));
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

int cpg_vec_axpby(unsigned long long n, double a, const double *x, double b, const double *y, double *out,
                  void *stream) {
    return guard([&] {
        if (!n) return;
        CPG_CHECK(x && out, CPG_ERR_ARG, "null argument");
        hipLaunchKernelGGL(k_axpby, dim3(ceil_div((size_t)n, 256)), dim3(256), 0, S(stream), (size_t)n, a, x, b, y,
                           out);
        CPG_LAUNCH_CHECK();
    });
}

int cpg_bsr_matvec(cpg_bsr A, const double *x, double *y, void *stream) {
    return guard([&] {
        CPG_CHECK(A, CPG_ERR_ARG, "null argument");
        if (A->n == 0) return;   // zero-length vectors may be null
        CPG_CHECK(x && y, CPG_ERR_ARG, "null argument");
        CPG_CHECK(x != y, CPG_ERR_ARG, "x and y alias");
        bsr_matvec(*A, x, y, S(stream));
    });
}

static int elem_of(int precision) { return precision == CPG_FP64 ? 8 : precision == CPG_FP32 ? 4 : 2; }

int cpg_stream_set_wpr(int precision, int bl, int wpr) {
    return guard([&] {
        CPG_CHECK(precision >= CPG_FP64 && precision <= CPG_BF16, CPG_ERR_ARG, "bad precision %d", precision);
        CPG_CHECK(wpr == 1 || wpr == 2 || wpr == 4, CPG_ERR_ARG, "waves per row must be 1, 2 or 4 (got %d)", wpr);
        CPG_CHECK(bl > 0, CPG_ERR_ARG, "bad block size %d", bl);
        set_stream_wpr(elem_of(precision), bl, wpr);
    });
}

int cpg_stream_get_wpr(int precision, int bl, int *wpr) {
    return guard([&] {
        CPG_CHECK(wpr && precision >= CPG_FP64 && precision <= CPG_BF16, CPG_ERR_ARG, "bad argument");
        *wpr = stream_wpr(elem_of(precision), bl);
    });
}

// Times the row stream of A's pattern (a matvec with blocks stored in `precision`) for each
// launch shape and keeps the fastest.  The block values do not change the timing, so the
// reduced-precision streams run over zero-filled blocks.
int cpg_stream_autotune(cpg_bsr A, int precision, int reps, int *best_wpr, float *ms, void *stream) {
    return guard([&] {
        CPG_CHECK(A && precision >= CPG_FP64 && precision <= CPG_BF16 && reps > 0, CPG_ERR_ARG, "bad argument");
        if (A->n == 0) return;
        hipStream_t s = S(stream);
        const int elem = elem_of(precision);
        const size_t N = (size_t)A->n * A->bl;
        DevBuf<double> x(N), y(N);
        CPG_HIP(hipMemsetAsync(x.p, 0, N * sizeof(double), s));
        DevBuf<unsigned char> F;
        DevBuf<float> sc;
        if (precision != CPG_FP64) {
            F.alloc(A->blocks.n * elem);
            CPG_HIP(hipMemsetAsync(F.p, 0, F.n, s));
            if (precision == CPG_BF16) {
                sc.alloc(A->nnzb);
                CPG_HIP(hipMemsetAsync(sc.p, 0, sc.n * sizeof(float), s));
            }
        }
        StreamArgs<double> a;
        a.S = RowStream{nullptr, A->rowptr.p, A->colind.p};
        a.items = A->n;
        a.nbmax = A->nbmax;
        a.src1 = x.p;
        a.out = y.p;
        auto run = [&](int wpr) {
            if (precision == CPG_FP64) launch_rowstream<double, double>(A->bl, a, A->blocks.p, nullptr, s, wpr);
            else if (precision == CPG_FP32)
                launch_rowstream<float, double>(A->bl, a, reinterpret_cast<const float *>(F.p), nullptr, s, wpr);
            else launch_rowstream<bf16, double>(A->bl, a, reinterpret_cast<const bf16 *>(F.p), sc.p, s, wpr);
        };
        hipEvent_t e0, e1;
        CPG_HIP(hipEventCreate(&e0));
        CPG_HIP(hipEventCreate(&e1));
        const int cand[3] = {1, 2, 4};
        float best = 1e30f;
        int bw = 4;
        for (int c = 0; c < 3; c++) {
            run(cand[c]);
            CPG_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) run(cand[c]);
            CPG_HIP(hipEventRecord(e1, s));
            CPG_HIP(hipEventSynchronize(e1));
            float t = 0.0f;
            CPG_HIP(hipEventElapsedTime(&t, e0, e1));
            t /= (float)reps;
            if (ms) ms[c] = t;
            if (t < best) { best = t; bw = cand[c]; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        set_stream_wpr(elem, A->bl, bw);
        if (best_wpr) *best_wpr = bw;
    });
}

}  // extern "C"
