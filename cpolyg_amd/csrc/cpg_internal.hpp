// Internal definitions shared by the translation units of libcpolyg_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cpolyg.h"

namespace cpg {

// ---- error plumbing ----------------------------------------------------------------
void set_error(const char *fmt, ...);

struct Error {
    int status;
};

#define CPG_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            ::cpg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
            throw ::cpg::Error{e_ == hipErrorOutOfMemory ? CPG_ERR_OOM : CPG_ERR_HIP};          \
        }                                                                                      \
    } while (0)

#define CPG_CHECK(cond, status, ...)        \
    do {                                    \
        if (!(cond)) {                      \
            ::cpg::set_error(__VA_ARGS__);  \
            throw ::cpg::Error{status};     \
        }                                   \
    } while (0)

#define CPG_LAUNCH_CHECK() CPG_HIP(hipGetLastError())

template <class F>
int guard(F &&f) {
    try {
        f();
        return CPG_OK;
    } catch (const Error &e) {
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return CPG_ERR_OOM;
    } catch (...) {
        set_error("unexpected C++ exception");
        return CPG_ERR_HIP;
    }
}

// ---- device buffers ----------------------------------------------------------------
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) CPG_HIP(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        CPG_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &h, hipStream_t s = nullptr) {
        if (h.size() > n) alloc(h.size());
        if (!h.empty()) {
            CPG_HIP(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
            CPG_HIP(hipStreamSynchronize(s));
        }
    }
};

inline hipStream_t S(void *stream) { return reinterpret_cast<hipStream_t>(stream); }

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Process-wide counter.  A handle takes a new value when it is created and whenever device
// buffers that captured graphs point into are reallocated, so a graph cache keyed on handle
// addresses tells a handle re-created at a reused address (or a reallocated one) from the one
// whose buffers it captured.
inline uint64_t next_generation() {
    static std::atomic<uint64_t> g{0};
    return ++g;
}

}  // namespace cpg

// ---- handle types ------------------------------------------------------------------
struct cpg_bsr_s {
    uint64_t gen = cpg::next_generation();
    int n = 0, bl = 0, nnzb = 0;
    int nbmax = 1;   // longest block row
    std::vector<int> h_rowptr, h_colind;
    cpg::DevBuf<int> rowptr, colind;
    cpg::DevBuf<double> blocks;
};

namespace cpg {
// shared kernels-launchers used across translation units
// `active` (optional device flag): the launches do nothing when it reads 0
void bsr_matvec(const cpg_bsr_s &A, const double *x, double *y, hipStream_t s, const int *active = nullptr);
void vec_fill(double *x, double v, size_t n, hipStream_t s);
// waves per block row of the row-stream kernels (rowstream.hpp), per element size and block size
int stream_wpr(int elem_bytes, int bl);
void set_stream_wpr(int elem_bytes, int bl, int wpr);
// out = M^-1 in for a factored preconditioner (in and out distinct); `active` as for bsr_matvec
// (DILU and block-Jacobi sweeps only)
void prec_apply_to(cpg_prec P, const double *in, double *out, hipStream_t s, const int *active = nullptr);
// reports (CPG_ERR_HIP) a single-launch block DILU apply whose dependency wait ran out since the last check
void prec_flow_check(cpg_prec P);
// the preconditioner's generation (next_generation): changes when its buffers are reallocated
uint64_t prec_generation(cpg_prec P);
// The factor streams of a DILU or block-Jacobi preconditioner as its sweeps read them: the
// forward sweep takes the rows frows[flev[l] .. flev[l+1]) of level l, row frows[p] from the
// packed blocks [fptr[p], fptr[p+1]) of F; the backward sweep likewise with blev, brows, bptr;
// scol[b] is the source tag of packed block b (stream_source in rowstream.hpp).  Block Jacobi:
// block i of F is D_i^-1 and the lists are unused.  The host vectors live as long as P.
struct PrecTables {
    int kind = 0, precision = 0, n = 0, bl = 0;
    bool diag32 = false;
    const void *F = nullptr;
    const float *scale = nullptr;
    const int *scol = nullptr;
    const std::vector<int> *flev = nullptr, *frows = nullptr, *fptr = nullptr;
    const std::vector<int> *blev = nullptr, *brows = nullptr, *bptr = nullptr;
};
PrecTables prec_tables(cpg_prec P);

// y = A x for the Krylov solvers; `active` as for bsr_matvec
struct LinOp {
    void (*fn)(void *ctx, const double *x, double *y, const int *active, hipStream_t s);
    void *ctx;
};
// the BSR matvec as a LinOp (ctx = const cpg_bsr_s *)
void bsr_linop(void *ctx, const double *x, double *y, const int *active, hipStream_t s);
// GMRES with the driver of K's mode (krylov.hip); x holds the initial guess unless x_zero
void gmres_solve(cpg_krylov K, const LinOp &A, cpg_prec P, const double *b, double *x, bool x_zero, double *tol,
                 int *maxit, hipStream_t s);
// out = M^-1 in (out distinct from in) for the flexible driver; `active` as for bsr_matvec
struct PcOp {
    void (*fn)(void *ctx, const double *in, double *out, const int *active, hipStream_t s);
    void *ctx;
};
// Block DILU of the stage-coupled Radau IIA system in its cell-interleaved form (3 stages of
// 40 x 40 blocks: bl 120).  The coupled block of slot q is coef[s][t] J_s[q] (+ the mass block
// on the diagonal), built on the fly from the stage Jacobians; P is a DILU handle created over
// the 120-wide pattern of the stage Jacobians (precond.hip)
struct CoupledBlocks {
    const double *J[3];   // stage Jacobian blocks (bl 40, one BSR pattern)
    const double *mass;   // per-cell bs x bs mass blocks
    double coef[9];       // coef[3 s + t] = -dt A[s][t]
    int bs;
    int n, nnzb;          // block rows and slots of the stage pattern
};
void prec_factor_coupled(cpg_prec P, const CoupledBlocks &C, hipStream_t s);
// flexible GMRES (K in CPG_GMRES_FLEXIBLE mode) with the preconditioner as a PcOp
void gmres_flexible_op(cpg_krylov K, const LinOp &A, const PcOp &M, const double *b, double *x, bool x_zero,
                       double *tol, int *maxit, hipStream_t s);

// Kernel timing from a dispatch's own start and end timestamps (hipExtLaunchKernel events, no
// marker packets between kernels): the caller arms a pair of events and the next launch that goes
// through launch_timed takes them (and sets used); other launches leave them armed.
struct DispatchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
    bool used = false;
};
DispatchTiming &dispatch_timing();   // per host thread

template <class F, class... Args>
void launch_timed(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, Args... args) {
    DispatchTiming &t = dispatch_timing();
    if (t.start) {
        hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, s, t.start, t.stop, 0u, args...);
        t.used = true;
        t.start = t.stop = nullptr;
    } else {
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
    }
}

}  // namespace cpg
