// hbs_api.cpp -- C-ABI of the engine: contexts, device CSR, SpMV/SpMM,
// BiCGStab (FP64, mixed FP32/FP64 refinement, adjoint) and one rank's share.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "hbs_err.h"
#include "hbs_internal.h"

using namespace hbs;

namespace {

thread_local std::string g_err;
int g_block_nnz = 2048;
int g_serial_max_avg = 64;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace

int hbs::set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(HBS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                \
    } while (0)

#define CHK(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != HBS_OK) return rc_; \
    } while (0)

template <typename T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(HBS_ERR_ALLOC, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    return HBS_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

int vec_elems_per_block(int n)
{
    // 256 threads x 16-byte packs, enough blocks to fill 256 CUs, at most ~2k
    // partials for the final reduction
    int u = (n + 512 * 2048 - 1) / (512 * 2048);
    u = std::max(1, std::min(16, u));
    return 512 * u;
}

int alloc_scratch(Scratch &sc, int nslots, int nrow, int nsys)
{
    CHK(dalloc(&sc.part, (size_t)std::max(nslots, 1) * kMaxDots));
    CHK(dalloc(&sc.chunk, (size_t)std::max(nrow, 1)));
    CHK(dalloc(&sc.ctr_row, (size_t)std::max(nrow, 1)));
    CHK(dalloc(&sc.ctr_sys, (size_t)std::max(nsys, 1)));
    CHK(dalloc(&sc.ndone, 1));
    HIPCHK(hipMemset(sc.ctr_row, 0, sizeof(unsigned) * std::max(nrow, 1)));
    HIPCHK(hipMemset(sc.ctr_sys, 0, sizeof(unsigned) * std::max(nsys, 1)));
    HIPCHK(hipMemset(sc.ndone, 0, sizeof(unsigned)));
    HIPCHK(hipMemset(sc.part, 0, sizeof(double) * std::max(nslots, 1) * kMaxDots));
    HIPCHK(hipMemset(sc.chunk, 0, sizeof(double) * std::max(nrow, 1)));
    return HBS_OK;
}

void free_scratch(Scratch &sc)
{
    dfree(sc.part);
    dfree(sc.chunk);
    dfree(sc.ctr_row);
    dfree(sc.ctr_sys);
    dfree(sc.ndone);
}

struct DevVecPlan {
    int n = -1;
    int nvb = 0;
    VecBlock *blocks = nullptr;
    SysRange *sys = nullptr;
};

int build_vec_plan(DevVecPlan &p, int n)
{
    dfree(p.blocks);
    dfree(p.sys);
    std::vector<VecBlock> vb;
    plan_vec(0, n, vec_elems_per_block(n), 0, vb);
    p.nvb = (int)vb.size();
    p.n = n;
    CHK(dalloc(&p.blocks, vb.size()));
    if (!vb.empty()) HIPCHK(hipMemcpy(p.blocks, vb.data(), vb.size() * sizeof(VecBlock), hipMemcpyHostToDevice));
    SysRange sr{0, 0, 0, p.nvb};
    CHK(dalloc(&p.sys, 1));
    HIPCHK(hipMemcpy(p.sys, &sr, sizeof sr, hipMemcpyHostToDevice));
    return HBS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// context

struct hbs_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    SolverState *st = nullptr;       // device, BLAS-1 results
    SolverState *hst = nullptr;      // pinned host
    Scratch sc;
    DevVecPlan vp;                   // cached plan for BLAS-1 dots
};

namespace {
struct DevGuard {
    explicit DevGuard(int d) { (void)hipSetDevice(d); }
};
}  // namespace

extern "C" const char *hbs_last_error(void) { return g_err.c_str(); }
extern "C" const char *hbs_version(void) { return "hbs 0.1 (gfx950)"; }

extern "C" int hbs_device_count(int *out)
{
    if (!out) return fail(HBS_ERR_ARG, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *out = n;
    return HBS_OK;
}

extern "C" int hbs_ctx_destroy(hbs_ctx ctx)
{
    if (!ctx) return HBS_OK;
    DevGuard g(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx->st);
    if (ctx->hst) (void)hipHostFree(ctx->hst);
    free_scratch(ctx->sc);
    dfree(ctx->vp.blocks);
    dfree(ctx->vp.sys);
    if (ctx->t0) (void)hipEventDestroy(ctx->t0);
    if (ctx->t1) (void)hipEventDestroy(ctx->t1);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HBS_OK;
}

extern "C" int hbs_ctx_create(int device, void *stream, hbs_ctx *out)
{
    if (!out) return fail(HBS_ERR_ARG, "null out");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(HBS_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(HBS_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    hbs_ctx c = new hbs_ctx_s();
    c->device = device;
    HIPCHK(hipSetDevice(device));
    int rc = HBS_OK;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = HBS_ERR_HIP;
        c->own_stream = true;
    }
    if (rc == HBS_OK && (hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess))
        rc = HBS_ERR_HIP;
    if (rc == HBS_OK) rc = dalloc(&c->st, 1);
    if (rc == HBS_OK && hipHostMalloc((void **)&c->hst, sizeof(SolverState)) != hipSuccess) rc = HBS_ERR_ALLOC;
    if (rc == HBS_OK) rc = alloc_scratch(c->sc, 4096, 1, 1);
    if (rc != HBS_OK) {
        std::string e = g_err;
        hbs_ctx_destroy(c);
        g_err = e.empty() ? "context creation failed" : e;
        return rc == HBS_OK ? HBS_ERR_HIP : rc;
    }
    *out = c;
    return HBS_OK;
}

extern "C" int hbs_ctx_sync(hbs_ctx ctx)
{
    if (!ctx) return fail(HBS_ERR_ARG, "null ctx");
    DevGuard g(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return HBS_OK;
}

extern "C" void *hbs_ctx_stream(hbs_ctx ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int hbs_dev_alloc(hbs_ctx ctx, size_t bytes, void **dev)
{
    if (!ctx || !dev) return fail(HBS_ERR_ARG, "null argument");
    DevGuard g(ctx->device);
    unsigned char *p = nullptr;
    CHK(dalloc(&p, bytes));
    *dev = p;
    return HBS_OK;
}

extern "C" int hbs_dev_free(hbs_ctx ctx, void *dev)
{
    if (!ctx) return fail(HBS_ERR_ARG, "null ctx");
    if (!dev) return HBS_OK;
    DevGuard g(ctx->device);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(dev));
    return HBS_OK;
}

extern "C" int hbs_h2d(hbs_ctx ctx, void *dev, const void *host, size_t bytes)
{
    if (!ctx || (bytes && (!dev || !host))) return fail(HBS_ERR_ARG, "null argument");
    if (!bytes) return HBS_OK;
    DevGuard g(ctx->device);
    HIPCHK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_d2h(hbs_ctx ctx, void *host, const void *dev, size_t bytes)
{
    if (!ctx || (bytes && (!dev || !host))) return fail(HBS_ERR_ARG, "null argument");
    if (!bytes) return HBS_OK;
    DevGuard g(ctx->device);
    HIPCHK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_d2d(hbs_ctx ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return fail(HBS_ERR_ARG, "null argument");
    if (!bytes) return HBS_OK;
    DevGuard g(ctx->device);
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_dev_zero(hbs_ctx ctx, void *dev, size_t bytes)
{
    if (!ctx || (bytes && !dev)) return fail(HBS_ERR_ARG, "null argument");
    DevGuard g(ctx->device);
    HIPCHK(launch_fill_zero(dev, bytes, ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_timer_start(hbs_ctx ctx)
{
    if (!ctx) return fail(HBS_ERR_ARG, "null ctx");
    DevGuard g(ctx->device);
    HIPCHK(hipEventRecord(ctx->t0, ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_timer_stop(hbs_ctx ctx, double *ms)
{
    if (!ctx || !ms) return fail(HBS_ERR_ARG, "null argument");
    DevGuard g(ctx->device);
    HIPCHK(hipEventRecord(ctx->t1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->t1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
    *ms = f;
    return HBS_OK;
}

// ---------------------------------------------------------------------------
// BLAS-1

extern "C" int hbs_dot_f64(hbs_ctx ctx, int n, const double *x, const double *y, double *out_host)
{
    if (!ctx || !out_host || n < 0 || (n > 0 && (!x || !y))) return fail(HBS_ERR_ARG, "bad argument");
    if (n == 0) { *out_host = 0.0; return HBS_OK; }
    DevGuard g(ctx->device);
    if (ctx->vp.n != n) CHK(build_vec_plan(ctx->vp, n));
    if (ctx->vp.nvb > 4096) return fail(HBS_ERR_ARG, "vector too long for the context scratch");
    VecArgs<double> a{};
    a.blocks = ctx->vp.blocks;
    a.sys = ctx->vp.sys;
    a.nvb = ctx->vp.nvb;
    a.st = ctx->st;
    a.sc = ctx->sc;
    HIPCHK(launch_dot(a, x, y, FIN_STORE, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->hst, ctx->st, sizeof(SolverState), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out_host = ctx->hst->out;
    return HBS_OK;
}

extern "C" int hbs_axpy_f64(hbs_ctx ctx, int n, double a, const double *x, double *y)
{
    if (!ctx || n < 0 || (n > 0 && (!x || !y))) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(ctx->device);
    HIPCHK(launch_axpy_f64(n, a, x, y, ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_scal_f64(hbs_ctx ctx, int n, double a, double *x)
{
    if (!ctx || n < 0 || (n > 0 && !x)) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(ctx->device);
    HIPCHK(launch_scal_f64(n, a, x, ctx->stream));
    return HBS_OK;
}

// ---------------------------------------------------------------------------
// device CSR

struct hbs_csr_s {
    hbs_ctx ctx = nullptr;
    int m = 0, n = 0, nnz = 0, flags = 0;
    int *rowptr = nullptr, *col = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;
    SpmvPlan plan;
    SysRange *sys = nullptr;
    Scratch sc;
    hbs_csr_s *tr = nullptr;
    std::map<std::pair<int, int>, int> spmm_tile;   // (k, dtype) -> tile
    hipEvent_t *evs = nullptr;
    int nevs = 0;
};

extern "C" int hbs_set_spmv_plan(int block_nnz, int serial_max_avg)
{
    if (block_nnz < 256 || block_nnz > kMaxBlockNnz || (block_nnz % 256) != 0)
        return fail(HBS_ERR_ARG, "block_nnz must be a multiple of 256 in [256, %d]", kMaxBlockNnz);
    if (serial_max_avg < 0) return fail(HBS_ERR_ARG, "serial_max_avg must be >= 0");
    g_block_nnz = block_nnz;
    g_serial_max_avg = serial_max_avg;
    return HBS_OK;
}

extern "C" int hbs_csr_destroy(hbs_csr a)
{
    if (!a) return HBS_OK;
    DevGuard g(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    dfree(a->rowptr);
    dfree(a->col);
    dfree(a->val);
    dfree(a->val32);
    dfree(a->plan.blocks);
    dfree(a->plan.aux);
    dfree(a->sys);
    free_scratch(a->sc);
    if (a->evs) {
        for (int i = 0; i < a->nevs; ++i) (void)hipEventDestroy(a->evs[i]);
        delete[] a->evs;
    }
    if (a->tr) hbs_csr_destroy(a->tr);
    delete a;
    return HBS_OK;
}

namespace {

int validate_csr(int m, int n, int nnz, const int *rowptr, const int *col, const double *val)
{
    if (m < 0 || n < 0 || nnz < 0) return fail(HBS_ERR_ARG, "negative dimension");
    if (!rowptr || (nnz > 0 && (!col || !val))) return fail(HBS_ERR_ARG, "null CSR array");
    if (rowptr[0] != 0 || rowptr[m] != nnz) return fail(HBS_ERR_ARG, "rowptr[0] must be 0 and rowptr[m] == nnz");
    for (int i = 0; i < m; ++i)
        if (rowptr[i + 1] < rowptr[i]) return fail(HBS_ERR_ARG, "rowptr not monotone at row %d", i);
    for (int j = 0; j < nnz; ++j)
        if (col[j] < 0 || col[j] >= n) return fail(HBS_ERR_ARG, "column index %d out of range at entry %d", col[j], j);
    return HBS_OK;
}

int csr_upload(hbs_ctx ctx, int m, int n, int nnz, const int *rowptr, const int *col, const double *val,
               int flags, hbs_csr *out)
{
    hbs_csr a = new hbs_csr_s();
    a->ctx = ctx;
    a->m = m;
    a->n = n;
    a->nnz = nnz;
    a->flags = flags;
    auto bail = [&](int rc) {
        std::string e = g_err;
        hbs_csr_destroy(a);
        g_err = e;
        return rc;
    };
    int rc;
    if ((rc = dalloc(&a->rowptr, (size_t)m + 1)) || (rc = dalloc(&a->col, nnz)) || (rc = dalloc(&a->val, nnz)))
        return bail(rc);
    if (hipMemcpy(a->rowptr, rowptr, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (nnz && hipMemcpy(a->col, col, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) ||
        (nnz && hipMemcpy(a->val, val, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess))
        return bail(fail(HBS_ERR_HIP, "CSR upload failed"));
    if (flags & HBS_CSR_F32) {
        if ((rc = dalloc(&a->val32, nnz))) return bail(rc);
        if (launch_to_f32(nnz, a->val, a->val32, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return bail(fail(HBS_ERR_HIP, "FP32 value copy failed"));
    }
    std::vector<RowBlock> blocks;
    std::vector<BlockAux> aux;
    a->plan.block_nnz = g_block_nnz;
    plan_rows(0, m, rowptr, g_block_nnz, g_serial_max_avg, 0, blocks, aux, &a->plan.serial_rows,
              &a->plan.max_row);
    a->plan.nblk = (int)blocks.size();
    if ((rc = dalloc(&a->plan.blocks, blocks.size())) || (rc = dalloc(&a->plan.aux, aux.size())) ||
        (rc = dalloc(&a->sys, 1)))
        return bail(rc);
    if (!blocks.empty() &&
        (hipMemcpy(a->plan.blocks, blocks.data(), blocks.size() * sizeof(RowBlock), hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(a->plan.aux, aux.data(), aux.size() * sizeof(BlockAux), hipMemcpyHostToDevice) != hipSuccess))
        return bail(fail(HBS_ERR_HIP, "plan upload failed"));
    SysRange sr{0, a->plan.nblk, 0, 0};
    if (hipMemcpy(a->sys, &sr, sizeof sr, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(HBS_ERR_HIP, "plan upload failed"));
    if ((rc = alloc_scratch(a->sc, a->plan.nblk, a->plan.nblk, 1))) return bail(rc);
    *out = a;
    return HBS_OK;
}

template <typename T>
SpmvArgs<T> spmv_args(hbs_csr a, const T *val, const Scratch &sc)
{
    SpmvArgs<T> s{};
    s.blocks = a->plan.blocks;
    s.aux = a->plan.aux;
    s.sys = a->sys;
    s.nblk = a->plan.nblk;
    s.rowptr = a->rowptr;
    s.col = a->col;
    s.val = val;
    s.sc = sc;
    s.alpha = T(1);
    s.beta = T(0);
    s.fin = FIN_NONE;
    return s;
}

}  // namespace

extern "C" int hbs_csr_create(hbs_ctx ctx, int m, int n, int nnz, const int *rowptr, const int *col,
                              const double *val, int flags, hbs_csr *out)
{
    if (!ctx || !out) return fail(HBS_ERR_ARG, "null argument");
    *out = nullptr;
    CHK(validate_csr(m, n, nnz, rowptr, col, val));
    DevGuard g(ctx->device);
    hbs_csr a = nullptr;
    CHK(csr_upload(ctx, m, n, nnz, rowptr, col, val, flags, &a));
    if (flags & HBS_CSR_TRANSPOSE) {
        // A^T as CSR: counting sort by column, rows visited in ascending order
        std::vector<int> tptr((size_t)n + 1, 0), tcol((size_t)nnz);
        std::vector<double> tval((size_t)nnz);
        for (int j = 0; j < nnz; ++j) tptr[(size_t)col[j] + 1]++;
        for (int c = 0; c < n; ++c) tptr[(size_t)c + 1] += tptr[c];
        std::vector<int> fill(tptr.begin(), tptr.end() - 1);
        for (int r = 0; r < m; ++r)
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const int d = fill[col[j]]++;
                tcol[d] = r;
                tval[d] = val[j];
            }
        int rc = csr_upload(ctx, n, m, nnz, tptr.data(), tcol.data(), tval.data(), 0, &a->tr);
        if (rc != HBS_OK) {
            std::string e = g_err;
            hbs_csr_destroy(a);
            g_err = e;
            return rc;
        }
    }
    *out = a;
    return HBS_OK;
}

extern "C" int hbs_csr_info(hbs_csr a, int *m, int *n, int *nnz, int *nblocks, int *serial_rows)
{
    if (!a) return fail(HBS_ERR_ARG, "null csr");
    if (m) *m = a->m;
    if (n) *n = a->n;
    if (nnz) *nnz = a->nnz;
    if (nblocks) *nblocks = a->plan.nblk;
    if (serial_rows) *serial_rows = a->plan.serial_rows;
    return HBS_OK;
}

extern "C" int hbs_csr_device_arrays(hbs_csr a, const int **rowptr, const int **col, const double **val)
{
    if (!a) return fail(HBS_ERR_ARG, "null csr");
    if (rowptr) *rowptr = a->rowptr;
    if (col) *col = a->col;
    if (val) *val = a->val;
    return HBS_OK;
}

extern "C" int hbs_csr_spmv_f64(hbs_csr a, double alpha, const double *x, double beta, double *y)
{
    if (!a || (a->m > 0 && (!x || !y))) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(a->ctx->device);
    SpmvArgs<double> s = spmv_args<double>(a, a->val, a->sc);
    s.x = x;
    s.y = y;
    s.alpha = alpha;
    s.beta = beta;
    HIPCHK(launch_spmv(s, EPI_AXPBY, a->plan.block_nnz, a->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_csr_spmv_f32(hbs_csr a, float alpha, const float *x, float beta, float *y)
{
    if (!a || (a->m > 0 && (!x || !y))) return fail(HBS_ERR_ARG, "bad argument");
    if (!a->val32) return fail(HBS_ERR_ARG, "matrix was created without HBS_CSR_F32");
    DevGuard g(a->ctx->device);
    SpmvArgs<float> s = spmv_args<float>(a, a->val32, a->sc);
    s.x = x;
    s.y = y;
    s.alpha = alpha;
    s.beta = beta;
    HIPCHK(launch_spmv(s, EPI_AXPBY, a->plan.block_nnz, a->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_csr_spmv_t_f64(hbs_csr a, const double *x, double *y)
{
    if (!a) return fail(HBS_ERR_ARG, "null csr");
    if (!a->tr) return fail(HBS_ERR_ARG, "matrix was created without HBS_CSR_TRANSPOSE");
    return hbs_csr_spmv_f64(a->tr, 1.0, x, 0.0, y);
}

namespace {

int ensure_events(hbs_csr a, int n)
{
    if (a->nevs >= n) return HBS_OK;
    if (a->evs) {
        for (int i = 0; i < a->nevs; ++i) (void)hipEventDestroy(a->evs[i]);
        delete[] a->evs;
    }
    a->evs = new hipEvent_t[n];
    a->nevs = n;
    for (int i = 0; i < n; ++i) HIPCHK(hipEventCreate(&a->evs[i]));
    return HBS_OK;
}

template <typename T>
int spmm_launch(hbs_csr a, int k, const T *X, T *Y, int tile)
{
    if constexpr (sizeof(T) == 8)
        return launch_spmm_f64(a->m, a->rowptr, a->col, a->val, k, X, Y, tile, a->ctx->stream) == hipSuccess
                   ? HBS_OK
                   : fail(HBS_ERR_HIP, "spmm launch failed");
    else
        return launch_spmm_f32(a->m, a->rowptr, a->col, a->val32, k, X, Y, tile, a->ctx->stream) == hipSuccess
                   ? HBS_OK
                   : fail(HBS_ERR_HIP, "spmm launch failed");
}

// Own autotuner: time each admissible RHS tile width (a few launches each,
// hipEvents) and keep the fastest per (k, dtype).
template <typename T>
int spmm_tuned(hbs_csr a, int k, const T *X, T *Y, int tile)
{
    if (tile != 0) {
        if (tile < 1 || tile > 64 || (tile & (tile - 1))) return fail(HBS_ERR_ARG, "tile must be a power of two <= 64");
        return spmm_launch<T>(a, k, X, Y, tile);
    }
    const auto key = std::make_pair(k, (int)sizeof(T));
    auto it = a->spmm_tile.find(key);
    if (it != a->spmm_tile.end()) return spmm_launch<T>(a, k, X, Y, it->second);
    CHK(ensure_events(a, 2));
    int best = 0;
    float best_ms = 1e30f;
    for (int t = 4; t <= 64; t *= 2) {
        if (t > 2 * k && t > 4) break;
        CHK(spmm_launch<T>(a, k, X, Y, t));   // warm
        HIPCHK(hipEventRecord(a->evs[0], a->ctx->stream));
        for (int r = 0; r < 3; ++r) CHK(spmm_launch<T>(a, k, X, Y, t));
        HIPCHK(hipEventRecord(a->evs[1], a->ctx->stream));
        HIPCHK(hipEventSynchronize(a->evs[1]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, a->evs[0], a->evs[1]));
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    a->spmm_tile[key] = best;
    return spmm_launch<T>(a, k, X, Y, best);
}

}  // namespace

extern "C" int hbs_csr_spmm_f64(hbs_csr a, int k, const double *X, double *Y, int tile)
{
    if (!a || k < 0 || (a->m > 0 && k > 0 && (!X || !Y))) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(a->ctx->device);
    return spmm_tuned<double>(a, k, X, Y, tile);
}

extern "C" int hbs_csr_spmm_f32(hbs_csr a, int k, const float *X, float *Y, int tile)
{
    if (!a || k < 0 || (a->m > 0 && k > 0 && (!X || !Y))) return fail(HBS_ERR_ARG, "bad argument");
    if (!a->val32) return fail(HBS_ERR_ARG, "matrix was created without HBS_CSR_F32");
    DevGuard g(a->ctx->device);
    return spmm_tuned<float>(a, k, X, Y, tile);
}

extern "C" int hbs_bench_spmv_f64(hbs_csr a, const double *x, double *y, int reps, double *avg_ms)
{
    if (!a || !avg_ms || reps <= 0) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(a->ctx->device);
    CHK(ensure_events(a, 2 * reps));
    CHK(hbs_csr_spmv_f64(a, 1.0, x, 0.0, y));   // warm-up, as spmv_gpu.c:116-118
    for (int r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(a->evs[2 * r], a->ctx->stream));
        CHK(hbs_csr_spmv_f64(a, 1.0, x, 0.0, y));
        HIPCHK(hipEventRecord(a->evs[2 * r + 1], a->ctx->stream));
    }
    HIPCHK(hipEventSynchronize(a->evs[2 * reps - 1]));
    double tot = 0.0;
    for (int r = 0; r < reps; ++r) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, a->evs[2 * r], a->evs[2 * r + 1]));
        tot += ms;
    }
    *avg_ms = tot / reps;
    return HBS_OK;
}

extern "C" int hbs_bench_spmv_f32(hbs_csr a, const float *x, float *y, int reps, double *avg_ms)
{
    if (!a || !avg_ms || reps <= 0) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(a->ctx->device);
    CHK(ensure_events(a, 2 * reps));
    CHK(hbs_csr_spmv_f32(a, 1.f, x, 0.f, y));
    for (int r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(a->evs[2 * r], a->ctx->stream));
        CHK(hbs_csr_spmv_f32(a, 1.f, x, 0.f, y));
        HIPCHK(hipEventRecord(a->evs[2 * r + 1], a->ctx->stream));
    }
    HIPCHK(hipEventSynchronize(a->evs[2 * reps - 1]));
    double tot = 0.0;
    for (int r = 0; r < reps; ++r) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, a->evs[2 * r], a->evs[2 * r + 1]));
        tot += ms;
    }
    *avg_ms = tot / reps;
    return HBS_OK;
}

extern "C" int hbs_bench_copy(hbs_ctx ctx, const double *src, double *dst, size_t n, int reps, double *avg_ms)
{
    if (!ctx || !avg_ms || reps <= 0 || !src || !dst) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(ctx->device);
    HIPCHK(launch_copy_f64(n, src, dst, ctx->stream));
    HIPCHK(hipEventRecord(ctx->t0, ctx->stream));
    for (int r = 0; r < reps; ++r) HIPCHK(launch_copy_f64(n, src, dst, ctx->stream));
    HIPCHK(hipEventRecord(ctx->t1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->t1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
    *avg_ms = ms / reps;
    return HBS_OK;
}

// ---------------------------------------------------------------------------
// BiCGStab

extern "C" void hbs_solve_opts_default(hbs_solve_opts *o)
{
    if (!o) return;
    o->tol = 1e-8;
    o->maxit = 1000;
    o->fixed_iters = 0;
    o->check_every = 8;
    o->use_graph = 1;
    o->inner_maxit = 1000;
    o->inner_tol = 1e-4;
    o->max_sweeps = 8;
}

struct hbs_solver_s {
    hbs_csr A = nullptr;
    int n = 0;
    double *X = nullptr, *R = nullptr, *R0 = nullptr, *P = nullptr, *V = nullptr, *S = nullptr, *T = nullptr;
    float *X32 = nullptr, *R32 = nullptr, *R032 = nullptr, *P32 = nullptr, *V32 = nullptr, *S32 = nullptr,
          *T32 = nullptr;
    SolverState *st = nullptr;    // device [2]: [0] main / inner, [1] outer
    SolverState *hst = nullptr;   // pinned host [3]
    Scratch sc;
    VecBlock *vblocks = nullptr;
    SysRange *sys = nullptr;
    int nvb = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // [0] FP64, [1] FP32
    int graph_chunk[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hbs_solver_s *adj = nullptr;
};

extern "C" int hbs_solver_destroy(hbs_solver s)
{
    if (!s) return HBS_OK;
    DevGuard g(s->A->ctx->device);
    (void)hipStreamSynchronize(s->A->ctx->stream);
    for (double **p : {&s->X, &s->R, &s->R0, &s->P, &s->V, &s->S, &s->T}) dfree(*p);
    for (float **p : {&s->X32, &s->R32, &s->R032, &s->P32, &s->V32, &s->S32, &s->T32}) dfree(*p);
    dfree(s->st);
    if (s->hst) (void)hipHostFree(s->hst);
    free_scratch(s->sc);
    dfree(s->vblocks);
    dfree(s->sys);
    for (int i = 0; i < 2; ++i)
        if (s->graph[i]) (void)hipGraphExecDestroy(s->graph[i]);
    for (hipEvent_t e : {s->ev[0], s->ev[1], s->e0, s->e1})
        if (e) (void)hipEventDestroy(e);
    if (s->adj) hbs_solver_destroy(s->adj);
    delete s;
    return HBS_OK;
}

extern "C" int hbs_solver_create(hbs_csr a, hbs_solver *out)
{
    if (!a || !out) return fail(HBS_ERR_ARG, "null argument");
    if (a->m != a->n) return fail(HBS_ERR_ARG, "BiCGStab needs a square matrix (%d x %d)", a->m, a->n);
    *out = nullptr;
    DevGuard g(a->ctx->device);
    hbs_solver s = new hbs_solver_s();
    s->A = a;
    s->n = a->m;
    auto bail = [&](int rc) {
        std::string e = g_err;
        hbs_solver_destroy(s);
        g_err = e;
        return rc;
    };
    int rc;
    for (double **p : {&s->X, &s->R, &s->R0, &s->P, &s->V, &s->S, &s->T})
        if ((rc = dalloc(p, s->n))) return bail(rc);
    if ((rc = dalloc(&s->st, 2))) return bail(rc);
    if (hipHostMalloc((void **)&s->hst, 3 * sizeof(SolverState)) != hipSuccess)
        return bail(fail(HBS_ERR_ALLOC, "pinned host allocation failed"));
    std::vector<VecBlock> vb;
    plan_vec(0, s->n, vec_elems_per_block(s->n), 0, vb);
    s->nvb = (int)vb.size();
    if ((rc = dalloc(&s->vblocks, vb.size())) || (rc = dalloc(&s->sys, 1))) return bail(rc);
    if (!vb.empty() &&
        hipMemcpy(s->vblocks, vb.data(), vb.size() * sizeof(VecBlock), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(HBS_ERR_HIP, "plan upload failed"));
    SysRange sr{0, a->plan.nblk, 0, s->nvb};
    if (hipMemcpy(s->sys, &sr, sizeof sr, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(HBS_ERR_HIP, "plan upload failed"));
    if ((rc = alloc_scratch(s->sc, std::max(a->plan.nblk, s->nvb), a->plan.nblk, 1))) return bail(rc);
    for (hipEvent_t *e : {&s->ev[0], &s->ev[1], &s->e0, &s->e1})
        if (hipEventCreate(e) != hipSuccess) return bail(fail(HBS_ERR_HIP, "event creation failed"));
    *out = s;
    return HBS_OK;
}

namespace {

template <typename T>
struct Work {
    T *X, *R, *R0, *P, *V, *S, *T_;
    const T *val;
};

Work<double> work64(hbs_solver s) { return {s->X, s->R, s->R0, s->P, s->V, s->S, s->T, s->A->val}; }
Work<float> work32(hbs_solver s) { return {s->X32, s->R32, s->R032, s->P32, s->V32, s->S32, s->T32, s->A->val32}; }

template <typename T>
VecArgs<T> vec_args(hbs_solver s, const Work<T> &w, SolverState *st)
{
    VecArgs<T> a{};
    a.blocks = s->vblocks;
    a.sys = s->sys;
    a.nvb = s->nvb;
    a.X = w.X;
    a.R = w.R;
    a.P = w.P;
    a.V = w.V;
    a.S = w.S;
    a.T_ = w.T_;
    a.R0 = w.R0;
    a.st = st;
    a.sc = s->sc;
    return a;
}

// One BiCGStab iteration: 5 kernels, every dot fused into the pass before it.
template <typename T>
hipError_t launch_iteration(hbs_solver s, const Work<T> &w, SolverState *st)
{
    hipStream_t q = s->A->ctx->stream;
    const int bnnz = s->A->plan.block_nnz;
    VecArgs<T> va = vec_args<T>(s, w, st);
    SpmvArgs<T> sa = spmv_args<T>(s->A, w.val, s->sc);
    sa.st = st;
    sa.check_done = 1;
    hipError_t e;
    if ((e = launch_vec_P(va, q)) != hipSuccess) return e;        // P = ((P - V w) b) + R
    sa.x = w.P;
    sa.y = w.V;
    sa.w = w.R0;
    sa.fin = FIN_ALPHA;
    if ((e = launch_spmv(sa, EPI_DOT1, bnnz, q)) != hipSuccess) return e;   // V = A P ; alpha
    if ((e = launch_vec_S(va, q)) != hipSuccess) return e;        // S = R - V a ; ||S||
    sa.x = w.S;
    sa.y = w.T_;
    sa.w = w.S;
    sa.fin = FIN_OMEGA;
    sa.skip_on_sconv = 1;
    if ((e = launch_spmv(sa, EPI_DOT2, bnnz, q)) != hipSuccess) return e;   // T = A S ; omega
    return launch_vec_XR(va, q);                                  // X, R ; rho', ||R||
}

template <typename T>
int run_chunk(hbs_solver s, const Work<T> &w, SolverState *st, int k, bool use_graph)
{
    hipStream_t q = s->A->ctx->stream;
    const int gi = sizeof(T) == 8 ? 0 : 1;
    if (!use_graph) {
        for (int i = 0; i < k; ++i) HIPCHK(launch_iteration<T>(s, w, st));
        return HBS_OK;
    }
    if (!s->graph[gi] || s->graph_chunk[gi] != k) {
        if (s->graph[gi]) {
            HIPCHK(hipGraphExecDestroy(s->graph[gi]));
            s->graph[gi] = nullptr;
        }
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        for (int i = 0; i < k && le == hipSuccess; ++i) le = launch_iteration<T>(s, w, st);
        hipError_t ce = hipStreamEndCapture(q, &gr);
        if (le != hipSuccess) return fail(HBS_ERR_HIP, "launch during capture failed: %s", hipGetErrorString(le));
        if (ce != hipSuccess) return fail(HBS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&s->graph[gi], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ie != hipSuccess) {
            s->graph[gi] = nullptr;
            return fail(HBS_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
        }
        s->graph_chunk[gi] = k;
    }
    HIPCHK(hipGraphLaunch(s->graph[gi], q));
    return HBS_OK;
}

// Launch iteration chunks until the device says the system is done.  The
// state of chunk c is read back after chunk c + 1 is queued, so the GPU never
// idles on the poll; iterations queued after convergence return at once.
template <typename T>
int iterate(hbs_solver s, const Work<T> &w, SolverState *st, int maxit, int check_every, bool use_graph)
{
    hipStream_t q = s->A->ctx->stream;
    const int k = check_every > 0 ? check_every : 8;
    long issued = 0;
    int c = 0, pending = -1;
    for (;;) {
        CHK(run_chunk<T>(s, w, st, k, use_graph));
        issued += k;
        HIPCHK(hipMemcpyAsync(&s->hst[c & 1], st, sizeof(SolverState), hipMemcpyDeviceToHost, q));
        HIPCHK(hipEventRecord(s->ev[c & 1], q));
        if (pending >= 0) {
            HIPCHK(hipEventSynchronize(s->ev[pending & 1]));
            if (s->hst[pending & 1].done) break;
        }
        pending = c++;
        if (issued >= maxit) {
            HIPCHK(hipEventSynchronize(s->ev[pending & 1]));
            break;
        }
    }
    return HBS_OK;
}

int read_state(hbs_solver s, int slot, SolverState *out)
{
    hipStream_t q = s->A->ctx->stream;
    HIPCHK(hipMemcpyAsync(&s->hst[2], s->st + slot, sizeof(SolverState), hipMemcpyDeviceToHost, q));
    HIPCHK(hipStreamSynchronize(q));
    *out = s->hst[2];
    return HBS_OK;
}

int status_of(const SolverState &st, bool fixed)
{
    if (st.stop == 4) return HBS_ERR_BREAKDOWN;
    if (!fixed && st.stop == 0) return HBS_ERR_NOCONV;
    return HBS_OK;
}

int check_vec_ptr(const void *p, const char *what)
{
    if (!p) return fail(HBS_ERR_ARG, "%s is null", what);
    return HBS_OK;
}

}  // namespace

extern "C" int hbs_bicgstab_f64(hbs_solver s, const double *b, double *x, const hbs_solve_opts *opts,
                                hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    CHK(check_vec_ptr(b, "b"));
    CHK(check_vec_ptr(x, "x"));
    hbs_solve_opts o;
    hbs_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.maxit < 0 || !(o.tol >= 0.0)) return fail(HBS_ERR_ARG, "bad tolerance or iteration cap");
    hbs_csr A = s->A;
    DevGuard g(A->ctx->device);
    hipStream_t q = A->ctx->stream;
    const int n = s->n;
    Work<double> w = work64(s);
    SolverState *st = s->st;

    HIPCHK(hipEventRecord(s->e0, q));
    HIPCHK(hipMemcpyAsync(s->X, x, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(launch_state_init(st, 1, o.tol, o.maxit, o.fixed_iters ? 1 : 0, s->sc.ndone, q));
    // R = B - A X ; ||R||^2 -> threshold and first beta (bicgstab_cpu.c:183-190)
    SpmvArgs<double> sa = spmv_args<double>(A, A->val, s->sc);
    sa.x = s->X;
    sa.y = s->R;
    sa.w = b;
    sa.st = st;
    sa.fin = FIN_INIT;
    HIPCHK(launch_spmv(sa, EPI_RESID, A->plan.block_nnz, q));
    HIPCHK(launch_init_vectors(vec_args<double>(s, w, st), q));
    if (o.maxit > 0) CHK(iterate<double>(s, w, st, o.maxit, o.check_every, o.use_graph != 0));
    // final relative residual sqrt(||B - AX||^2 / ||B||^2) (bicgstab_cpu.c:251-255)
    sa.y = s->V;
    sa.fin = FIN_RESID;
    HIPCHK(launch_spmv(sa, EPI_RESID, A->plan.block_nnz, q));
    HIPCHK(launch_dot(vec_args<double>(s, w, st), b, b, FIN_STORE, q));
    HIPCHK(hipMemcpyAsync(x, s->X, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipEventRecord(s->e1, q));
    SolverState hs;
    CHK(read_state(s, 0, &hs));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
    if (info) {
        info->iters = hs.iter;
        info->stop = hs.stop;
        info->sweeps = 0;
        info->status = status_of(hs, o.fixed_iters != 0);
        info->relres = hs.out > 0 ? std::sqrt(hs.resid2 / hs.out) : std::sqrt(hs.resid2);
        info->rel_r0 = hs.rr0 > 0 ? std::sqrt(hs.resid2 / hs.rr0) : 0.0;
        info->time_ms = ms;
    }
    int status = status_of(hs, o.fixed_iters != 0);
    if (status == HBS_ERR_BREAKDOWN) return fail(status, "BiCGStab breakdown after %d iterations", hs.iter);
    if (status == HBS_ERR_NOCONV)
        return fail(status, "BiCGStab reached %d iterations, ||r||/||r0|| = %.3e", hs.iter,
                    hs.rr0 > 0 ? std::sqrt(hs.resid2 / hs.rr0) : 0.0);
    return HBS_OK;
}

extern "C" int hbs_bicgstab_mixed(hbs_solver s, const double *b, double *x, const hbs_solve_opts *opts,
                                  hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    CHK(check_vec_ptr(b, "b"));
    CHK(check_vec_ptr(x, "x"));
    hbs_csr A = s->A;
    if (!A->val32) return fail(HBS_ERR_ARG, "mixed precision needs a matrix created with HBS_CSR_F32");
    hbs_solve_opts o;
    hbs_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.maxit < 0 || !(o.tol >= 0.0) || o.max_sweeps < 1) return fail(HBS_ERR_ARG, "bad options");
    DevGuard g(A->ctx->device);
    hipStream_t q = A->ctx->stream;
    const int n = s->n;
    if (!s->X32) {
        for (float **p : {&s->X32, &s->R32, &s->R032, &s->P32, &s->V32, &s->S32, &s->T32}) CHK(dalloc(p, n));
    }
    Work<double> w64 = work64(s);
    Work<float> w32 = work32(s);
    SolverState *inner = s->st, *outer = s->st + 1;

    HIPCHK(hipEventRecord(s->e0, q));
    HIPCHK(hipMemcpyAsync(s->X, x, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(launch_state_init(outer, 1, o.tol, 0, 0, nullptr, q));
    // FP64 outer residual r = b - A x (reuses R as the FP64 residual vector)
    SpmvArgs<double> sa = spmv_args<double>(A, A->val, s->sc);
    sa.x = s->X;
    sa.y = s->R;
    sa.w = b;
    sa.st = outer;
    sa.fin = FIN_RESID;
    HIPCHK(launch_spmv(sa, EPI_RESID, A->plan.block_nnz, q));
    HIPCHK(launch_dot(vec_args<double>(s, w64, outer), b, b, FIN_STORE, q));
    SolverState hs;
    CHK(read_state(s, 1, &hs));
    const double r0 = std::sqrt(hs.resid2);
    const double bnorm2 = hs.out;
    const double thr = o.tol * r0;
    double rnorm = r0;
    int total_iters = 0, sweeps = 0, stop = 2;
    bool converged = !(r0 > 0.0);
    bool breakdown = false;
    while (!converged && sweeps < o.max_sweeps) {
        // inner tolerance: the requested one, or just what is still missing
        double itol = std::max(o.inner_tol, 0.5 * thr / rnorm);
        if (itol > 0.5) itol = 0.5;
        const int imax = std::min(o.inner_maxit, std::max(1, o.maxit - total_iters));
        HIPCHK(launch_state_init(inner, 1, itol, imax, 0, s->sc.ndone, q));
        HIPCHK(launch_ir_start(vec_args<float>(s, w32, inner), s->R, outer, q));
        CHK(iterate<float>(s, w32, inner, imax, o.check_every, o.use_graph != 0));
        HIPCHK(launch_ir_update(vec_args<float>(s, w32, inner), s->X, q));
        HIPCHK(launch_spmv(sa, EPI_RESID, A->plan.block_nnz, q));
        SolverState hi;
        CHK(read_state(s, 0, &hi));
        CHK(read_state(s, 1, &hs));
        total_iters += hi.iter;
        ++sweeps;
        rnorm = std::sqrt(hs.resid2);
        stop = hi.stop;
        if (hi.stop == 4 && hi.iter == 0) {
            breakdown = true;
            break;
        }
        if (rnorm < thr) converged = true;
        if (total_iters >= o.maxit) break;
    }
    HIPCHK(hipMemcpyAsync(x, s->X, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipEventRecord(s->e1, q));
    HIPCHK(hipEventSynchronize(s->e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
    int status = converged ? HBS_OK : (breakdown ? HBS_ERR_BREAKDOWN : HBS_ERR_NOCONV);
    if (info) {
        info->iters = total_iters;
        info->stop = breakdown ? 4 : stop;
        info->sweeps = sweeps;
        info->status = status;
        info->relres = bnorm2 > 0 ? rnorm / std::sqrt(bnorm2) : rnorm;
        info->rel_r0 = r0 > 0 ? rnorm / r0 : 0.0;
        info->time_ms = ms;
    }
    if (status != HBS_OK)
        return fail(status, "mixed-precision BiCGStab: %d sweeps, %d iterations, ||r||/||r0|| = %.3e", sweeps,
                    total_iters, r0 > 0 ? rnorm / r0 : 0.0);
    return HBS_OK;
}

extern "C" int hbs_adjoint_f64(hbs_solver s, const double *gx, const double *x, double *lambda,
                               double *grad_val, const hbs_solve_opts *opts, hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    CHK(check_vec_ptr(gx, "g"));
    CHK(check_vec_ptr(x, "x"));
    CHK(check_vec_ptr(lambda, "lambda"));
    hbs_csr A = s->A;
    if (!A->tr) return fail(HBS_ERR_ARG, "adjoint needs a matrix created with HBS_CSR_TRANSPOSE");
    DevGuard g(A->ctx->device);
    if (!s->adj) CHK(hbs_solver_create(A->tr, &s->adj));
    HIPCHK(launch_fill_zero(lambda, sizeof(double) * (size_t)s->n, A->ctx->stream));
    int rc = hbs_bicgstab_f64(s->adj, gx, lambda, opts, info);
    if (rc != HBS_OK) return rc;
    if (grad_val)
        HIPCHK(launch_grad_pattern(A->m, A->rowptr, A->col, lambda, x, grad_val, A->ctx->stream));
    HIPCHK(hipStreamSynchronize(A->ctx->stream));
    return HBS_OK;
}

// ---------------------------------------------------------------------------
// one rank's share

struct hbs_shard_s {
    hbs_ctx ctx = nullptr;
    int m_loc = 0, n_shr = 0, n_send = 0;
    hbs_csr merged = nullptr;     // m_loc x (m_loc + n_shr)
    int *send_J = nullptr;        // device [n_send]
    int *recv_src = nullptr;      // device [n_shr]: recv position of shr column c
    double *stage = nullptr;      // device [n_shr]: halo as received
    double *x_ext = nullptr;      // device [m_loc + n_shr], for split x_loc / x_shr callers
    double *d_send = nullptr;     // device [n_send], packed halo before D2H
};

extern "C" int hbs_shard_destroy(hbs_shard s)
{
    if (!s) return HBS_OK;
    DevGuard g(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->merged) hbs_csr_destroy(s->merged);
    dfree(s->send_J);
    dfree(s->recv_src);
    dfree(s->stage);
    dfree(s->x_ext);
    dfree(s->d_send);
    delete s;
    return HBS_OK;
}

// Upload of a SHARD_CSC.  loc and shr become one CSR over the extended
// vector [x_loc ; x_shr]: row i holds its loc entries in ascending local
// column, then its shr entries in the order the reference's CPU halo SpMV
// adds them (MPI_SHARD_CSC_mpi_spmxv_seq, hhp_cpu.c:563-584: the loc column
// scatter first, then recv blocks in rank order, entries in recv order), so
// the order-exact row sum reproduces that function when blocks arrive in
// rank order.
extern "C" int hbs_shard_create(hbs_ctx ctx, const hbs_host_shard *h, hbs_shard *out)
{
    if (!ctx || !h || !out) return fail(HBS_ERR_ARG, "null argument");
    *out = nullptr;
    const int m = h->loc.m;
    const int ns = h->shr.n;
    if (m <= 0 || h->loc.n != m || h->shr.m != m || ns < 0) return fail(HBS_ERR_ARG, "inconsistent shard shape");
    if (!h->loc.I || (ns > 0 && (!h->shr.I || !h->recv.J || !h->recv.I))) return fail(HBS_ERR_ARG, "null shard array");
    const int nrecv = h->recv.num > 0 ? h->recv.I[h->recv.num] : 0;
    if (nrecv != ns) return fail(HBS_ERR_ARG, "recv list covers %d of %d shr columns", nrecv, ns);
    const int nsend = h->send.num > 0 ? h->send.I[h->send.num] : 0;
    std::vector<int> cnt((size_t)m + 1, 0);
    for (int c = 0; c < m; ++c)
        for (int j = h->loc.I[c]; j < h->loc.I[c + 1]; ++j) {
            const int r = h->loc.J[j];
            if (r < 0 || r >= m) return fail(HBS_ERR_ARG, "loc row index out of range");
            cnt[(size_t)r + 1]++;
        }
    std::vector<int> pos((size_t)ns, -1);
    for (int p = 0; p < ns; ++p) {
        const int c = h->recv.J[p];
        if (c < 0 || c >= ns || pos[(size_t)c] >= 0) return fail(HBS_ERR_ARG, "recv.J is not a permutation of the shr columns");
        pos[(size_t)c] = p;
    }
    for (int c = 0; c < ns; ++c)
        for (int j = h->shr.I[c]; j < h->shr.I[c + 1]; ++j) {
            const int r = h->shr.J[j];
            if (r < 0 || r >= m) return fail(HBS_ERR_ARG, "shr row index out of range");
            cnt[(size_t)r + 1]++;
        }
    for (int i = 0; i < m; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
    const int nnz = cnt[(size_t)m];
    std::vector<int> col((size_t)std::max(nnz, 1));
    std::vector<double> val((size_t)std::max(nnz, 1));
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int c = 0; c < m; ++c)
        for (int j = h->loc.I[c]; j < h->loc.I[c + 1]; ++j) {
            const int d = fill[(size_t)h->loc.J[j]]++;
            col[(size_t)d] = c;
            val[(size_t)d] = h->loc.val[j];
        }
    for (int p = 0; p < ns; ++p) {
        const int c = h->recv.J[p];
        for (int j = h->shr.I[c]; j < h->shr.I[c + 1]; ++j) {
            const int d = fill[(size_t)h->shr.J[j]]++;
            col[(size_t)d] = m + c;
            val[(size_t)d] = h->shr.val[j];
        }
    }
    for (int i = 0; i < nsend; ++i)
        if (h->send.J[i] < 0 || h->send.J[i] >= m) return fail(HBS_ERR_ARG, "send.J entry out of range");
    DevGuard g(ctx->device);
    hbs_shard s = new hbs_shard_s();
    s->ctx = ctx;
    s->m_loc = m;
    s->n_shr = ns;
    s->n_send = nsend;
    auto bail = [&](int rc) {
        std::string e = g_err;
        hbs_shard_destroy(s);
        g_err = e;
        return rc;
    };
    int rc = hbs_csr_create(ctx, m, m + ns, nnz, cnt.data(), col.data(), val.data(), 0, &s->merged);
    if (rc) return bail(rc);
    if ((rc = dalloc(&s->send_J, nsend)) || (rc = dalloc(&s->recv_src, ns)) || (rc = dalloc(&s->stage, ns)))
        return bail(rc);
    if ((nsend && hipMemcpy(s->send_J, h->send.J, sizeof(int) * nsend, hipMemcpyHostToDevice) != hipSuccess) ||
        (ns && hipMemcpy(s->recv_src, pos.data(), sizeof(int) * ns, hipMemcpyHostToDevice) != hipSuccess))
        return bail(fail(HBS_ERR_HIP, "shard list upload failed"));
    *out = s;
    return HBS_OK;
}

extern "C" int hbs_shard_info(hbs_shard s, int *m_loc, int *n_shr, int *n_send, int *nnz)
{
    if (!s) return fail(HBS_ERR_ARG, "null shard");
    if (m_loc) *m_loc = s->m_loc;
    if (n_shr) *n_shr = s->n_shr;
    if (n_send) *n_send = s->n_send;
    if (nnz) *nnz = s->merged->nnz;
    return HBS_OK;
}

extern "C" int hbs_shard_pack(hbs_shard s, const double *x_loc, double *send)
{
    if (!s || (s->n_send > 0 && (!x_loc || !send))) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    HIPCHK(launch_gather_f64(s->n_send, x_loc, s->send_J, send, s->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_shard_unpack(hbs_shard s, const double *recv, double *x_ext)
{
    if (!s || (s->n_shr > 0 && (!recv || !x_ext))) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    HIPCHK(launch_gather_f64(s->n_shr, recv, s->recv_src, x_ext + s->m_loc, s->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_shard_spmv(hbs_shard s, double *x_ext, const double *recv_host, double *y)
{
    if (!s || !x_ext || !y) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    if (recv_host && s->n_shr > 0) {
        HIPCHK(hipMemcpyAsync(s->stage, recv_host, sizeof(double) * s->n_shr, hipMemcpyHostToDevice, s->ctx->stream));
        CHK(hbs_shard_unpack(s, s->stage, x_ext));
    }
    return hbs_csr_spmv_f64(s->merged, 1.0, x_ext, 0.0, y);
}

extern "C" int hbs_shard_dot_partial(hbs_shard s, const double *a, const double *b, double *out)
{
    if (!s) return fail(HBS_ERR_ARG, "null shard");
    return hbs_dot_f64(s->ctx, s->m_loc, a, b, out);
}

extern "C" int hbs_shard_csr(hbs_shard s, hbs_csr *out)
{
    if (!s || !out) return fail(HBS_ERR_ARG, "null argument");
    *out = s->merged;
    return HBS_OK;
}

namespace {
int shard_scratch(hbs_shard s)
{
    if (!s->x_ext) CHK(dalloc(&s->x_ext, (size_t)s->m_loc + s->n_shr));
    if (!s->d_send) CHK(dalloc(&s->d_send, (size_t)std::max(s->n_send, 1)));
    return HBS_OK;
}
}  // namespace

extern "C" int hbs_shard_pack_host(hbs_shard s, const double *x_loc, double *send_host)
{
    if (!s || (s->n_send > 0 && (!x_loc || !send_host))) return fail(HBS_ERR_ARG, "bad argument");
    if (s->n_send == 0) return HBS_OK;
    DevGuard g(s->ctx->device);
    CHK(shard_scratch(s));
    CHK(hbs_shard_pack(s, x_loc, s->d_send));
    HIPCHK(hipMemcpyAsync(send_host, s->d_send, sizeof(double) * s->n_send, hipMemcpyDeviceToHost,
                          s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_shard_spmv2(hbs_shard s, const double *x_loc, const double *recv_host, double *x_shr,
                               double alpha, double beta, double *y)
{
    if (!s || !x_loc || !y) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    CHK(shard_scratch(s));
    hipStream_t q = s->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->x_ext, x_loc, sizeof(double) * s->m_loc, hipMemcpyDeviceToDevice, q));
    if (s->n_shr > 0) {
        if (recv_host) {
            HIPCHK(hipMemcpyAsync(s->stage, recv_host, sizeof(double) * s->n_shr, hipMemcpyHostToDevice, q));
            CHK(hbs_shard_unpack(s, s->stage, s->x_ext));
            if (x_shr)
                HIPCHK(hipMemcpyAsync(x_shr, s->x_ext + s->m_loc, sizeof(double) * s->n_shr,
                                      hipMemcpyDeviceToDevice, q));
        } else {
            if (!x_shr) return fail(HBS_ERR_ARG, "no halo: pass recv_host or x_shr");
            HIPCHK(hipMemcpyAsync(s->x_ext + s->m_loc, x_shr, sizeof(double) * s->n_shr,
                                  hipMemcpyDeviceToDevice, q));
        }
    }
    return hbs_csr_spmv_f64(s->merged, alpha, s->x_ext, beta, y);
}
