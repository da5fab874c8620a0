// hbs_api.cpp -- C-ABI of the engine: contexts, device memory, BLAS-1,
// device CSR, SpMV / SpMM and their benchmarks.  BiCGStab is in
// hbs_solver.cpp, one rank's share in hbs_shard.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hbs_objects.h"

using namespace hbs;

namespace {
thread_local std::string g_err;
int g_block_nnz = 0;      // 0: by average row length (defaults from sweeps on MI355X)
int g_serial_max_avg = 64;
int g_max_groups = 0;     // 0: kMaxGroups, or the resident workgroups of the 512-thread SpMV
int g_stream_nt = 0;
int g_spmv_npt = 0;       // entries per thread per pass of 512-thread SpMV blocks (0: 4)
int g_sort_pct = 70;      // kind-2 blocks: sorted gathers touch <= 70% as many lines as entries
int g_xseg = 2;           // x-segmented SpMV: 0 never, 1 every one-system matrix, 2 scattered columns
int g_xseg_split = 1;     // k_xseg workgroups per x segment
int g_spmm_order = 1;     // multi-column SpMM on a 3-D grid: 1 y-slab runs (plan_grid_runs), 0 rows in order
int g_spmm_dynamic = 1;   // 64-column SpMM sweeps: 1 steps from per-XCD counters, 0 fixed stride
int g_block3 = 1;         // 3x3-block matrices: kind-3 row blocks (1) or plain CSR-stream blocks (0)
int g_diag_split = 1;     // FP32 SpMVs of square x-segmented matrices: diagonal apart (1) or in the entries (0)
// csr_upload options
constexpr int kUpForceXseg = 1;   // x-segmented whatever the column scatter
constexpr int kUpNoSplit = 2;     // no FP32 off-diagonal part (the part itself)
}  // namespace

#define fail ::hbs::set_error

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

int hbs::DevVecPlan::build(int nn)
{
    release();
    std::vector<VecBlock> vb;
    const int e = vec_elems_per_block(nn);
    plan_vec(0, nn, e, 0, vb);
    nvb = (int)vb.size();
    n = nn;
    epb = e;
    CHK(dalloc(&blocks, vb.size()));
    CHK(upload(blocks, vb));
    SysRange sr{0, 0, 0, nvb};
    CHK(dalloc(&sys, 1));
    CHK(upload(sys, std::vector<SysRange>{sr}));
    CHK(dalloc(&part, (size_t)std::max(nvb, 1) * kMaxDots));
    return HBS_OK;
}

void hbs::DevVecPlan::release()
{
    dfree(blocks);
    dfree(sys);
    dfree(part);
    n = -1;
    nvb = 0;
}

// ---------------------------------------------------------------------------
// context

extern "C" const char *hbs_last_error(void) { return g_err.c_str(); }
extern "C" const char *hbs_version(void) { return "hbs 0.2 (gfx950)"; }

extern "C" int hbs_device_count(int *out)
{
    if (!out) return fail(HBS_ERR_ARG, "null out");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
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
    ctx->vp.release();
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
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(HBS_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(HBS_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    HIPCHK(hipSetDevice(device));
    hbs_ctx c = new hbs_ctx_s();
    c->device = device;
    int rc = HBS_OK;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        // default priority, non-blocking with respect to the null stream
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
            rc = fail(HBS_ERR_HIP, "stream creation failed");
        c->own_stream = true;
    }
    if (rc == HBS_OK && (hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess))
        rc = fail(HBS_ERR_HIP, "event creation failed");
    if (rc == HBS_OK) rc = dalloc(&c->st, 1);
    if (rc == HBS_OK && hipHostMalloc((void **)&c->hst, sizeof(SolverState)) != hipSuccess)
        rc = fail(HBS_ERR_ALLOC, "pinned host allocation failed");
    if (rc != HBS_OK) {
        std::string e = g_err;
        hbs_ctx_destroy(c);
        g_err = e;
        return rc;
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

extern "C" int hbs_pinned_alloc(hbs_ctx ctx, size_t bytes, void **host)
{
    if (!ctx || !host) return fail(HBS_ERR_ARG, "null argument");
    *host = nullptr;
    DevGuard g(ctx->device);
    if (hipHostMalloc(host, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *host = nullptr;
        return fail(HBS_ERR_ALLOC, "pinned host allocation of %zu bytes failed", bytes);
    }
    return HBS_OK;
}

extern "C" int hbs_pinned_free(hbs_ctx ctx, void *host)
{
    if (!ctx) return fail(HBS_ERR_ARG, "null ctx");
    if (!host) return HBS_OK;
    DevGuard g(ctx->device);
    HIPCHK(hipHostFree(host));
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
    if (n == 0) {
        *out_host = 0.0;
        return HBS_OK;
    }
    DevGuard g(ctx->device);
    if (ctx->vp.n != n || ctx->vp.epb != vec_elems_per_block(n)) CHK(ctx->vp.build(n));
    VecArgs<double> a{};
    a.blocks = ctx->vp.blocks;
    a.sys = ctx->vp.sys;
    a.nvb = ctx->vp.nvb;
    a.part_out = ctx->vp.part;
    HIPCHK(launch_dot(a, x, y, ctx->stream));
    HIPCHK(launch_finalize(ctx->st, ctx->vp.sys, 1, ctx->vp.part, ctx->vp.nvb, 0, FIN_STORE, ctx->stream));
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

extern "C" int hbs_set_spmv_plan(int block_nnz, int serial_max_avg)
{
    if (block_nnz != 0 && (block_nnz < 256 || block_nnz > kMaxBlockNnz || (block_nnz % 256) != 0))
        return fail(HBS_ERR_ARG, "block_nnz must be 0 or a multiple of 256 in [256, %d]", kMaxBlockNnz);
    if (serial_max_avg < 0) return fail(HBS_ERR_ARG, "serial_max_avg must be >= 0");
    g_block_nnz = block_nnz;
    g_serial_max_avg = serial_max_avg;
    return HBS_OK;
}

extern "C" int hbs_set_spmv_groups(int max_groups)
{
    if (max_groups < 0 || max_groups > 65536) return fail(HBS_ERR_ARG, "max_groups must be in [0, 65536]");
    g_max_groups = max_groups;
    return HBS_OK;
}

extern "C" int hbs_set_spmv_sort(int max_line_pct)
{
    if (max_line_pct < 0 || max_line_pct > 100) return fail(HBS_ERR_ARG, "max_line_pct must be in [0, 100]");
    g_sort_pct = max_line_pct;
    return HBS_OK;
}

extern "C" int hbs_set_vec_blocks(int target)
{
    if (target < 1 || target > 65536) return fail(HBS_ERR_ARG, "target must be in [1, 65536]");
    g_vec_target = target;
    return HBS_OK;
}

extern "C" int hbs_set_spmv_stream_nt(int on)
{
    g_stream_nt = on ? 1 : 0;
    return HBS_OK;
}

extern "C" int hbs_set_spmv_pass(int entries_per_thread)
{
    if (entries_per_thread != 0 && entries_per_thread != 4 && entries_per_thread != 8)
        return fail(HBS_ERR_ARG, "entries_per_thread must be 0, 4 or 8");
    g_spmv_npt = entries_per_thread;
    return HBS_OK;
}

extern "C" int hbs_set_spmv_block3(int mode)
{
    if (mode != 0 && mode != 1) return fail(HBS_ERR_ARG, "mode must be 0 or 1");
    g_block3 = mode;
    return HBS_OK;
}

extern "C" int hbs_detect_block3(int m, int n, const int *rowptr, const int *col, int *yes)
{
    if (m < 0 || n < 0 || !rowptr || !col || !yes) return fail(HBS_ERR_ARG, "bad argument");
    *yes = detect_b3(m, n, rowptr, col) ? 1 : 0;
    return HBS_OK;
}

extern "C" int hbs_csr_block3(hbs_csr a, int *on)
{
    if (!a || !on) return fail(HBS_ERR_ARG, "bad argument");
    *on = a->bval == nullptr ? 0 : a->b3var ? 2 : 1;
    return HBS_OK;
}

extern "C" int hbs_set_spmm_order(int mode)
{
    if (mode != 0 && mode != 1) return fail(HBS_ERR_ARG, "mode must be 0 (rows in order) or 1 (grid runs)");
    g_spmm_order = mode;
    return HBS_OK;
}

extern "C" int hbs_set_spmm_tuning(int slab_lines, int groups, int dynamic, int wide_loads)
{
    if (slab_lines < 0 || groups < 0 || (groups > 0 && (groups < 8 || groups % 8 != 0)) || dynamic < 0 ||
        dynamic > kCtrPerXcd || (wide_loads != 0 && wide_loads != 1))
        return fail(HBS_ERR_ARG, "slab_lines >= 0; groups 0 or a multiple of 8; dynamic 0..%d; wide_loads 0 or 1",
                    kCtrPerXcd);
    g_grid_slab = slab_lines;
    g_spmm_groups = groups;
    g_spmm_dynamic = dynamic > 0;
    if (dynamic > 0) g_spmm_counters = dynamic;
    g_spmm_v16 = wide_loads;
    return HBS_OK;
}

extern "C" int hbs_plan_grid_runs(int m, const int *rowptr, const int *col, int *runs, int cap, int *nruns,
                                  int *len)
{
    if (m < 0 || !rowptr || !col || !nruns || !len || cap < 0 || (cap > 0 && !runs))
        return fail(HBS_ERR_ARG, "bad argument");
    std::vector<int> r;
    int l = 0;
    const int nr = plan_grid_runs(m, rowptr, col, r, l);
    if (nr > cap) return fail(HBS_ERR_ARG, "%d runs do not fit the %d given", nr, cap);
    if (nr > 0) std::memcpy(runs, r.data(), sizeof(int) * (size_t)nr);
    *nruns = nr;
    *len = l;
    return HBS_OK;
}

extern "C" int hbs_csr_spmm_runs(hbs_csr a, int *nruns, int *len)
{
    if (!a || !nruns || !len) return fail(HBS_ERR_ARG, "bad argument");
    *nruns = a->rr_nruns;
    *len = a->rr_len;
    return HBS_OK;
}

extern "C" int hbs_set_xseg_split(int split)
{
    if (split < 1 || split > 16) return fail(HBS_ERR_ARG, "split must be in [1, 16]");
    g_xseg_split = split;
    return HBS_OK;
}

extern "C" int hbs_set_spmv_xseg(int mode)
{
    if (mode < 0 || mode > 2) return fail(HBS_ERR_ARG, "mode must be 0, 1 or 2");
    g_xseg = mode;
    return HBS_OK;
}

extern "C" int hbs_set_xseg_fuse(int mode)
{
    if (mode < 0 || mode > 1) return fail(HBS_ERR_ARG, "mode must be 0 or 1");
    g_xseg_fuse = mode;
    return HBS_OK;
}

extern "C" int hbs_set_spmv_diag_split(int mode)
{
    if (mode < 0 || mode > 1) return fail(HBS_ERR_ARG, "mode must be 0 or 1");
    g_diag_split = mode;
    return HBS_OK;
}

extern "C" int hbs_csr_diag_split(hbs_csr a, int *mode)
{
    if (!a || !mode) return fail(HBS_ERR_ARG, "null argument");
    *mode = !a->od ? 0 : a->diag32 ? 2 : 1;
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
    dfree(a->scol);
    dfree(a->sval);
    dfree(a->sval32);
    dfree(a->sperm);
    dfree(a->ssrc);
    dfree(a->xs_sbase);
    dfree(a->xs_cbase);
    dfree(a->xs_gbase);
    dfree(a->xs_gch);
    dfree(a->xs_off);
    dfree(a->xs_src);
    dfree(a->xs_col);
    dfree(a->xs_pos);
    dfree(a->rperm);
    dfree(a->rr_run);
    dfree(a->bcol);
    dfree(a->bval);
    dfree(a->bval32);
    dfree(a->bsrc);
    dfree(a->spmm_ctr);
    dfree(a->xs_delta);
    dfree(a->xs_val);
    dfree(a->xs_prod);
    dfree(a->xs_val32);
    dfree(a->xs_prod32);
    dfree(a->plan.blocks);
    dfree(a->plan.aux);
    dfree(a->plan.chunks);
    dfree(a->plan.groups);
    dfree(a->sys);
    dfree(a->chunk_sums);
    dfree(a->part);
    dfree(a->tr_perm);
    if (a->evs) {
        for (int i = 0; i < a->nevs; ++i) (void)hipEventDestroy(a->evs[i]);
        delete[] a->evs;
    }
    if (a->tr) hbs_csr_destroy(a->tr);
    if (a->od) hbs_csr_destroy(a->od);
    dfree(a->diag32);
    dfree(a->od_src);
    dfree(a->diag_src);
    delete a;
    return HBS_OK;
}

namespace {

int validate_csr(int m, int n, int nnz, const int *rowptr, const int *col, const double *val)
{
    if (m < 0 || n < 0 || nnz < 0) return fail(HBS_ERR_ARG, "negative dimension");
    // the SpMV gathers x through 32-bit byte offsets from its base
    if ((long long)n * (long long)sizeof(double) > (long long)UINT32_MAX)
        return fail(HBS_ERR_ARG, "n = %d columns exceed the SpMV's 32-bit x offsets (at most %lld)", n,
                    (long long)(UINT32_MAX / sizeof(double)));
    if (!rowptr || (nnz > 0 && (!col || !val))) return fail(HBS_ERR_ARG, "null CSR array");
    if (rowptr[0] != 0 || rowptr[m] != nnz) return fail(HBS_ERR_ARG, "rowptr[0] must be 0 and rowptr[m] == nnz");
    for (int i = 0; i < m; ++i)
        if (rowptr[i + 1] < rowptr[i]) return fail(HBS_ERR_ARG, "rowptr not monotone at row %d", i);
    for (int j = 0; j < nnz; ++j)
        if (col[j] < 0 || col[j] >= n) return fail(HBS_ERR_ARG, "column index %d out of range at entry %d", col[j], j);
    return HBS_OK;
}

// Systems: sys_rows[nsys + 1] row offsets of independent diagonal blocks
// (nullptr: one system of all m rows).  Each system gets its own row blocks
// and a share of the SpMV workgroups proportional to its entries.
// Entries per row block (measured on MI355X).  Short rows (under 10 on
// average) gather little per row and do best with blocks of 1024 entries
// and many of them.  Longer rows gain from big blocks: column-sorted gathers
// share more x lines.  Up to 6144 entries, three 512-thread workgroups (4
// entries per thread per pass) fit a CU's LDS and keep their loads in
// flight.  The blocks are sized so that the resident workgroups (`resident`,
// 0 if unknown) each get the same whole number of them, with 1% slack for
// blocks ending at row boundaries, and a small matrix still fills the GPU.
int default_block_nnz(int m, int nnz, int resident)
{
    if ((long long)nnz < 10LL * std::max(m, 1)) return 1024;
    if (resident <= 0) resident = 1024;
    const long long per_round = 6144LL * resident;
    const long long k = std::max(1LL, ((long long)nnz + per_round - 1) / per_round);
    const long long bn = ((long long)((double)nnz / (double)(k * resident) * 1.01) + 63) / 64 * 64;
    return (int)std::min<long long>(6144, std::max<long long>(2048, bn));
}

int csr_upload(hbs_ctx ctx, int m, int n, int nnz, const int *rowptr, const int *col, const double *val, int flags,
               int nsys, const int *sys_rows, hbs_csr *out, int up = 0);

// FP32 SpMVs of a square x-segmented matrix: the diagonal apart and the
// off-diagonal entries (CSR order) as an x-segmented matrix of their own, so
// the product pass and the row sums move no diagonal entry; each row adds
// d_i x_i to its off-diagonal sum (x_i alone when every diagonal entry is
// 1, as after a Jacobi scaling).  A row's first entry in column i is its
// diagonal; further ones stay off-diagonal entries, and a row without one
// has d_i = 0.  A no-op when the off-diagonal part would not be x-segmented.
int split_diagonal(hbs_csr a, const int *rowptr, const int *col, const double *val)
{
    const int m = a->m, nnz = a->nnz;
    std::vector<int> orp((size_t)m + 1, 0), ocol, osrc, dsrc((size_t)m, nnz);
    ocol.reserve((size_t)nnz);
    osrc.reserve((size_t)nnz);
    bool unit = true;
    for (int r = 0; r < m; ++r) {
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            if (col[j] == r && dsrc[(size_t)r] == nnz) {
                dsrc[(size_t)r] = j;
            } else {
                ocol.push_back(col[j]);
                osrc.push_back(j);
            }
        }
        orp[(size_t)r + 1] = (int)ocol.size();
        unit = unit && dsrc[(size_t)r] < nnz && (float)val[dsrc[(size_t)r]] == 1.0f;
    }
    const int onz = (int)ocol.size();
    if (onz == 0) return HBS_OK;
    std::vector<double> oval((size_t)onz);
    for (int k = 0; k < onz; ++k) oval[(size_t)k] = val[osrc[(size_t)k]];
    hbs_csr od = nullptr;
    CHK(csr_upload(a->ctx, m, m, onz, orp.data(), ocol.data(), oval.data(), HBS_CSR_F32, 1, nullptr, &od,
                   kUpForceXseg | kUpNoSplit));
    if (od->xs_nseg == 0) {
        hbs_csr_destroy(od);
        return HBS_OK;
    }
    a->od = od;
    int rc;
    if ((rc = dalloc(&a->od_src, (size_t)onz)) || (rc = upload(a->od_src, osrc)) ||
        (rc = dalloc(&a->diag_src, (size_t)m)) || (rc = upload(a->diag_src, dsrc)))
        return rc;
    if (!unit) {
        if ((rc = dalloc(&a->diag32, (size_t)m))) return rc;
        if (launch_gather_to_f32(m, a->val, a->diag_src, a->diag32, a->ctx->stream) != hipSuccess ||
            hipStreamSynchronize(a->ctx->stream) != hipSuccess)
            return fail(HBS_ERR_HIP, "FP32 diagonal copy failed");
    }
    return HBS_OK;
}

int csr_upload(hbs_ctx ctx, int m, int n, int nnz, const int *rowptr, const int *col, const double *val, int flags,
               int nsys, const int *sys_rows, hbs_csr *out, int up)
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
    // col / val / val32 carry kRowPad zero entries past nnz: the 64-column
    // SpMM reads whole chunks of a row's entries through the scalar cache
    const size_t padded = (size_t)nnz + kRowPad;
    if ((rc = dalloc(&a->rowptr, (size_t)m + 1)) || (rc = dalloc(&a->col, padded)) || (rc = dalloc(&a->val, padded)))
        return bail(rc);
    if (hipMemcpy(a->rowptr, rowptr, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (nnz && hipMemcpy(a->col, col, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) ||
        (nnz && hipMemcpy(a->val, val, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemset(a->col + nnz, 0, sizeof(int) * kRowPad) != hipSuccess ||
        hipMemset(a->val + nnz, 0, sizeof(double) * kRowPad) != hipSuccess)
        return bail(fail(HBS_ERR_HIP, "CSR upload failed"));
    if (flags & HBS_CSR_F32) {
        if ((rc = dalloc(&a->val32, padded))) return bail(rc);
        if (hipMemset(a->val32 + nnz, 0, sizeof(float) * kRowPad) != hipSuccess)
            return bail(fail(HBS_ERR_HIP, "CSR upload failed"));
        if (launch_to_f32(nnz, a->val, a->val32, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return bail(fail(HBS_ERR_HIP, "FP32 value copy failed"));
    }
    HostPlan hp;
    std::vector<SysRange> ranges;
    // x-segmented SpMV for one system whose columns spread over an x larger
    // than an XCD's L2 share (8 MiB of FP64): the gathers of 128-byte-line
    // windows of kSegBlockNnz entries touch a new line for more than half
    // the entries
    bool xseg = (nsys <= 1 || !sys_rows) && nnz > 0 &&
                (g_xseg == 1 || (up & kUpForceXseg) ||
                 (g_xseg == 2 && (long long)n * 8 >= (8LL << 20) && column_scatter(col, nnz, kSegBlockNnz) > 0.5));
    int block_nnz =
        xseg ? (g_block_nnz > 2048 ? g_block_nnz : kSegBlockNnz)
             : g_block_nnz > 0 ? g_block_nnz : default_block_nnz(m, nnz, spmv_resident_groups(6144, 4, ctx->device));
    if (!sys_rows) {
        a->sys_rows = {0, m};
        nsys = 1;
    } else {
        a->sys_rows.assign(sys_rows, sys_rows + nsys + 1);
    }
    a->nsys = nsys;
    // 3x3-block matrices (k_spmv3): row blocks of whole block rows, at most
    // block_nnz entries and one block per thread, one column index per block
    // of 9 entries (the stream reads 8 + 4/9 bytes per entry instead of 12)
    const bool b3 = g_block3 && !xseg && nsys == 1 && spmv_threads(block_nnz) == kWideThreads &&
                    block_nnz >= 3 * kLongRow && detect_b3(m, n, rowptr, col);
    const int group_nnz = block_nnz;   // row blocks' entry limit
    if (b3) block_nnz = kB3BlockNnz;   // the launches' LDS
    int npt = b3 ? kB3Npt
                 : spmv_threads(block_nnz) == kThreads ? (block_nnz <= 4 * kThreads ? 4 : 8)
                                                       : (xseg ? kPreNpt : g_spmv_npt ? g_spmv_npt : 4);
    if (prepare_spmv(block_nnz, npt) != hipSuccess)
        return bail(fail(HBS_ERR_HIP, "SpMV LDS limit for %d-entry row blocks not granted", block_nnz));
    // SpMV workgroups: 512-thread ones as many as are resident at once, each
    // looping over its run of blocks with the next block's loads in flight;
    // 256-thread ones up to kMaxGroups
    int max_groups = g_max_groups;
    if (max_groups == 0) {
        max_groups = kMaxGroups;
        if (spmv_threads(block_nnz) == kWideThreads) {
            const int res = spmv_resident_groups(block_nnz, npt, ctx->device, xseg);
            if (res > 0) max_groups = std::min(max_groups, res);
        }
    }
    for (int q = 0; q < nsys; ++q) {
        const int r0 = a->sys_rows[(size_t)q], r1 = a->sys_rows[(size_t)q + 1];
        const int rb0 = (int)hp.blocks.size(), g0 = (int)hp.groups.size();
        if (b3) a->b3var += plan_rows_b3(r0, r1 - r0, rowptr, group_nnz, q, hp);
        else plan_rows(r0, r1 - r0, rowptr, block_nnz, g_serial_max_avg, q, hp);
        const long long nz = (long long)rowptr[r1] - rowptr[r0];
        const int want = nsys == 1 ? max_groups
                                   : (int)std::max(1LL, nnz > 0 ? (long long)max_groups * nz / nnz : 1LL);
        plan_groups(rb0, (int)hp.blocks.size(), want, q, hp);
        ranges.push_back(SysRange{g0, (int)hp.groups.size(), 0, 0});
    }
    a->sys_host = ranges;
    // step counters of the 64-column SpMM sweeps, zero until a sweep uses
    // them (each launch leaves them at zero)
    if (g_spmm_dynamic) {
        const size_t nctr = (size_t)kSpmmCtrTiles * kXcds * kCtrSet * kCtrStride;
        if ((rc = dalloc(&a->spmm_ctr, nctr))) return bail(rc);
        if (hipMemset(a->spmm_ctr, 0, sizeof(int) * nctr) != hipSuccess)
            return bail(fail(HBS_ERR_HIP, "SpMM counter initialisation failed"));
    }
    // multi-column SpMM traversal of a 3-D grid operator
    if (g_spmm_order == 1 && nsys == 1 && m == n) {
        std::vector<int> runs;
        int len = 0;
        if (plan_grid_runs(m, rowptr, col, runs, len) > 0) {
            if ((rc = dalloc(&a->rr_run, runs.size())) || (rc = upload(a->rr_run, runs))) return bail(rc);
            a->rr_len = len;
            a->rr_nruns = (int)runs.size();
        }
    }
    XsegHost xh;
    if (xseg) {
        // the segment's chunk tables must fit k_xseg's LDS beside x
        plan_xseg(col, n, hp, xh);
        if (xseg_lds_bytes(8, xh.max_chunks, xh.max_granules) > (size_t)kXsegMaxLds ||
            prepare_xseg(xh.max_chunks, xh.max_granules) != hipSuccess) {
            // CSR-stream SpMV over the same row blocks
            xseg = false;
            npt = 4;
            if (prepare_spmv(block_nnz, npt) != hipSuccess)
                return bail(fail(HBS_ERR_HIP, "SpMV LDS limit for %d-entry row blocks not granted", block_nnz));
        }
    }
    if (xseg) {
        a->xs_nseg = xh.nseg;
        a->xs_split = g_xseg_split;
        a->xs_max_chunks = xh.max_chunks;
        a->xs_max_granules = xh.max_granules;
        std::vector<double> xv((size_t)nnz);
        for (int g = 0; g < nnz; ++g) xv[(size_t)g] = val[xh.src[(size_t)g]];
        if ((rc = dalloc(&a->xs_sbase, xh.sbase.size())) || (rc = dalloc(&a->xs_cbase, xh.cbase.size())) ||
            (rc = dalloc(&a->xs_gbase, xh.gbase.size())) || (rc = upload(a->xs_cbase, xh.cbase)) ||
            (rc = dalloc(&a->xs_gch, xh.gch.size())) || (rc = dalloc(&a->xs_off, xh.off.size())) ||
            (rc = dalloc(&a->xs_src, (size_t)nnz)) || (rc = dalloc(&a->xs_col, (size_t)nnz)) ||
            (rc = dalloc(&a->xs_pos, padded)) || (rc = dalloc(&a->xs_delta, (size_t)nnz)) ||
            (rc = dalloc(&a->xs_val, (size_t)nnz)) || (rc = dalloc(&a->xs_prod, padded)) ||
            (rc = upload(a->xs_sbase, xh.sbase)) || (rc = upload(a->xs_gbase, xh.gbase)) ||
            (rc = upload(a->xs_gch, xh.gch)) || (rc = upload(a->xs_off, xh.off)) || (rc = upload(a->xs_src, xh.src)) ||
            (rc = upload(a->xs_col, xh.xcol)) || (rc = upload(a->xs_pos, xh.ppos)) ||
            (rc = upload(a->xs_delta, xh.delta)) || (rc = upload(a->xs_val, xv)))
            return bail(rc);
        // serial row sums of power-law blocks in descending row length
        std::vector<unsigned short> rp16;
        if (plan_row_order(rowptr, hp, m, rp16) > 0) {
            if ((rc = dalloc(&a->rperm, rp16.size())) || (rc = upload(a->rperm, rp16))) return bail(rc);
        }
        // the passes read a block's products clamped to its last entry, or
        // to entry nnz (padding) for a block without entries
        if (hipMemset(a->xs_pos + nnz, 0, sizeof(unsigned short) * kRowPad) != hipSuccess ||
            hipMemset(a->xs_prod + nnz, 0, sizeof(double) * kRowPad) != hipSuccess)
            return bail(fail(HBS_ERR_HIP, "x-segment buffers"));
        if (a->val32) {
            if ((rc = dalloc(&a->xs_val32, (size_t)nnz)) || (rc = dalloc(&a->xs_prod32, padded))) return bail(rc);
            if (hipMemset(a->xs_prod32 + nnz, 0, sizeof(float) * kRowPad) != hipSuccess)
                return bail(fail(HBS_ERR_HIP, "x-segment buffers"));
            if (launch_to_f32(nnz, a->xs_val, a->xs_val32, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                return bail(fail(HBS_ERR_HIP, "FP32 value copy failed"));
        }
        if (a->val32 && m == n && g_diag_split && !(up & kUpNoSplit) && (rc = split_diagonal(a, rowptr, col, val)))
            return bail(rc);
    }
    if (b3) {
        // k_spmv3's threads past a row block's blocks read up to kWideThreads
        // values and columns past it: padding of zero values and column 0
        std::vector<int> bsrc, bcol;
        b3_layout(hp, rowptr, col, bsrc, bcol);
        std::vector<double> bv((size_t)nnz + kWideThreads, 0.0);
        for (int j = 0; j < nnz; ++j) bv[(size_t)j] = val[bsrc[(size_t)j]];
        bcol.resize(bcol.size() + kWideThreads, 0);
        if ((rc = dalloc(&a->bcol, bcol.size())) || (rc = dalloc(&a->bval, bv.size())) ||
            (rc = dalloc(&a->bsrc, (size_t)nnz)) || (rc = upload(a->bcol, bcol)) || (rc = upload(a->bval, bv)) ||
            (rc = upload(a->bsrc, bsrc)))
            return bail(rc);
        if (a->val32) {
            if ((rc = dalloc(&a->bval32, bv.size()))) return bail(rc);
            if (launch_to_f32((int)bv.size(), a->bval, a->bval32, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                return bail(fail(HBS_ERR_HIP, "FP32 value copy failed"));
        }
    }
    // column-sorted gathers where they share cache lines
    std::vector<int> src;
    std::vector<unsigned short> spos;
    if (!xseg && plan_sorted(col, nnz, g_sort_pct, hp, src, spos) > 0) {
        // When the column index and the position in the block fit in 32 bits
        // together, the position rides in the column's high bits and the
        // separate position stream is not needed.
        auto bits = [](long long v) {
            int b = 1;
            while (b < 31 && (1LL << b) <= v) ++b;
            return b;
        };
        const int cb = bits((long long)n - 1), pb = bits((long long)block_nnz - 1);
        a->pos_shift = cb + pb <= 32 ? cb : 0;
        std::vector<int> sc((size_t)nnz);
        std::vector<double> sv((size_t)nnz);
        for (int j = 0; j < nnz; ++j) {
            const unsigned c = (unsigned)col[src[(size_t)j]];
            sc[(size_t)j] = (int)(a->pos_shift ? c | ((unsigned)spos[(size_t)j] << a->pos_shift) : c);
            sv[(size_t)j] = val[src[(size_t)j]];
        }
        if ((rc = dalloc(&a->scol, nnz)) || (rc = dalloc(&a->sval, nnz)) || (rc = dalloc(&a->ssrc, nnz)) ||
            (rc = upload(a->scol, sc)) || (rc = upload(a->sval, sv)) || (rc = upload(a->ssrc, src)))
            return bail(rc);
        if (!a->pos_shift && ((rc = dalloc(&a->sperm, nnz)) || (rc = upload(a->sperm, spos)))) return bail(rc);
        if (a->val32) {
            if ((rc = dalloc(&a->sval32, nnz))) return bail(rc);
            if (launch_to_f32(nnz, a->sval, a->sval32, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                return bail(fail(HBS_ERR_HIP, "FP32 value copy failed"));
        }
    }
    SpmvPlan &p = a->plan;
    p.nsys = nsys;
    a->stream_nt = g_stream_nt;
    p.block_nnz = block_nnz;
    p.npt = npt;
    p.nblk = (int)hp.blocks.size();
    p.ngroups = (int)hp.groups.size();
    p.nchunks = (int)hp.chunks.size();
    p.serial_rows = hp.serial_rows;
    p.max_row = hp.max_row;
    if ((rc = dalloc(&p.blocks, hp.blocks.size())) || (rc = dalloc(&p.aux, hp.aux.size())) ||
        (rc = dalloc(&p.chunks, hp.chunks.size())) || (rc = dalloc(&p.groups, hp.groups.size())) ||
        (rc = dalloc(&a->sys, (size_t)nsys)) || (rc = dalloc(&a->chunk_sums, hp.chunks.size())) ||
        (rc = dalloc(&a->part, (size_t)std::max(p.ngroups, 1) * kMaxDots)))
        return bail(rc);
    if ((rc = upload(p.blocks, hp.blocks)) || (rc = upload(p.aux, hp.aux)) || (rc = upload(p.chunks, hp.chunks)) ||
        (rc = upload(p.groups, hp.groups)) || (rc = upload(a->sys, ranges)))
        return bail(rc);
    *out = a;
    return HBS_OK;
}

}  // namespace

namespace {

int csr_create(hbs_ctx ctx, int m, int n, int nnz, const int *rowptr, const int *col, const double *val, int flags,
               int nsys, const int *sys_rows, hbs_csr *out)
{
    if (!ctx || !out) return fail(HBS_ERR_ARG, "null argument");
    *out = nullptr;
    CHK(validate_csr(m, n, nnz, rowptr, col, val));
    DevGuard g(ctx->device);
    hbs_csr a = nullptr;
    CHK(csr_upload(ctx, m, n, nnz, rowptr, col, val, flags, nsys, sys_rows, &a));
    if (flags & HBS_CSR_TRANSPOSE) {
        // A^T as CSR: counting sort by column, rows visited in ascending order
        std::vector<int> tptr((size_t)n + 1, 0), tcol((size_t)nnz);
        std::vector<double> tval((size_t)nnz);
        for (int j = 0; j < nnz; ++j) tptr[(size_t)col[j] + 1]++;
        for (int c = 0; c < n; ++c) tptr[(size_t)c + 1] += tptr[(size_t)c];
        std::vector<int> fill(tptr.begin(), tptr.end() - 1), perm((size_t)nnz);
        for (int r = 0; r < m; ++r)
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const int d = fill[(size_t)col[j]]++;
                tcol[(size_t)d] = r;
                tval[(size_t)d] = val[j];
                perm[(size_t)j] = d;
            }
        int rc = csr_upload(ctx, n, m, nnz, tptr.data(), tcol.data(), tval.data(), 0, nsys, sys_rows, &a->tr);
        if (rc == HBS_OK) rc = dalloc(&a->tr_perm, (size_t)nnz);
        if (rc == HBS_OK) rc = upload(a->tr_perm, perm);
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

}  // namespace

extern "C" int hbs_csr_create(hbs_ctx ctx, int m, int n, int nnz, const int *rowptr, const int *col,
                              const double *val, int flags, hbs_csr *out)
{
    return csr_create(ctx, m, n, nnz, rowptr, col, val, flags, 0, nullptr, out);
}

extern "C" int hbs_csr_create_systems(hbs_ctx ctx, int nsys, const int *sys_rows, int nnz, const int *rowptr,
                                      const int *col, const double *val, int flags, hbs_csr *out)
{
    if (!out) return fail(HBS_ERR_ARG, "null argument");
    *out = nullptr;
    if (nsys < 1 || !sys_rows) return fail(HBS_ERR_ARG, "need nsys >= 1 and sys_rows[nsys + 1]");
    if (sys_rows[0] != 0) return fail(HBS_ERR_ARG, "sys_rows[0] must be 0");
    for (int q = 0; q < nsys; ++q)
        if (sys_rows[q + 1] <= sys_rows[q]) return fail(HBS_ERR_ARG, "system %d has no rows", q);
    const int m = sys_rows[nsys];
    CHK(validate_csr(m, m, nnz, rowptr, col, val));
    for (int q = 0; q < nsys; ++q)
        for (int r = sys_rows[q]; r < sys_rows[q + 1]; ++r)
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j)
                if (col[j] < sys_rows[q] || col[j] >= sys_rows[q + 1])
                    return fail(HBS_ERR_ARG, "entry %d of row %d leaves system %d (column %d)", j, r, q, col[j]);
    return csr_create(ctx, m, m, nnz, rowptr, col, val, flags, nsys, sys_rows, out);
}

extern "C" int hbs_csr_set_values(hbs_csr a, const double *val_dev)
{
    if (!a || (a->nnz > 0 && !val_dev)) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(a->ctx->device);
    hipStream_t q = a->ctx->stream;
    if (a->nnz == 0) return HBS_OK;
    HIPCHK(hipMemcpyAsync(a->val, val_dev, sizeof(double) * (size_t)a->nnz, hipMemcpyDeviceToDevice, q));
    if (a->tr) HIPCHK(launch_scatter_f64(a->nnz, a->val, a->tr_perm, a->tr->val, q));
    // the copies derived from val: FP32, column-sorted, of A and of A^T
    for (hbs_csr t : {a, a->tr}) {
        if (!t) continue;
        if (t->val32) HIPCHK(launch_to_f32(t->nnz, t->val, t->val32, q));
        if (t->sval) HIPCHK(launch_gather_f64(t->nnz, t->val, t->ssrc, t->sval, q));
        if (t->sval32) HIPCHK(launch_to_f32(t->nnz, t->sval, t->sval32, q));
        if (t->xs_val) HIPCHK(launch_gather_f64(t->nnz, t->val, t->xs_src, t->xs_val, q));
        if (t->xs_val32) HIPCHK(launch_to_f32(t->nnz, t->xs_val, t->xs_val32, q));
        if (t->bval) HIPCHK(launch_gather_f64(t->nnz, t->val, t->bsrc, t->bval, q));
        if (t->bval32) HIPCHK(launch_to_f32(t->nnz, t->bval, t->bval32, q));
    }
    if (hbs_csr od = a->od) {
        // the FP32 off-diagonal part and the diagonal (whose new values need
        // not be 1: from here on the row sums read them)
        HIPCHK(launch_gather_f64(od->nnz, a->val, a->od_src, od->val, q));
        HIPCHK(launch_to_f32(od->nnz, od->val, od->val32, q));
        HIPCHK(launch_gather_f64(od->nnz, od->val, od->xs_src, od->xs_val, q));
        HIPCHK(launch_to_f32(od->nnz, od->xs_val, od->xs_val32, q));
        if (!a->diag32) {
            CHK(dalloc(&a->diag32, (size_t)a->m));
            ++a->diag_gen;   // graphs captured with the unit-diagonal row sums are stale
        }
        HIPCHK(launch_gather_to_f32(a->m, a->val, a->diag_src, a->diag32, q));
    }
    return HBS_OK;
}

extern "C" int hbs_csr_systems(hbs_csr a, int *nsys)
{
    if (!a || !nsys) return fail(HBS_ERR_ARG, "null argument");
    *nsys = a->nsys;
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

extern "C" int hbs_csr_plan_info(hbs_csr a, int *ngroups, int *nchunks, int *block_nnz, int *max_row)
{
    if (!a) return fail(HBS_ERR_ARG, "null csr");
    if (ngroups) *ngroups = a->plan.ngroups;
    if (nchunks) *nchunks = a->plan.nchunks;
    if (block_nnz) *block_nnz = a->plan.block_nnz;
    if (max_row) *max_row = a->plan.max_row;
    return HBS_OK;
}

extern "C" int hbs_csr_xseg(hbs_csr a, int *nseg)
{
    if (!a || !nseg) return fail(HBS_ERR_ARG, "null argument");
    *nseg = a->xs_nseg;
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
    SpmvArgs<double> s = spmv_args<double>(a, a->val);
    s.x = x;
    s.y = y;
    s.alpha = alpha;
    s.beta = beta;
    HIPCHK(run_spmv(s, a, EPI_AXPBY, nullptr, a->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_csr_spmv_f32(hbs_csr a, float alpha, const float *x, float beta, float *y)
{
    if (!a || (a->m > 0 && (!x || !y))) return fail(HBS_ERR_ARG, "bad argument");
    if (!a->val32) return fail(HBS_ERR_ARG, "matrix was created without HBS_CSR_F32");
    DevGuard g(a->ctx->device);
    SpmvArgs<float> s = spmv_args<float>(a, a->val32);
    s.x = x;
    s.y = y;
    s.alpha = alpha;
    s.beta = beta;
    HIPCHK(run_spmv(s, a, EPI_AXPBY, nullptr, a->ctx->stream));
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
    hipError_t e;
    if constexpr (sizeof(T) == 8)
        e = launch_spmm_f64(a->m, a->rowptr, a->col, a->val, k, X, Y, tile, a->row_runs(), a->ctx->stream);
    else
        e = launch_spmm_f32(a->m, a->rowptr, a->col, a->val32, k, X, Y, tile, a->row_runs(), a->ctx->stream);
    return e == hipSuccess ? HBS_OK : fail(HBS_ERR_HIP, "spmm launch failed: %s", hipGetErrorString(e));
}

// Own autotuner: time each admissible RHS tile width (hipEvents, a few
// launches each) and keep the fastest per (k, dtype).
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

template <typename T>
int bench_spmv(hbs_csr a, const T *x, T *y, int reps, double *avg_ms)
{
    if (!a || !avg_ms || reps <= 0) return fail(HBS_ERR_ARG, "bad argument");
    if (sizeof(T) == 4 && !a->val32) return fail(HBS_ERR_ARG, "matrix was created without HBS_CSR_F32");
    DevGuard g(a->ctx->device);
    CHK(ensure_events(a, 2));
    SpmvArgs<T> s = spmv_args<T>(a, sizeof(T) == 8 ? (const T *)(const void *)a->val : (const T *)(const void *)a->val32);
    s.x = x;
    s.y = y;
    HIPCHK(run_spmv(s, a, EPI_AXPBY, nullptr, a->ctx->stream));   // warm-up, as spmv_gpu.c:116-118
    // one event pair around all reps back to back: the average launch
    // duration in steady state, without an event packet per launch
    HIPCHK(hipEventRecord(a->evs[0], a->ctx->stream));
    for (int r = 0; r < reps; ++r) HIPCHK(run_spmv(s, a, EPI_AXPBY, nullptr, a->ctx->stream));
    HIPCHK(hipEventRecord(a->evs[1], a->ctx->stream));
    HIPCHK(hipEventSynchronize(a->evs[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a->evs[0], a->evs[1]));
    *avg_ms = (double)ms / reps;
    return HBS_OK;
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
    return bench_spmv<double>(a, x, y, reps, avg_ms);
}

extern "C" int hbs_bench_spmv_f32(hbs_csr a, const float *x, float *y, int reps, double *avg_ms)
{
    return bench_spmv<float>(a, x, y, reps, avg_ms);
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
