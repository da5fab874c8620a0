// hbs_objects.h -- the library's handle types and host-side helpers shared by
// hbs_api.cpp, hbs_solver.cpp and hbs_shard.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "hbs_err.h"
#include "hbs_internal.h"

namespace hbs {

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return ::hbs::set_error(HBS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                    __FILE__, __LINE__);                                                \
    } while (0)

#define CHK(expr)                      \
    do {                               \
        int rc_ = (expr);              \
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
        return set_error(HBS_ERR_ALLOC, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    return HBS_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <typename T>
int upload(T *dev, const std::vector<T> &host)
{
    if (host.empty()) return HBS_OK;
    HIPCHK(hipMemcpy(dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
    return HBS_OK;
}

struct DevGuard {
    explicit DevGuard(int d) { (void)hipSetDevice(d); }
};

// Vector workgroups of n elements (one system) on the device.
struct DevVecPlan {
    int n = -1;
    int epb = -1;               // elements per vector workgroup
    int nvb = 0;
    VecBlock *blocks = nullptr;
    SysRange *sys = nullptr;    // {0, 0, 0, nvb}
    double *part = nullptr;     // [nvb * kMaxDots]
    int build(int n);
    void release();
};

}  // namespace hbs

struct hbs_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hbs::SolverState *st = nullptr;   // device, BLAS-1 results
    hbs::SolverState *hst = nullptr;  // pinned host
    hbs::DevVecPlan vp;               // cached plan for BLAS-1 dots
};

struct hbs_csr_s {
    hbs_ctx ctx = nullptr;
    int m = 0, n = 0, nnz = 0, flags = 0;
    int *rowptr = nullptr, *col = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;
    int *scol = nullptr;              // column-sorted copy of kind-2 blocks (else CSR order)
    double *sval = nullptr;
    float *sval32 = nullptr;
    unsigned short *sperm = nullptr;  // position in the block of each sorted entry (pos_shift 0)
    int pos_shift = 0;                // > 0: scol = col | position << pos_shift, no sperm
    int *ssrc = nullptr;              // CSR entry of each sorted position (refreshes sval)
    // 3x3-block matrices (detect_b3): values block-major and each block's
    // first column, read by the kind-3 row blocks instead of col / val
    int *bcol = nullptr;
    double *bval = nullptr;
    float *bval32 = nullptr;
    int *bsrc = nullptr;              // CSR entry of each block-major value (refreshes bval)
    int b3var = 0;                    // some row blocks hold block rows of different lengths
    // x-segmented SpMV (columns scattered over all of x): k_xseg forms the
    // products per x segment, k_spmv<PRE> sums them (layout in XsegArgs)
    int xs_nseg = 0;                  // 0: CSR-stream SpMV
    int xs_split = 1;                 // k_xseg workgroups per segment
    int xs_max_chunks = 0, xs_max_granules = 0;
    int *xs_sbase = nullptr, *xs_cbase = nullptr, *xs_gbase = nullptr, *xs_gch = nullptr, *xs_off = nullptr;
    int *xs_src = nullptr;            // CSR entry of each segment-major entry (refreshes xs_val)
    unsigned short *xs_col = nullptr, *xs_pos = nullptr;
    unsigned short *rperm = nullptr;   // serial row-sum order of x-segmented matrices (plan_row_order)
    unsigned char *xs_delta = nullptr;
    double *xs_val = nullptr, *xs_prod = nullptr;
    float *xs_val32 = nullptr, *xs_prod32 = nullptr;
    // FP32 SpMVs of a square x-segmented matrix: the off-diagonal entries as
    // an x-segmented matrix of their own (od), the diagonal added per row
    // (d_i x_i; x_i alone while every diagonal entry is 1, diag32 null)
    hbs_csr_s *od = nullptr;
    float *diag32 = nullptr;
    int *od_src = nullptr;            // [od->nnz] CSR entry of each off-diagonal entry (refreshes od->val)
    int *diag_src = nullptr;          // [m] CSR entry of each row's diagonal, nnz (a zero) for none
    int diag_gen = 0;                 // counts changes of the diagonal's form (unit -> values)
    hbs::SpmvPlan plan;
    hbs::SysRange *sys = nullptr;     // [nsys] {g0, g1, 0, 0}
    double *chunk_sums = nullptr;     // [nchunks]
    double *part = nullptr;           // [ngroups * kMaxDots]
    hbs_csr_s *tr = nullptr;
    int *tr_perm = nullptr;           // [nnz] position of entry j of A in A^T
    std::map<std::pair<int, int>, int> spmm_tile;   // (k, dtype) -> tile
    // multi-column SpMM traversal: runs of rr_len consecutive rows (a 3-D
    // grid's x-lines in y-slab order, plan_grid_runs); null: rows in order
    int *rr_run = nullptr;
    int rr_len = 0, rr_nruns = 0;
    // step counters of the 64-column SpMM sweeps (zero between launches):
    // kSpmmCtrTiles column tiles x 8 XCDs x kCtrSet, kCtrStride ints apart.
    // The multi-column SpMMs of one matrix run one at a time (stream order).
    int *spmm_ctr = nullptr;
    hbs::RowRuns row_runs() const
    {
        return hbs::RowRuns{rr_run, rr_len, rr_nruns, spmm_ctr, spmm_ctr ? hbs::kSpmmCtrTiles : 0,
                            hbs::g_spmm_counters};
    }
    int stream_nt = 0;                // non-temporal col/val loads
    int nsys = 1;                     // independent diagonal blocks (systems)
    std::vector<int> sys_rows;        // [nsys + 1] row offsets of the systems
    std::vector<hbs::SysRange> sys_host;   // host copy of sys
    hipEvent_t *evs = nullptr;
    int nevs = 0;
};

namespace hbs {
constexpr int kSlotA = 0, kSlotB = 1, kSlotOuter = 2, kNumSlots = 4;
constexpr int kMirrorSlots = 5;   // host mirror: two chunk polls, then states read after a solve or sweep
struct BlockWork;                    // k right-hand-side workspace (hbs_block.cpp)
void block_work_free(BlockWork *w);
}  // namespace hbs

struct hbs_solver_s {
    hbs_csr A = nullptr;
    int n = 0;
    int nsys = 1;                 // independent systems of A (diagonal blocks)
    double *X = nullptr, *R = nullptr, *R0 = nullptr, *P = nullptr, *V = nullptr, *S = nullptr, *T = nullptr;
    double *Xb = nullptr;              // mixed: the second FP64 iterate (sweeps alternate X and Xb)
    float *X32 = nullptr, *R32 = nullptr, *R032 = nullptr, *P32 = nullptr, *V32 = nullptr, *S32 = nullptr,
          *T32 = nullptr;
    hbs::SolverState *st = nullptr;    // device [kNumSlots][nsys]
    hbs::SolverState *hst = nullptr;   // pinned host [kMirrorSlots][nsys]: poll buffers 0, 1; read buffers 2..4
    hbs::SolverState *hst_dev = nullptr;   // device view of hst (host-mapped, coherent)
    hbs::VecBlock *vblocks = nullptr;
    hbs::SysRange *sys = nullptr;      // [nsys] {g0, g1, v0, v1}
    int nvb = 0;
    // FP32 sweeps (one system): vector workgroups of twice the elements, so
    // each moves as many bytes as an FP64 one and the fixed per-workgroup
    // work (partial re-sums, state) is spread over as many bytes
    hbs::VecBlock *vblocks32 = nullptr;
    hbs::SysRange *sys32 = nullptr;    // {g0, g1, 0, nvb32}
    int nvb32 = 0, epb32 = 0;
    // partials, one buffer per producer so a consumer never sees them overwritten
    double *partA = nullptr;      // SpMV1: (r0, v)
    double *partB = nullptr;      // k_vec_S: ||s||^2
    double *partC = nullptr;      // SpMV2: (t, s), (t, t)
    double *partD = nullptr;      // k_vec_XR / start: (r, r0), ||r||^2
    double *partE = nullptr;      // residual SpMV: ||b - Ax||^2
    double *partF = nullptr;      // dots: (b, b)
    double *chunk_sums = nullptr; // long-row chunk sums
    static constexpr int kGraphLens = 33;             // chunks of 1..32 iterations
    hipGraphExec_t graph[2][kGraphLens][2] = {};       // [FP64, FP32][chunk length][poll slot]
    // x-segmented iterations with the P and S updates in the product passes:
    // per precision, the SpMV's workgroups and the product pass's partials
    // ({g0, g1, 0, nseg * split}) for the ||s|| test; the partials buffer
    hbs::SysRange *sysx[2] = {nullptr, nullptr};
    double *partX = nullptr;
    int graph_diag_gen = 0;                            // A->diag_gen the FP32 graphs were captured at
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int seq = 0;                  // chunks launched without graphs: the mirror's completion stamp
    hbs_solver_s *adj = nullptr;
    hbs::BlockWork *blk = nullptr;
    hbs::SolverState *slot(int k) const { return st + (size_t)k * nsys; }
};

namespace hbs {

// Host wait for an event, bounded by the solver wait limit
// (hbs_set_wait_limit): HBS_ERR_HIP when it passes or the stream fails.
int wait_event_limited(hipEvent_t ev);

// SpMV launch arguments of matrix a with values val (FP64 or FP32 copy).
// The matrix whose streams T's SpMVs read: FP32 SpMVs of a matrix with
// its diagonal apart read the off-diagonal part.
template <typename T>
hbs_csr spmv_mat(hbs_csr a)
{
    return sizeof(T) == 4 && a->od ? a->od : a;
}

// SpMV launch arguments of matrix a with values val (FP64 or FP32 copy; an
// FP32 SpMV of a matrix with its diagonal apart reads a->od's FP32 values
// and the diagonal).
template <typename T>
SpmvArgs<T> spmv_args(hbs_csr a, const T *val)
{
    if constexpr (sizeof(T) == 4) {
        if (a->od) {
            SpmvArgs<T> s = spmv_args<T>(a->od, a->od->val32);
            s.dmode = a->diag32 ? 2 : 1;
            s.diag = a->diag32;
            return s;
        }
    }
    SpmvArgs<T> s{};
    s.blocks = a->plan.blocks;
    s.aux = a->plan.aux;
    s.groups = a->plan.groups;
    s.ngroups = a->plan.ngroups;
    s.sys = a->sys;
    s.rowptr = a->rowptr;
    s.col = a->col;
    s.val = val;
    s.scol = a->scol;
    if constexpr (sizeof(T) == 8) s.sval = reinterpret_cast<const T *>(a->sval);
    else s.sval = reinterpret_cast<const T *>(a->sval32);
    s.sperm = a->sperm;
    s.pos_shift = a->pos_shift;
    s.chunk_sums = a->chunk_sums;
    if (a->xs_nseg > 0) {
        if constexpr (sizeof(T) == 8) s.prod = reinterpret_cast<const T *>(a->xs_prod);
        else s.prod = reinterpret_cast<const T *>(a->xs_prod32);
        s.ppos = a->xs_pos;
        s.rperm = a->rperm;
    }
    s.bcol = a->bcol;
    if constexpr (sizeof(T) == 8) s.bval = reinterpret_cast<const T *>(a->bval);
    else s.bval = reinterpret_cast<const T *>(a->bval32);
    s.b3var = a->b3var;
    s.part_out = a->part;
    s.stream_nt = a->stream_nt;
    s.alpha = T(1);
    s.beta = T(0);
    s.fin_in = FIN_NONE;
    return s;
}

// Queue one SpMV: for x-segmented matrices the product pass, then the
// long-row chunk sums (if any), then the row blocks.
template <typename T>
hipError_t run_spmv(const SpmvArgs<T> &s, const hbs_csr a_in, int epi, const SolverState *chunk_st, hipStream_t q,
                    const XsegArgs<T> *fuse = nullptr)
{
    const hbs_csr a = spmv_mat<T>(a_in);
    if (a->xs_nseg > 0) {
        XsegArgs<T> x{};
        x.nseg = a->xs_nseg;
        x.split = a->xs_split;
        x.seg_cols = kSegCols;
        x.n = a->n;
        x.max_chunks = a->xs_max_chunks;
        x.max_granules = a->xs_max_granules;
        x.sbase = a->xs_sbase;
        x.cbase = a->xs_cbase;
        x.gbase = a->xs_gbase;
        x.gch = a->xs_gch;
        x.off = a->xs_off;
        x.xcol = a->xs_col;
        x.delta = a->xs_delta;
        if constexpr (sizeof(T) == 8) x.xval = reinterpret_cast<const T *>(a->xs_val);
        else x.xval = reinterpret_cast<const T *>(a->xs_val32);
        x.x = s.x;
        x.prod = const_cast<T *>(s.prod);
        x.st = chunk_st;
        if (fuse) {
            // the product pass updates its x (a solver iteration's P or S)
            x.fuse = fuse->fuse;
            x.xv_out = fuse->xv_out;
            x.R = fuse->R;
            x.V = fuse->V;
            x.st_in = fuse->st_in;
            x.st_out = fuse->st_out;
            x.part_in = fuse->part_in;
            x.nparts_in = fuse->nparts_in;
            x.pin0 = fuse->pin0;
            x.pin1 = fuse->pin1;
            x.part_out = fuse->part_out;
        }
        if (!x.xval || !x.prod) return hipErrorInvalidValue;
        hipError_t e = launch_xseg(x, q);
        if (e != hipSuccess) return e;
    }
    if (a->plan.nchunks > 0) {
        ChunkArgs<T> c{};
        c.chunks = a->plan.chunks;
        c.nchunks = a->plan.nchunks;
        c.col = s.col;
        c.val = s.val;
        c.x = s.x;
        c.chunk_sums = const_cast<double *>(s.chunk_sums);
        c.st = chunk_st;
        c.prod = s.prod;
        hipError_t e = launch_chunks(c, a->plan.block_nnz, q);
        if (e != hipSuccess) return e;
    }
    return launch_spmv(s, epi, a->plan.block_nnz, a->plan.npt, q);
}

}  // namespace hbs
