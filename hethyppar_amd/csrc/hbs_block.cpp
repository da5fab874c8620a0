// hbs_block.cpp -- BiCGStab on k right-hand sides (hbs_bicgstab_block_f64):
// one independent BiCGStab per column, columns in groups of up to 64 that
// share every SpMM pass over A.  Per iteration of a group: k_bvec P, SpMM
// with (r0, v), alpha, k_bvec S with ||s||^2, the ||s|| test, SpMM with
// (t, s), (t, t), omega, k_bvec XR with (r, r0), ||r||^2, the ||r|| test --
// the scalar steps each one workgroup per column.  Iteration chunks are
// captured into hipGraphs; the host polls the column states once per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "hbs_objects.h"

using namespace hbs;

#define fail ::hbs::set_error

namespace hbs {

struct BlockWork {
    int n = 0, k = 0;
    double *X = nullptr, *R = nullptr, *R0 = nullptr, *P = nullptr, *V = nullptr, *S = nullptr, *T = nullptr;
    SolverState *st = nullptr;    // device [kBlockCols]
    SolverState *hst = nullptr;   // pinned host [3][kBlockCols]
    double *part = nullptr;       // [kMaxDots][kBlockCols][kMaxBlockGroups], consumed before the next producer
    int ngroups = 0;              // a multiple of 8: rows are dealt per XCD
    int sgroups = 0;              // the SpMMs' workgroups (a multiple of 8, at most ngroups; block_tune)
    int tile = 0, rb = 0;         // columns per group, SpMM rows per lane step (block_tune)
    int v16 = 0;                  // 64-column groups gathered in 16-byte pieces (block_tune)
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;   // (first column, columns, chunk)
    hipEvent_t ev[2] = {nullptr, nullptr};
};

void block_work_free(BlockWork *w)
{
    if (!w) return;
    for (double **p : {&w->X, &w->R, &w->R0, &w->P, &w->V, &w->S, &w->T, &w->part}) dfree(*p);
    dfree(w->st);
    if (w->hst) (void)hipHostFree(w->hst);
    for (auto &g : w->graphs) (void)hipGraphExecDestroy(g.second);
    for (hipEvent_t e : w->ev)
        if (e) (void)hipEventDestroy(e);
    delete w;
}

}  // namespace hbs

namespace {

int pow2_ceil(int v)
{
    int p = 1;
    while (p < v) p *= 2;
    return p;
}

constexpr int kMaxBlockGroups = 2048;
int g_block_tile = 0, g_block_rb = 0;   // forced block-solve tuning (0: tuned)
int g_block_sgroups = 0;                 // forced SpMM workgroups of the block solve (0: tuned)

// Workgroups of the streaming and SpMM kernels for groups of `tile`
// columns: up to kMaxBlockGroups, a multiple of the 8 XCDs; a pass of 4
// waves covers 4 * (64 / kp) rows.
int block_groups(int n, int tile)
{
    const int kp = pow2_ceil(std::min(tile, kBlockCols));
    const long long pass = (long long)kWaves * (64 / kp);
    const long long want = std::min<long long>(kMaxBlockGroups, (n + pass - 1) / pass);
    return (int)std::max<long long>(8, (want + 7) / 8 * 8);
}

// Whether the SpMMs of groups of `tile` columns (k in all) gather X in
// 16-byte pieces, two columns per lane (k_bspmm2): 64-column groups, k
// even (so every group's first column and width, and the row length, keep
// the column pairs 16-byte aligned), wide loads on (hbs_set_spmm_tuning)
// and no forced rows per lane step (those are the one-column-per-lane
// form's).
bool block_v16(int k, int tile)
{
    return g_spmm_v16 && g_block_rb == 0 && pow2_ceil(std::min(tile, kBlockCols)) == 64 && k % 2 == 0;
}

int block_work(hbs_solver s, int k, BlockWork **out)
{
    BlockWork *w = s->blk;
    if (w && w->k == k) {
        *out = w;
        return HBS_OK;
    }
    if (w) {
        block_work_free(w);
        s->blk = nullptr;
    }
    w = new BlockWork();
    w->n = s->n;
    w->k = k;
    w->ngroups = block_groups(w->n, k);
    w->sgroups = w->ngroups;
    const size_t nk = (size_t)w->n * k;
    int rc = HBS_OK;
    for (double **p : {&w->X, &w->R, &w->R0, &w->P, &w->V, &w->S, &w->T})
        if (rc == HBS_OK) rc = dalloc(p, nk);
    if (rc == HBS_OK) rc = dalloc(&w->part, (size_t)kMaxDots * kBlockCols * kMaxBlockGroups);
    if (rc == HBS_OK) rc = dalloc(&w->st, kBlockCols);
    if (rc == HBS_OK && hipHostMalloc((void **)&w->hst, 3 * sizeof(SolverState) * kBlockCols) != hipSuccess)
        rc = fail(HBS_ERR_ALLOC, "pinned host allocation failed");
    for (hipEvent_t &e : w->ev)
        if (rc == HBS_OK && hipEventCreate(&e) != hipSuccess) rc = fail(HBS_ERR_HIP, "event creation failed");
    if (rc != HBS_OK) {
        block_work_free(w);
        return rc;
    }
    s->blk = w;
    *out = w;
    return HBS_OK;
}

BlockArgs group_args(hbs_solver s, BlockWork *w, int c0, int kc)
{
    BlockArgs a{};
    a.n = w->n;
    a.ld = w->k;
    a.kc = kc;
    a.kp = pow2_ceil(kc);
    a.ngroups = w->ngroups;
    a.rowptr = s->A->rowptr;
    a.col = s->A->col;
    a.val = s->A->val;
    a.X = w->X + c0;
    a.R = w->R + c0;
    a.R0 = w->R0 + c0;
    a.P = w->P + c0;
    a.V = w->V + c0;
    a.S = w->S + c0;
    a.T = w->T + c0;
    a.st = w->st;
    a.skip_done = 1;
    a.part = w->part;
    a.rr = s->A->row_runs();
    a.rb = w->rb > 0 ? w->rb : 2;
    a.v16 = w->v16 && a.kp == 64 && kc % 2 == 0 && c0 % 2 == 0 && a.ld % 2 == 0;
    return a;
}

// The block solve's autotuner: for k right-hand sides, the column-group
// width (the SpMM's RHS tile: 64, 32, 16 ... columns, at most k rounded up
// to a power of two) and the SpMM rows per lane step (1, 2, 4; with the
// tile's 64 / tile rows per wave, the rows a wavefront works on at once;
// 64-column tiles in the 16-byte form always take two rows per step).
// Each candidate times the (r0, v) SpMM of one group, best of 3 after a
// warm-up, on the workspace's own blocks; the score is that time times the
// groups the k columns need.  Forced choices (hbs_set_block_tuning) skip the
// timing.  Once per workspace (k).
int block_tune(hbs_solver s, BlockWork *w, hipStream_t q)
{
    const int kmax = pow2_ceil(std::min(w->k, kBlockCols));
    std::vector<int> tiles, rbs;
    if (g_block_tile > 0) tiles.push_back(std::min(g_block_tile, kmax));
    else
        for (int t = kmax; t >= std::max(1, kmax / 4); t /= 2) tiles.push_back(t);
    if (g_block_rb > 0) rbs.push_back(g_block_rb);
    else rbs = {1, 2, 4};
    // the grid's runs hold whole steps of the 64-column sweep
    const RowRuns rr = s->A->row_runs();
    if (tiles.size() == 1 && rbs.size() == 1) {
        w->tile = tiles[0];
        w->rb = rr.run && rr.len % rbs[0] != 0 ? 2 : rbs[0];
        w->v16 = block_v16(w->k, w->tile);
        w->ngroups = block_groups(w->n, w->tile);
        w->sgroups = g_block_sgroups > 0 ? std::min(g_block_sgroups, w->ngroups) : w->ngroups;
        return HBS_OK;
    }
    const size_t nk = (size_t)w->n * w->k;
    HIPCHK(launch_fill_zero(w->P, sizeof(double) * nk, q));
    HIPCHK(launch_fill_zero(w->R0, sizeof(double) * nk, q));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(HBS_ERR_HIP, "event creation failed");
    }
    double best = -1.0;
    int bt = tiles[0], brb = 2, bsg = block_groups(w->n, tiles[0]), rc = HBS_OK;
    for (int t : tiles) {
        // SpMM workgroups: the streaming kernels' count, a half and a
        // quarter of it (fewer rows in flight per XCD keep more of the X
        // rows the sweep re-reads in its L2)
        const int g0 = block_groups(w->n, t);
        std::vector<int> sgs;
        if (g_block_sgroups > 0) sgs.push_back(std::min(g_block_sgroups, g0));
        else
            for (int d = 1; d <= 4; d *= 2)
                if (g0 / d >= 8 && (d == 1 || g0 / d >= 256) && (g0 / d) % 8 == 0) sgs.push_back(g0 / d);
        const bool v16 = block_v16(w->k, t);
        for (int sg : sgs)
        for (int rb : v16 ? std::vector<int>{2} : rbs) {
            if (rr.run && rr.len % rb != 0) continue;
            BlockArgs a = group_args(s, w, 0, std::min(t, w->k));
            a.ngroups = sg;
            a.rb = rb;
            a.v16 = v16 && a.kp == 64;
            a.skip_done = 0;
            a.x = w->P;
            a.y = w->V;
            a.w = w->R0;
            float ms = 1e30f;
            hipError_t e = launch_block_spmm(a, BEPI_DOT1, q);
            for (int r = 0; r < 3 && e == hipSuccess; ++r) {
                float m = 0.f;
                if ((e = hipEventRecord(e0, q)) != hipSuccess || (e = launch_block_spmm(a, BEPI_DOT1, q)) != hipSuccess ||
                    (e = hipEventRecord(e1, q)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
                    (e = hipEventElapsedTime(&m, e0, e1)) != hipSuccess)
                    break;
                ms = std::min(ms, m);
            }
            if (e != hipSuccess) {
                rc = fail(HBS_ERR_HIP, "block SpMM tuning launch failed: %s", hipGetErrorString(e));
                break;
            }
            const double score = (double)ms * ((w->k + t - 1) / t);
            if (best < 0.0 || score < best) {
                best = score;
                bt = t;
                brb = rb;
                bsg = sg;
            }
        }
        if (rc != HBS_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != HBS_OK) return rc;
    w->tile = bt;
    w->rb = brb;
    w->v16 = block_v16(w->k, bt);
    w->ngroups = block_groups(w->n, bt);
    w->sgroups = bsg;
    return HBS_OK;
}

hipError_t block_iteration(const BlockArgs &g, BlockWork *w, hipStream_t q)
{
    hipError_t e;
    BlockArgs a = g;
    a.ngroups = w->sgroups;   // the SpMMs' workgroups, and the partials they write
    if ((e = launch_block_vec(g, BVEC_P, q)) != hipSuccess) return e;
    a.x = g.P;
    a.y = g.V;
    a.w = g.R0;
    if ((e = launch_block_spmm(a, BEPI_DOT1, q)) != hipSuccess) return e;
    if ((e = launch_block_fin(w->st, g.kc, w->part, w->sgroups, FIN_ALPHA, q)) != hipSuccess) return e;
    if ((e = launch_block_vec(g, BVEC_S, q)) != hipSuccess) return e;
    if ((e = launch_block_fin(w->st, g.kc, w->part, w->ngroups, FIN_S, q)) != hipSuccess) return e;
    a.x = g.S;
    a.y = g.T;
    a.w = g.S;
    if ((e = launch_block_spmm(a, BEPI_DOT2, q)) != hipSuccess) return e;
    if ((e = launch_block_fin(w->st, g.kc, w->part, w->sgroups, FIN_OMEGA, q)) != hipSuccess) return e;
    if ((e = launch_block_vec(g, BVEC_XR, q)) != hipSuccess) return e;
    return launch_block_fin(w->st, g.kc, w->part, w->ngroups, FIN_XR, q);
}

int block_chunk(const BlockArgs &g, BlockWork *w, int c0, int chunk, bool use_graph, hipStream_t q)
{
    if (!use_graph) {
        for (int i = 0; i < chunk; ++i) HIPCHK(block_iteration(g, w, q));
        return HBS_OK;
    }
    const auto key = std::make_tuple(c0, g.kc, chunk);
    auto it = w->graphs.find(key);
    if (it == w->graphs.end()) {
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        for (int i = 0; i < chunk && le == hipSuccess; ++i) le = block_iteration(g, w, q);
        hipError_t ce = hipStreamEndCapture(q, &gr);
        if (le != hipSuccess) return fail(HBS_ERR_HIP, "launch during capture failed: %s", hipGetErrorString(le));
        if (ce != hipSuccess) return fail(HBS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
        hipGraphExec_t ex = nullptr;
        hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ie != hipSuccess) return fail(HBS_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
        it = w->graphs.emplace(key, ex).first;
    }
    HIPCHK(hipGraphLaunch(it->second, q));
    return HBS_OK;
}

// Chunks of iterations until every column of the group is done; the states
// of chunk c are read after chunk c + 1 is queued.
int block_iterate(const BlockArgs &g, BlockWork *w, int c0, int maxit, int check_every, bool use_graph,
                  hipStream_t q)
{
    const int k = check_every > 0 ? check_every : 4;
    const size_t bytes = sizeof(SolverState) * (size_t)g.kc;
    long issued = 0;
    int c = 0, pending = -1;
    for (;;) {
        CHK(block_chunk(g, w, c0, k, use_graph, q));
        issued += k;
        HIPCHK(hipMemcpyAsync(w->hst + (size_t)(c & 1) * kBlockCols, w->st, bytes, hipMemcpyDeviceToHost, q));
        HIPCHK(hipEventRecord(w->ev[c & 1], q));
        if (pending >= 0) {
            CHK(wait_event_limited(w->ev[pending & 1]));
            const SolverState *h = w->hst + (size_t)(pending & 1) * kBlockCols;
            bool all = true;
            for (int i = 0; i < g.kc && all; ++i) all = h[i].done != 0;
            if (all) break;
        }
        pending = c++;
        if (issued >= maxit) {
            CHK(wait_event_limited(w->ev[pending & 1]));
            break;
        }
    }
    return HBS_OK;
}

}  // namespace

extern "C" int hbs_set_block_tuning(int tile, int rows_per_step)
{
    if (tile < 0 || tile > kBlockCols || (tile & (tile - 1)) != 0 ||
        (rows_per_step != 0 && rows_per_step != 1 && rows_per_step != 2 && rows_per_step != 4))
        return fail(HBS_ERR_ARG, "tile 0 or a power of two <= %d; rows_per_step 0, 1, 2 or 4", kBlockCols);
    g_block_tile = tile;
    g_block_rb = rows_per_step;
    return HBS_OK;
}

extern "C" int hbs_set_block_spmm_groups(int groups)
{
    if (groups < 0 || groups % 8 != 0) return fail(HBS_ERR_ARG, "groups: 0 (tuned) or a positive multiple of 8");
    g_block_sgroups = groups;
    return HBS_OK;
}

extern "C" int hbs_block_spmm_groups(hbs_solver s, int k, int *groups)
{
    if (!s || !groups) return fail(HBS_ERR_ARG, "null argument");
    const BlockWork *w = s->blk;
    if (!w || w->k != k || w->tile == 0) return fail(HBS_ERR_ARG, "no block solve with k = %d tuned yet", k);
    *groups = w->sgroups;
    return HBS_OK;
}

extern "C" int hbs_block_tuning(hbs_solver s, int k, int *tile, int *rows_per_step, int *rows_per_wave)
{
    if (!s || !tile || !rows_per_step || !rows_per_wave) return fail(HBS_ERR_ARG, "null argument");
    const BlockWork *w = s->blk;
    if (!w || w->k != k || w->tile == 0) return fail(HBS_ERR_ARG, "no block solve with k = %d tuned yet", k);
    *tile = w->tile;
    *rows_per_step = w->v16 ? 2 : w->rb;
    *rows_per_wave = w->v16 ? 2 : (64 / pow2_ceil(w->tile)) * w->rb;
    return HBS_OK;
}

extern "C" int hbs_bicgstab_block_f64(hbs_solver s, int k, const double *B, double *X, const hbs_solve_opts *opts,
                                      hbs_solve_info *infos)
{
    if (!s || !infos) return fail(HBS_ERR_ARG, "null argument");
    if (k < 1) return fail(HBS_ERR_ARG, "need k >= 1 right-hand sides (got %d)", k);
    if (!B || !X) return fail(HBS_ERR_ARG, "B and X must be device blocks");
    if ((size_t)s->n * (size_t)k > (size_t)INT32_MAX * 8) return fail(HBS_ERR_ARG, "n * k too large");
    hbs_solve_opts o;
    hbs_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.maxit < 0 || !(o.tol >= 0.0)) return fail(HBS_ERR_ARG, "bad tolerance or iteration cap");
    DevGuard dg(s->A->ctx->device);
    hipStream_t q = s->A->ctx->stream;
    BlockWork *w = nullptr;
    CHK(block_work(s, k, &w));
    const size_t nk = (size_t)w->n * k;
    if (w->tile == 0) CHK(block_tune(s, w, q));
    hipEvent_t e0 = s->e0, e1 = s->e1;
    HIPCHK(hipEventRecord(e0, q));
    HIPCHK(hipMemcpyAsync(w->X, X, sizeof(double) * nk, hipMemcpyDeviceToDevice, q));
    std::vector<SolverState> hs((size_t)k);
    for (int c0 = 0; c0 < k; c0 += w->tile) {
        const int kc = std::min(w->tile, k - c0);
        const BlockArgs g = group_args(s, w, c0, kc);
        HIPCHK(launch_state_init(w->st, kc, o.tol, o.maxit, o.fixed_iters ? 1 : 0, q));
        // R = B - A X ; R0 = R ; P = V = 0 ; iteration 0's rho and threshold
        BlockArgs a = g;
        a.skip_done = 0;
        a.ngroups = w->sgroups;
        a.x = g.X;
        a.y = g.R;
        a.w = B + c0;
        HIPCHK(launch_block_spmm(a, BEPI_RESID, q));
        a.ngroups = w->ngroups;
        HIPCHK(launch_block_vec(a, BVEC_INIT, q));
        HIPCHK(launch_block_fin(w->st, kc, w->part, w->sgroups, FIN_XR, q));   // the SpMM's partials
        CHK(block_iterate(g, w, c0, std::max(o.maxit, 1), o.check_every, o.use_graph != 0, q));
        // sqrt(||B - AX||^2 / ||B||^2) per column
        a.x = g.X;
        a.y = g.V;
        a.w = B + c0;
        a.ngroups = w->sgroups;
        HIPCHK(launch_block_spmm(a, BEPI_RESID, q));
        HIPCHK(launch_block_fin(w->st, kc, w->part, w->sgroups, FIN_RESID, q));
        a.ngroups = w->ngroups;
        a.x = B + c0;
        HIPCHK(launch_block_vec(a, BVEC_DOT, q));
        HIPCHK(launch_block_fin(w->st, kc, w->part, w->ngroups, FIN_STORE, q));
        SolverState *h = w->hst + 2 * (size_t)kBlockCols;
        HIPCHK(hipMemcpyAsync(h, w->st, sizeof(SolverState) * kc, hipMemcpyDeviceToHost, q));
        HIPCHK(hipStreamSynchronize(q));
        std::memcpy(&hs[(size_t)c0], h, sizeof(SolverState) * kc);
    }
    HIPCHK(hipMemcpyAsync(X, w->X, sizeof(double) * nk, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipEventRecord(e1, q));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    int worst = HBS_OK, wc = 0;
    for (int c = 0; c < k; ++c) {
        const SolverState &st = hs[(size_t)c];
        hbs_solve_info &in = infos[c];
        in.iters = std::max(st.iter, 0);
        in.stop = st.stop;
        in.sweeps = 0;
        in.status = st.stop == 4 ? HBS_ERR_BREAKDOWN : ((!o.fixed_iters && st.stop == 0) ? HBS_ERR_NOCONV : HBS_OK);
        in.relres = st.out > 0 ? std::sqrt(st.resid2 / st.out) : std::sqrt(st.resid2);
        in.rel_r0 = st.rr0 > 0 ? std::sqrt(st.resid2 / st.rr0) : 0.0;
        in.time_ms = ms;
        if (in.status != worst && (in.status == HBS_ERR_BREAKDOWN || (in.status == HBS_ERR_NOCONV && worst == HBS_OK))) {
            worst = in.status;
            wc = c;
        }
    }
    if (worst == HBS_ERR_BREAKDOWN)
        return fail(worst, "block BiCGStab: breakdown in column %d after %d iterations", wc, infos[wc].iters);
    if (worst == HBS_ERR_NOCONV)
        return fail(worst, "block BiCGStab: column %d reached %d iterations, ||r||/||r0|| = %.3e", wc,
                    infos[wc].iters, infos[wc].rel_r0);
    return HBS_OK;
}
