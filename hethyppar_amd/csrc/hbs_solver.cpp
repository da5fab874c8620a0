// hbs_solver.cpp -- BiCGStab on the device: the FP64 drop-in for the
// reference loop, the mixed FP32/FP64 iterative-refinement form, and the
// adjoint solve on A^T.
//
// One iteration is five kernels (k_vec_P, SpMV+(r0,v), k_vec_S, SpMV+(t,s),
// (t,t), k_vec_XR); every dot is fused into the pass that produces its
// operand and every scalar is derived on the device, so the host only polls
// a state copy once per chunk of iterations.  Chunks are captured once into
// a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "hbs_objects.h"

using namespace hbs;

#define fail ::hbs::set_error

extern "C" void hbs_solve_opts_default(hbs_solve_opts *o)
{
    if (!o) return;
    o->tol = 1e-8;
    o->maxit = 1000;
    o->fixed_iters = 0;
    o->check_every = 4;
    o->use_graph = 2;
    o->inner_maxit = 1000;
    o->inner_tol = 1e-4;
    o->max_sweeps = 8;
}

extern "C" int hbs_solver_destroy(hbs_solver s)
{
    if (!s) return HBS_OK;
    DevGuard g(s->A->ctx->device);
    (void)hipStreamSynchronize(s->A->ctx->stream);
    for (double **p : {&s->X, &s->R, &s->R0, &s->P, &s->V, &s->S, &s->T, &s->Xb}) dfree(*p);
    for (float **p : {&s->X32, &s->R32, &s->R032, &s->P32, &s->V32, &s->S32, &s->T32}) dfree(*p);
    for (double **p : {&s->partA, &s->partB, &s->partC, &s->partD, &s->partE, &s->partF, &s->chunk_sums, &s->partX})
        dfree(*p);
    for (hbs::SysRange *&r : s->sysx) dfree(r);
    dfree(s->st);
    if (s->hst) (void)hipHostFree(s->hst);
    dfree(s->vblocks);
    dfree(s->sys);
    dfree(s->vblocks32);
    dfree(s->sys32);
    for (int i = 0; i < 2; ++i)
        for (auto &bylen : s->graph[i])
            for (hipGraphExec_t g : bylen)
                if (g) (void)hipGraphExecDestroy(g);
    for (hipEvent_t e : {s->ev[0], s->ev[1], s->e0, s->e1})
        if (e) (void)hipEventDestroy(e);
    if (s->adj) hbs_solver_destroy(s->adj);
    if (s->blk) block_work_free(s->blk);
    delete s;
    return HBS_OK;
}

extern "C" int hbs_solver_create(hbs_csr a, hbs_solver *out)
{
    if (!a || !out) return fail(HBS_ERR_ARG, "null argument");
    if (a->m != a->n) return fail(HBS_ERR_ARG, "BiCGStab needs a square matrix (%d x %d)", a->m, a->n);
    if (a->m <= 0) return fail(HBS_ERR_ARG, "empty matrix");
    *out = nullptr;
    DevGuard g(a->ctx->device);
    hbs_solver s = new hbs_solver_s();
    s->A = a;
    s->n = a->m;
    auto bail = [&](int rc) {
        std::string e = hbs_last_error();
        hbs_solver_destroy(s);
        set_error(rc, "%s", e.c_str());
        return rc;
    };
    int rc;
    s->nsys = a->nsys;
    for (double **p : {&s->X, &s->R, &s->R0, &s->P, &s->V, &s->S, &s->T})
        if ((rc = dalloc(p, s->n))) return bail(rc);
    if ((rc = dalloc(&s->st, (size_t)kNumSlots * s->nsys))) return bail(rc);
    // host-mapped and coherent: k_apply_xr writes the polled states straight into it
    if (hipHostMalloc((void **)&s->hst, kMirrorSlots * sizeof(SolverState) * (size_t)s->nsys,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&s->hst_dev, s->hst, 0) != hipSuccess)
        return bail(fail(HBS_ERR_ALLOC, "pinned host allocation failed"));
    std::memset((void *)s->hst, 0, kMirrorSlots * sizeof(SolverState) * (size_t)s->nsys);   // no stamp yet
    // vector workgroups system by system; one block size for the launch
    std::vector<VecBlock> vb;
    std::vector<SysRange> ranges = a->sys_host;
    const int epb = vec_elems_per_block(s->n);
    for (int q = 0; q < s->nsys; ++q) {
        ranges[(size_t)q].v0 = (int)vb.size();
        plan_vec(a->sys_rows[(size_t)q], a->sys_rows[(size_t)q + 1] - a->sys_rows[(size_t)q], epb, q, vb);
        ranges[(size_t)q].v1 = (int)vb.size();
    }
    s->nvb = (int)vb.size();
    // SpMV partials and chunk sums of A's plan or, for FP32 SpMVs of a
    // matrix with its diagonal apart, of its off-diagonal part's
    const int ng = std::max(a->plan.ngroups, a->od ? a->od->plan.ngroups : 0);
    const int nch = std::max(a->plan.nchunks, a->od ? a->od->plan.nchunks : 0);
    if ((rc = dalloc(&s->vblocks, vb.size())) || (rc = dalloc(&s->sys, (size_t)s->nsys))) return bail(rc);
    if ((rc = upload(s->vblocks, vb)) || (rc = upload(s->sys, ranges))) return bail(rc);
    const size_t ps = (size_t)std::max(ng, 1) * kMaxDots, pv = (size_t)std::max(s->nvb, 1) * kMaxDots;
    if ((rc = dalloc(&s->partA, ps)) || (rc = dalloc(&s->partB, pv)) || (rc = dalloc(&s->partC, ps)) ||
        (rc = dalloc(&s->partD, pv)) || (rc = dalloc(&s->partE, ps)) || (rc = dalloc(&s->partF, pv)) ||
        (rc = dalloc(&s->chunk_sums, nch)))
        return bail(rc);
    for (hipEvent_t *e : {&s->ev[0], &s->ev[1], &s->e0, &s->e1})
        if (hipEventCreate(e) != hipSuccess) return bail(fail(HBS_ERR_HIP, "event creation failed"));
    // x-segmented iterations with the P and S updates in the product passes
    if (s->nsys == 1 && g_xseg_fuse) {
        int nx = 0;
        for (int gi = 0; gi < 2; ++gi) {
            const hbs_csr M = gi == 0 ? a : (a->od ? a->od : a);
            if (M->xs_nseg <= 0 || (gi == 1 && !a->val32)) continue;
            const int nxw = M->xs_nseg * M->xs_split;
            std::vector<SysRange> r{SysRange{M->sys_host[0].g0, M->sys_host[0].g1, 0, nxw}};
            if ((rc = dalloc(&s->sysx[gi], 1)) || (rc = upload(s->sysx[gi], r))) return bail(rc);
            nx = std::max(nx, nxw);
        }
        if (nx > 0 && (rc = dalloc(&s->partX, (size_t)nx))) return bail(rc);
    }
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

// The vector workgroups (and system ranges naming them) of T's iterations:
// FP32 sweeps have their own plan of larger workgroups.
template <typename T>
const VecBlock *vblocks_of(hbs_solver s) { return sizeof(T) == 4 && s->sys32 ? s->vblocks32 : s->vblocks; }
template <typename T>
const SysRange *sys_of(hbs_solver s) { return sizeof(T) == 4 && s->sys32 ? s->sys32 : s->sys; }
template <typename T>
int nvb_of(hbs_solver s) { return sizeof(T) == 4 && s->sys32 ? s->nvb32 : s->nvb; }
template <typename T>
int epb_of(hbs_solver s) { return sizeof(T) == 4 && s->sys32 ? s->epb32 : vec_elems_per_block(s->n); }

template <typename T>
VecArgs<T> vec_args(hbs_solver s, const Work<T> &w)
{
    VecArgs<T> a{};
    a.blocks = vblocks_of<T>(s);
    a.sys = sys_of<T>(s);
    a.nvb = nvb_of<T>(s);
    a.upb = vec_packs_per_thread(epb_of<T>(s), (int)sizeof(T));
    a.X = w.X;
    a.R = w.R;
    a.P = w.P;
    a.V = w.V;
    a.S = w.S;
    a.T_ = w.T_;
    a.R0 = w.R0;
    return a;
}

template <typename T>
SpmvArgs<T> solver_spmv(hbs_solver s, const T *val)
{
    SpmvArgs<T> sa = spmv_args<T>(s->A, val);
    sa.sys = sys_of<T>(s);
    sa.chunk_sums = s->chunk_sums;
    return sa;
}

// Whether T's iterations update P and S in the product passes of an
// x-segmented SpMV (k_xseg with XFUSE_P / XFUSE_S) instead of k_vec_P and
// k_vec_S: one x-segmented system, hbs_set_xseg_fuse on (at creation).
template <typename T>
bool xseg_fused(hbs_solver s)
{
    return s->sysx[sizeof(T) == 4 ? 1 : 0] != nullptr && s->partX != nullptr;
}

// One BiCGStab iteration of an x-segmented system with the P and S updates
// in the product passes: the same steps, scalars and state slots as below,
// with the ||s|| partials from the product pass of T = A S.
template <typename T>
hipError_t launch_iteration_fused(hbs_solver s, const Work<T> &w)
{
    hipStream_t q = s->A->ctx->stream;
    SolverState *A = s->slot(kSlotA), *B = s->slot(kSlotB);
    const hbs_csr M = spmv_mat<T>(s->A);
    const int ng = M->plan.ngroups, nvb = nvb_of<T>(s), nxw = M->xs_nseg * M->xs_split;
    hipError_t e;
    // FIN_XR of the previous iteration, P = ((P - V w) b) + R in the product
    // pass ; V = A P ; (r0, v)
    XsegArgs<T> f{};
    f.fuse = XFUSE_P;
    f.xv_out = w.P;
    f.R = w.R;
    f.V = w.V;
    f.st_in = A;
    f.st_out = B;
    f.part_in = s->partD;
    f.nparts_in = nvb;
    f.pin0 = 0;
    f.pin1 = nvb;
    SpmvArgs<T> sa = solver_spmv<T>(s, w.val);
    sa.x = w.P;
    sa.y = w.V;
    sa.w = w.R0;
    sa.st_in = B;
    sa.check_done = 1;
    sa.part_out = s->partA;
    if ((e = run_spmv(sa, s->A, EPI_DOT1, B, q, &f)) != hipSuccess) return e;
    // alpha ; S = R - V a in the product pass, ||S||^2 ; T = A S ; ||s||
    // test ; (t, s), (t, t)
    XsegArgs<T> g{};
    g.fuse = XFUSE_S;
    g.xv_out = w.S;
    g.R = w.R;
    g.V = w.V;
    g.st_in = B;
    g.st_out = A;
    g.part_in = s->partA;
    g.nparts_in = ng;
    g.pin0 = M->sys_host[0].g0;
    g.pin1 = M->sys_host[0].g1;
    g.part_out = s->partX;
    sa.x = w.S;
    sa.y = w.T_;
    sa.w = w.S;
    sa.st_in = A;
    sa.st_out = B;
    sa.fin_in = FIN_S;
    sa.part_in = s->partX;
    sa.nparts_in = nxw;
    sa.sys = s->sysx[sizeof(T) == 4 ? 1 : 0];
    sa.part_out = s->partC;
    if ((e = run_spmv(sa, s->A, EPI_DOT2, A, q, &g)) != hipSuccess) return e;
    // omega ; X, R ; (r, r0), ||r||^2
    VecArgs<T> va = vec_args<T>(s, w);
    va.st_in = B;
    va.st_out = A;
    va.part_in = s->partC;
    va.nparts_in = ng;
    va.in_spmv = 1;
    va.part_out = s->partD;
    return launch_vec_XR(va, q);
}

// One BiCGStab iteration, state slots alternating A -> B -> A.
template <typename T>
hipError_t launch_iteration(hbs_solver s, const Work<T> &w)
{
    if (xseg_fused<T>(s)) return launch_iteration_fused<T>(s, w);
    hipStream_t q = s->A->ctx->stream;
    SolverState *A = s->slot(kSlotA), *B = s->slot(kSlotB);
    const int ng = spmv_mat<T>(s->A)->plan.ngroups;
    hipError_t e;
    VecArgs<T> va = vec_args<T>(s, w);
    // P = ((P - V w) b) + R, after FIN_XR of the previous iteration
    va.st_in = A;
    va.st_out = B;
    va.part_in = s->partD;
    va.nparts_in = nvb_of<T>(s);
    va.in_spmv = 0;
    if ((e = launch_vec_P(va, q)) != hipSuccess) return e;
    // V = A P ; (r0, v)
    SpmvArgs<T> sa = solver_spmv<T>(s, w.val);
    sa.x = w.P;
    sa.y = w.V;
    sa.w = w.R0;
    sa.st_in = B;
    sa.check_done = 1;
    sa.part_out = s->partA;
    if ((e = run_spmv(sa, s->A, EPI_DOT1, B, q)) != hipSuccess) return e;
    // alpha ; S = R - V a ; ||S||^2
    va.st_in = B;
    va.st_out = A;
    va.part_in = s->partA;
    va.nparts_in = ng;
    va.in_spmv = 1;
    va.part_out = s->partB;
    if ((e = launch_vec_S(va, q)) != hipSuccess) return e;
    // ||s|| test ; T = A S ; (t, s), (t, t)
    sa.x = w.S;
    sa.y = w.T_;
    sa.w = w.S;
    sa.st_in = A;
    sa.st_out = B;
    sa.fin_in = FIN_S;
    sa.part_in = s->partB;
    sa.nparts_in = nvb_of<T>(s);
    sa.part_out = s->partC;
    if ((e = run_spmv(sa, s->A, EPI_DOT2, A, q)) != hipSuccess) return e;
    // omega ; X, R ; (r, r0), ||r||^2
    va.st_in = B;
    va.st_out = A;
    va.part_in = s->partC;
    va.nparts_in = ng;
    va.in_spmv = 1;
    va.part_out = s->partD;
    return launch_vec_XR(va, q);
}

// k iterations, then the state the host polls (FIN_XR of the chunk's last
// iteration applied), mirrored into poll slot `slot` of the host buffer and
// stamped with `stamp` (0: none).
template <typename T>
hipError_t launch_chunk(hbs_solver s, const Work<T> &w, int k, int slot, int stamp)
{
    hipError_t e = hipSuccess;
    for (int i = 0; i < k && e == hipSuccess; ++i) e = launch_iteration<T>(s, w);
    if (e != hipSuccess) return e;
    return launch_apply_xr(s->slot(kSlotA), sys_of<T>(s), s->nsys, s->partD, nvb_of<T>(s),
                           s->hst_dev + (size_t)slot * s->nsys, stamp, s->A->ctx->stream);
}

// Queue one chunk of k iterations polled through slot `slot`.  Without
// graphs the chunk's states carry a new stamp, returned in *stamp; with
// graphs (whose launch arguments are fixed at capture) *stamp is 0 and the
// host waits on an event instead.
template <typename T>
int run_chunk(hbs_solver s, const Work<T> &w, int k, int slot, bool use_graph, int *stamp)
{
    hipStream_t q = s->A->ctx->stream;
    const int gi = sizeof(T) == 8 ? 0 : 1;
    *stamp = 0;
    // Chunks longer than the kept graphs launch their kernels one by one: an
    // executable graph made for one launch could only be destroyed once that
    // launch has finished, which the host learns only later.
    if (!use_graph || k >= hbs_solver_s::kGraphLens) {
        s->seq = s->seq == INT_MAX ? 1 : s->seq + 1;
        *stamp = s->seq;
        HIPCHK(launch_chunk<T>(s, w, k, slot, s->seq));
        return HBS_OK;
    }
    // FP32 graphs captured before the matrix's diagonal changed form run
    // the old row sums: drop them (no graph of this solver is in flight
    // between solves)
    if (gi == 1 && s->graph_diag_gen != s->A->diag_gen) {
        for (auto &bylen : s->graph[1])
            for (hipGraphExec_t &g : bylen)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
        s->graph_diag_gen = s->A->diag_gen;
    }
    // one graph per chunk length and poll slot, captured once and kept
    hipGraphExec_t &ge = s->graph[gi][k][slot];
    if (!ge) {
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_chunk<T>(s, w, k, slot, 0);
        hipError_t ce = hipStreamEndCapture(q, &gr);
        if (le != hipSuccess) return fail(HBS_ERR_HIP, "launch during capture failed: %s", hipGetErrorString(le));
        if (ce != hipSuccess) return fail(HBS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ie != hipSuccess) {
            ge = nullptr;
            return fail(HBS_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
        }
    }
    HIPCHK(hipGraphLaunch(ge, q));
    return HBS_OK;
}

// Longest host wait for one chunk's stamp before the solve gives up with
// HBS_ERR_HIP (hbs_set_wait_limit).
std::atomic<double> g_wait_limit_s{600.0};

// Wait until every mirrored state of poll slot `slot` carries `stamp`
// (written by the chunk's last kernel after the states), checking the
// stream for errors and the time limit now and then.
int wait_mirror(hbs_solver s, size_t first, int count, int stamp)
{
    const volatile SolverState *h = s->hst + first;
    const auto t0 = std::chrono::steady_clock::now();
    const double limit = g_wait_limit_s.load();
    for (long spin = 0;; ++spin) {
        bool all = true;
        for (int i = 0; i < count && all; ++i) all = h[i].pad == stamp;
        if (all) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return HBS_OK;
        }
        if ((spin & 1023) == 1023) {
            const hipError_t e = hipStreamQuery(s->A->ctx->stream);
            if (e != hipSuccess && e != hipErrorNotReady)
                return fail(HBS_ERR_HIP, "solver stream failed: %s", hipGetErrorString(e));
            if (e == hipSuccess) {
                // the stream drained: the stamp must be visible now
                for (int i = 0; i < count; ++i)
                    if (h[i].pad != stamp) return fail(HBS_ERR_HIP, "chunk state stamp %d missing", stamp);
                return HBS_OK;
            }
            const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (waited > limit)
                return fail(HBS_ERR_HIP, "no chunk state after %.3f s (wait limit %.3f s); work is still queued",
                            waited, limit);
            std::this_thread::yield();
        }
    }
}

int wait_stamp(hbs_solver s, int slot, int stamp)
{
    return wait_mirror(s, (size_t)slot * s->nsys, s->nsys, stamp);
}

int next_stamp(hbs_solver s)
{
    s->seq = s->seq == INT_MAX ? 1 : s->seq + 1;
    return s->seq;
}

}  // namespace

// Wait for an event recorded after a chunk, within the same time limit as
// wait_stamp.
int hbs::wait_event_limited(hipEvent_t ev)
{
    const auto t0 = std::chrono::steady_clock::now();
    const double limit = g_wait_limit_s.load();
    for (long spin = 0;; ++spin) {
        const hipError_t e = hipEventQuery(ev);
        if (e == hipSuccess) return HBS_OK;
        if (e != hipErrorNotReady) return fail(HBS_ERR_HIP, "solver stream failed: %s", hipGetErrorString(e));
        if ((spin & 63) == 63) {
            const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (waited > limit)
                return fail(HBS_ERR_HIP, "no chunk state after %.3f s (wait limit %.3f s); work is still queued",
                            waited, limit);
            std::this_thread::yield();
        }
    }
}

namespace {

bool all_done(const SolverState *h, int nsys)
{
    for (int i = 0; i < nsys; ++i)
        if (!h[i].done) return false;
    return true;
}

// Whether iteration chunks replay captured graphs.  Auto (2): from 4M
// entries an iteration's kernels run long enough (~75 us at 8.5M on
// MI355X) that launching them one by one keeps the GPU fed, and each graph
// launch would add ~13 us between chunks; smaller matrices replay graphs.
bool graphs(const hbs_csr a, int use_graph)
{
    if (use_graph == 2) return (long long)a->nnz < 4000000LL;
    return use_graph != 0;
}

// Iterations at least this long (host-measured, between polls) use the
// optimistic convergence prediction in iterate() and, once the predicted
// end is within one chunk, chunks of half the predicted rest (an iteration
// queued past convergence then costs more than a poll).
constexpr double kLongIterUs = 150.0;

// Queue iteration chunks until the device states say every system is done.
// One chunk is kept queued ahead of the one whose state the host reads, so
// the GPU never idles on the poll -- until the residual history of a single
// system predicts convergence within the chunks already queued; then the
// host waits instead of queueing iterations that would only return at once,
// sizes the next chunk to the iterations still predicted, and queues `tail`
// (the work that follows convergence) right behind the last chunk.
// *tail_queued says whether that tail still follows the last chunk.
template <typename T>
int iterate(hbs_solver s, const Work<T> &w, int maxit, int check_every, bool use_graph,
            const std::function<int()> &tail = {}, bool *tail_queued = nullptr)
{
    hipStream_t q = s->A->ctx->stream;
    const int k = std::max(1, check_every > 0 ? check_every : 4);
    long issued = 0;
    int slot = 0, nfly = 0;
    int fly_slot[2] = {0, 0}, fly_len[2] = {0, 0}, fly_stamp[2] = {0, 0};   // queued chunks, oldest first
    bool spec = false;   // tail queued after the last queued chunk
    auto queue = [&](int len) -> int {
        spec = false;
        int stamp = 0;
        CHK(run_chunk<T>(s, w, len, slot, use_graph, &stamp));
        if (!stamp) HIPCHK(hipEventRecord(s->ev[slot], q));
        fly_slot[nfly] = slot;
        fly_len[nfly] = len;
        fly_stamp[nfly] = stamp;
        ++nfly;
        slot ^= 1;
        issued += len;
        return HBS_OK;
    };
    int next = (int)std::min<long>(k, maxit);
    bool ahead = true;
    double rr_prev = -1.0;
    int it_prev = 0;
    // wall time per iteration, from the polls: long iterations make a chunk
    // queued past convergence cost more than a poll the GPU waits for
    using clk = std::chrono::steady_clock;
    clk::time_point t_prev{};
    int it_time_prev = -1;
    double us_per_iter = 0.0;
    CHK(queue(next));
    for (;;) {
        if (ahead && nfly < 2 && issued < maxit) CHK(queue((int)std::min<long>(next, maxit - issued)));
        if (fly_stamp[0]) CHK(wait_stamp(s, fly_slot[0], fly_stamp[0]));
        else CHK(wait_event_limited(s->ev[fly_slot[0]]));
        const SolverState *h = s->hst + (size_t)fly_slot[0] * s->nsys;
        const bool done = all_done(h, s->nsys);
        fly_slot[0] = fly_slot[1];
        fly_len[0] = fly_len[1];
        fly_stamp[0] = fly_stamp[1];
        --nfly;
        if (done || (issued >= maxit && nfly == 0)) break;
        ahead = true;
        next = k;
        const clk::time_point t_now = clk::now();
        if (it_time_prev >= 0 && h->iter > it_time_prev)
            us_per_iter = std::chrono::duration<double, std::micro>(t_now - t_prev).count() / (h->iter - it_time_prev);
        t_prev = t_now;
        it_time_prev = h->iter;
        if (s->nsys == 1 && !h->fixed && h->iter > 0 && h->rr > 0.0 && h->thr > 0.0) {
            // log decay of ||r||^2 per iteration over the last interval;
            // for long iterations, the faster of that and the rate since the
            // start (BiCGStab residuals are irregular, and there a chunk
            // queued past convergence costs more than a missed poll)
            double per = 0.0;
            if (rr_prev > 0.0 && h->iter > it_prev && h->rr < rr_prev)
                per = std::log(h->rr / rr_prev) / (h->iter - it_prev);
            if (us_per_iter >= kLongIterUs && h->rr0 > 0.0 && h->rr < h->rr0)
                per = std::min(per, std::log(h->rr / h->rr0) / h->iter);
            if (per < 0.0) {
                const double need = std::log(h->thr * h->thr / h->rr) / per;   // iterations after h
                const double left = need - (nfly ? fly_len[0] : 0);
                if (left <= 0.5) ahead = false;
                else if (us_per_iter >= kLongIterUs && left <= k)
                    next = (int)std::max(1.0, std::ceil(0.5 * left));   // the end in halves
                else next = (int)std::min<double>(k, std::max(1.0, std::ceil(left)));
            }
            rr_prev = h->rr;
            it_prev = h->iter;
        }
        if (nfly == 0 && issued < maxit) CHK(queue((int)std::min<long>(next, maxit - issued)));
        if (!ahead && tail && !spec) {
            CHK(tail());
            spec = true;
        }
    }
    if (tail_queued) *tail_queued = spec;
    return HBS_OK;
}

int status_of(const SolverState &st, bool fixed)
{
    if (st.stop == 4) return HBS_ERR_BREAKDOWN;
    if (!fixed && st.stop == 0) return HBS_ERR_NOCONV;
    return HBS_OK;
}

// r = b - A x (FP64) with ||r||^2 into resid2 of slot k (every system).
hipError_t residual(hbs_solver s, const double *x, const double *b, double *r, int k)
{
    hipStream_t q = s->A->ctx->stream;
    SpmvArgs<double> sa = solver_spmv<double>(s, s->A->val);
    sa.x = x;
    sa.y = r;
    sa.w = b;
    sa.part_out = s->partE;
    hipError_t e = run_spmv(sa, s->A, EPI_RESID, nullptr, q);
    if (e != hipSuccess) return e;
    return launch_finalize(s->slot(k), s->sys, s->nsys, s->partE, s->A->plan.ngroups, 1, FIN_RESID, q);
}

// r = b - A x (FP64) with ||r||^2 into resid2 and ||b||^2 into out of slot
// k (every system), in one pass.
hipError_t residual_b(hbs_solver s, const double *x, const double *b, double *r, int k)
{
    hipStream_t q = s->A->ctx->stream;
    SpmvArgs<double> sa = solver_spmv<double>(s, s->A->val);
    sa.x = x;
    sa.y = r;
    sa.w = b;
    sa.part_out = s->partE;
    hipError_t e = run_spmv(sa, s->A, EPI_RESIDB, nullptr, q);
    if (e != hipSuccess) return e;
    return launch_finalize(s->slot(k), s->sys, s->nsys, s->partE, s->A->plan.ngroups, 1, FIN_RESIDB, q);
}

void fill_info(const SolverState &hs, bool fixed, float ms, hbs_solve_info *info)
{
    info->iters = std::max(hs.iter, 0);
    info->stop = hs.stop;
    info->sweeps = 0;
    info->status = status_of(hs, fixed);
    info->relres = hs.out > 0 ? std::sqrt(hs.resid2 / hs.out) : std::sqrt(hs.resid2);
    info->rel_r0 = hs.rr0 > 0 ? std::sqrt(hs.resid2 / hs.rr0) : 0.0;
    info->time_ms = ms;
}

// FP64 BiCGStab of every system of A at once; infos[nsys] (always filled).
int solve_f64(hbs_solver s, const double *b, double *x, const hbs_solve_opts &o, hbs_solve_info *infos)
{
    hbs_csr A = s->A;
    hipStream_t q = A->ctx->stream;
    const int n = s->n;
    Work<double> w = work64(s);
    // Without graphs (whose captured launches keep the workspace pointers)
    // the iterate lives in the caller's x itself when its packs are 16-byte
    // aligned and it is not b: no copies in and out.
    const bool use_graphs = graphs(A, o.use_graph);
    const bool direct = !use_graphs && ((uintptr_t)x % 16) == 0 && x != b;
    if (direct) w.X = x;

    HIPCHK(hipEventRecord(s->e0, q));
    HIPCHK(launch_solve_start(direct ? nullptr : x, w.X, n, s->st, 2 * s->nsys, o.tol, o.maxit,
                              o.fixed_iters ? 1 : 0, q));
    // R = B - A X ; R0 = R ; P = V = 0 ; (R, R0) = ||R||^2 pending in slot A
    // (bicgstab_cpu.c:176-188)
    {
        SpmvArgs<double> sa = solver_spmv<double>(s, A->val);
        sa.x = w.X;
        sa.y = s->R;
        sa.w = b;
        sa.part_out = s->partE;
        HIPCHK(run_spmv(sa, A, EPI_RESID, nullptr, q));
        VecArgs<double> va = vec_args<double>(s, w);
        va.st_out = s->slot(kSlotA);
        va.part_out = s->partD;
        HIPCHK(launch_init_vectors(va, q));
    }
    // final relative residual sqrt(||B - AX||^2 / ||B||^2) (bicgstab_cpu.c:251-255)
    // into S (dead between iterations), ||B||^2, x = X; queued by iterate
    // behind the chunk predicted to be the last
    // The final states go to the host-mapped mirror (read buffer 2) before
    // e1, so the host reads them once e1 has completed.
    auto tail = [&]() -> int {
        HIPCHK(residual_b(s, w.X, b, s->S, kSlotA));
        HIPCHK(launch_mirror(s->slot(kSlotA), nullptr, s->nsys, s->hst_dev + 2 * (size_t)s->nsys, 0, q));
        if (!direct) HIPCHK(hipMemcpyAsync(x, w.X, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
        HIPCHK(hipEventRecord(s->e1, q));
        return HBS_OK;
    };
    // x aliasing b: no tail ahead of the convergence the host has seen, since
    // its copy of the iterate into x would overwrite the b that a later tail
    // (after a misprediction) measures the residual against
    bool tail_done = false;
    CHK(iterate<double>(s, w, std::max(o.maxit, 1), o.check_every, use_graphs,
                        x == b ? std::function<int()>{} : std::function<int()>{tail}, &tail_done));
    if (!tail_done) CHK(tail());
    CHK(wait_event_limited(s->e1));
    std::vector<SolverState> hs((size_t)s->nsys);
    std::memcpy(hs.data(), (const void *)(s->hst + 2 * (size_t)s->nsys), sizeof(SolverState) * (size_t)s->nsys);
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
    for (int i = 0; i < s->nsys; ++i) fill_info(hs[(size_t)i], o.fixed_iters != 0, ms, &infos[i]);
    return HBS_OK;
}

int check_opts(const hbs_solve_opts *opts, hbs_solve_opts *o)
{
    hbs_solve_opts_default(o);
    if (opts) *o = *opts;
    if (o->maxit < 0 || !(o->tol >= 0.0)) return fail(HBS_ERR_ARG, "bad tolerance or iteration cap");
    return HBS_OK;
}

int report(const hbs_solve_info &in, const char *what, int sys)
{
    if (in.status == HBS_ERR_BREAKDOWN)
        return fail(in.status, "%s: breakdown after %d iterations (system %d)", what, in.iters, sys);
    if (in.status == HBS_ERR_NOCONV)
        return fail(in.status, "%s reached %d iterations, ||r||/||r0|| = %.3e (system %d)", what, in.iters,
                    in.rel_r0, sys);
    return HBS_OK;
}

// Worst status first (breakdown, then not converged), else the first system.
int worst_system(const std::vector<hbs_solve_info> &infos)
{
    int w = 0;
    for (int i = 1; i < (int)infos.size(); ++i) {
        const int a = infos[(size_t)i].status, b = infos[(size_t)w].status;
        if (a != b && (a == HBS_ERR_BREAKDOWN || (a == HBS_ERR_NOCONV && b == HBS_OK))) w = i;
    }
    return w;
}

}  // namespace

extern "C" int hbs_set_wait_limit(double seconds)
{
    if (!(seconds > 0.0)) return fail(HBS_ERR_ARG, "wait limit must be positive");
    g_wait_limit_s.store(seconds);
    return HBS_OK;
}

extern "C" int hbs_bicgstab_f64(hbs_solver s, const double *b, double *x, const hbs_solve_opts *opts,
                                hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    if (!b || !x) return fail(HBS_ERR_ARG, "b and x must be device vectors");
    hbs_solve_opts o;
    CHK(check_opts(opts, &o));
    DevGuard g(s->A->ctx->device);
    std::vector<hbs_solve_info> infos((size_t)s->nsys);
    CHK(solve_f64(s, b, x, o, infos.data()));
    // several systems: the most iterations, the largest residuals, the worst status
    hbs_solve_info agg = infos[(size_t)worst_system(infos)];
    for (const hbs_solve_info &in : infos) {
        agg.iters = std::max(agg.iters, in.iters);
        agg.relres = std::max(agg.relres, in.relres);
        agg.rel_r0 = std::max(agg.rel_r0, in.rel_r0);
    }
    if (info) *info = agg;
    return report(infos[(size_t)worst_system(infos)], "BiCGStab", worst_system(infos));
}

extern "C" int hbs_bicgstab_systems_f64(hbs_solver s, const double *b, double *x, const hbs_solve_opts *opts,
                                        hbs_solve_info *infos)
{
    if (!s || !infos) return fail(HBS_ERR_ARG, "null argument");
    if (!b || !x) return fail(HBS_ERR_ARG, "b and x must be device vectors");
    hbs_solve_opts o;
    CHK(check_opts(opts, &o));
    DevGuard g(s->A->ctx->device);
    std::vector<hbs_solve_info> in((size_t)s->nsys);
    CHK(solve_f64(s, b, x, o, in.data()));
    std::memcpy(infos, in.data(), sizeof(hbs_solve_info) * in.size());
    const int w = worst_system(in);
    return report(in[(size_t)w], "BiCGStab", w);
}

extern "C" int hbs_bicgstab_mixed(hbs_solver s, const double *b, double *x, const hbs_solve_opts *opts,
                                  hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    if (!b || !x) return fail(HBS_ERR_ARG, "b and x must be device vectors");
    hbs_csr A = s->A;
    if (!A->val32) return fail(HBS_ERR_ARG, "mixed precision needs a matrix created with HBS_CSR_F32");
    if (s->nsys != 1) return fail(HBS_ERR_ARG, "mixed precision solves one system (matrix has %d)", s->nsys);
    hbs_solve_opts o;
    hbs_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.maxit < 1 || !(o.tol >= 0.0) || o.max_sweeps < 1) return fail(HBS_ERR_ARG, "bad options");
    DevGuard g(A->ctx->device);
    hipStream_t q = A->ctx->stream;
    const int n = s->n;
    if (!s->X32) {
        for (float **p : {&s->X32, &s->R32, &s->R032, &s->P32, &s->V32, &s->S32, &s->T32}) CHK(dalloc(p, n));
        CHK(dalloc(&s->Xb, n));
        // FP32 vector workgroups: twice the FP64 elements (at most 8192),
        // never more workgroups than the FP64 plan, whose partial buffers
        // they share
        std::vector<VecBlock> vb;
        const int epb = std::min(2 * vec_elems_per_block(n), 8192);
        plan_vec(0, n, epb, 0, vb);
        std::vector<SysRange> r = spmv_mat<float>(A)->sys_host;   // the FP32 SpMVs' workgroups
        r[0].v0 = 0;
        r[0].v1 = (int)vb.size();
        if ((int)vb.size() > s->nvb) return fail(HBS_ERR_ARG, "FP32 vector plan larger than the FP64 one");
        // sys32 is set last: the FP32 iterations use the plan only once it
        // is whole
        VecBlock *blocks = nullptr;
        SysRange *sys = nullptr;
        CHK(dalloc(&blocks, vb.size()));
        s->vblocks32 = blocks;
        CHK(upload(blocks, vb));
        CHK(dalloc(&sys, 1));
        if (int rc = upload(sys, r)) {
            dfree(sys);
            return rc;
        }
        s->nvb32 = (int)vb.size();
        s->epb32 = epb;
        s->sys32 = sys;
    }
    Work<float> w32 = work32(s);
    SolverState *outer = s->slot(kSlotOuter);
    // The host reads states from the host-mapped mirror: read buffer 3 the
    // sweep's inner state, 4 the outer one, each stamped after a fence.
    SolverState *const m_inner = s->hst_dev + 3;
    const volatile SolverState *const h_inner = s->hst + 3, *const h_outer = s->hst + 4;
    // the iterate alternates between X and Xb: a sweep reads one and writes
    // the other, so its update can be queued again after a misprediction
    double *xc = s->X, *xn = s->Xb;

    HIPCHK(hipEventRecord(s->e0, q));
    HIPCHK(hipMemcpyAsync(xc, x, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(launch_state_init(outer, 1, o.tol, 0, 0, q));
    // FP64 outer residual r = b - A x in R with ||r||^2 and ||b||^2 (one
    // pass).  The host does not wait for them: the first sweep's inner
    // tolerance does not depend on ||r0||, and the first sweep's state
    // carries ||r0|| (its scale) to the host with its result.
    HIPCHK(residual_b(s, xc, b, s->R, kSlotOuter));
    int stamp = 0;
    SolverState hs;
    double r0 = 0.0, bnorm2v = 0.0, thr = 0.0, rnorm = 0.0;
    int total_iters = 0, sweeps = 0, stop = 2;
    bool converged = false;
    bool breakdown = false;
    while (!converged && sweeps < o.max_sweeps && total_iters < o.maxit) {
        // inner tolerance: the requested one, or just what is still missing
        // (0.5 thr / ||r||, which is 0.5 tol for the first sweep)
        double itol = std::max(o.inner_tol, sweeps == 0 ? 0.5 * o.tol : 0.5 * thr / rnorm);
        if (itol > 0.5) itol = 0.5;
        const int imax = std::max(1, std::min(o.inner_maxit, o.maxit - total_iters));
        HIPCHK(launch_state_init(s->st, 2, itol, imax, 0, q));
        VecArgs<float> va = vec_args<float>(s, w32);
        va.st_out = s->slot(kSlotA);
        va.part_out = s->partD;
        HIPCHK(launch_ir_start(va, s->R, outer, q));
        va.st_in = s->slot(kSlotA);
        // the sweep's tail: x_next = x + scale * d, r = b - A x_next (FP64),
        // then both states to the mirror.  Queued behind the chunk predicted
        // to be the sweep's last, and again if more chunks follow it.
        auto tail = [&]() -> int {
            HIPCHK(launch_ir_update(va, xc, xn, q));
            HIPCHK(residual(s, xn, b, s->R, kSlotOuter));
            stamp = next_stamp(s);
            HIPCHK(launch_mirror(s->slot(kSlotA), outer, 1, m_inner, stamp, q));
            return HBS_OK;
        };
        bool tail_done = false;
        CHK(iterate<float>(s, w32, imax, o.check_every, graphs(s->A, o.use_graph), tail, &tail_done));
        if (!tail_done) CHK(tail());
        CHK(wait_mirror(s, 3, 2, stamp));
        SolverState hi;
        std::memcpy(&hi, (const void *)h_inner, sizeof hi);
        std::memcpy(&hs, (const void *)h_outer, sizeof hs);
        if (sweeps == 0) {
            // ||r0|| as the first sweep's start scaled by it, ||b||^2 as
            // the outer residual left it
            r0 = hi.scale;
            bnorm2v = hs.out;
            thr = o.tol * r0;
            if (!(r0 > 0.0)) {
                // x solves the system already: the sweep (on r = 0) is
                // discarded and x stays as it came
                rnorm = r0;
                converged = true;
                break;
            }
        }
        std::swap(xc, xn);
        total_iters += std::max(hi.iter, 0);
        ++sweeps;
        rnorm = std::sqrt(hs.resid2);
        stop = hi.stop;
        if (rnorm < thr) {
            converged = true;
        } else if (hi.stop == 4 && hi.iter <= 1) {
            breakdown = true;
            break;
        }
    }
    HIPCHK(hipMemcpyAsync(x, xc, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipEventRecord(s->e1, q));
    CHK(wait_event_limited(s->e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
    const int status = converged ? HBS_OK : (breakdown ? HBS_ERR_BREAKDOWN : HBS_ERR_NOCONV);
    if (info) {
        info->iters = total_iters;
        info->stop = breakdown ? 4 : stop;
        info->sweeps = sweeps;
        info->status = status;
        info->relres = bnorm2v > 0 ? rnorm / std::sqrt(bnorm2v) : rnorm;
        info->rel_r0 = r0 > 0 ? rnorm / r0 : 0.0;
        info->time_ms = ms;
    }
    if (status != HBS_OK)
        return fail(status, "mixed-precision BiCGStab: %d sweeps, %d iterations, ||r||/||r0|| = %.3e", sweeps,
                    total_iters, r0 > 0 ? rnorm / r0 : 0.0);
    return HBS_OK;
}

extern "C" int hbs_adjoint_f64(hbs_solver s, const double *gx, const double *x, double *lambda, double *grad_val,
                               const hbs_solve_opts *opts, hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    if (!gx || !x || !lambda) return fail(HBS_ERR_ARG, "g, x and lambda must be device vectors");
    hbs_csr A = s->A;
    if (!A->tr) return fail(HBS_ERR_ARG, "adjoint needs a matrix created with HBS_CSR_TRANSPOSE");
    DevGuard g(A->ctx->device);
    if (!s->adj) CHK(hbs_solver_create(A->tr, &s->adj));
    HIPCHK(launch_fill_zero(lambda, sizeof(double) * (size_t)s->n, A->ctx->stream));
    CHK(hbs_bicgstab_f64(s->adj, gx, lambda, opts, info));
    if (grad_val) HIPCHK(launch_grad_pattern(A->m, A->rowptr, A->col, lambda, x, grad_val, A->ctx->stream));
    HIPCHK(hipStreamSynchronize(A->ctx->stream));
    return HBS_OK;
}
