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
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hbs_objects.h"

using namespace hbs;

#define fail ::hbs::set_error

namespace {
constexpr int kSlotA = 0, kSlotB = 1, kSlotOuter = 2, kNumSlots = 4;
}

struct hbs_solver_s {
    hbs_csr A = nullptr;
    int n = 0;
    double *X = nullptr, *R = nullptr, *R0 = nullptr, *P = nullptr, *V = nullptr, *S = nullptr, *T = nullptr;
    float *X32 = nullptr, *R32 = nullptr, *R032 = nullptr, *P32 = nullptr, *V32 = nullptr, *S32 = nullptr,
          *T32 = nullptr;
    SolverState *st = nullptr;    // device [kNumSlots]
    SolverState *hst = nullptr;   // pinned host [3]
    VecBlock *vblocks = nullptr;
    SysRange *sys = nullptr;      // {0, ngroups, 0, nvb}
    int nvb = 0;
    // partials, one buffer per producer so a consumer never sees them overwritten
    double *partA = nullptr;      // SpMV1: (r0, v)
    double *partB = nullptr;      // k_vec_S: ||s||^2
    double *partC = nullptr;      // SpMV2: (t, s), (t, t)
    double *partD = nullptr;      // k_vec_XR / start: (r, r0), ||r||^2
    double *partE = nullptr;      // residual SpMV: ||b - Ax||^2
    double *partF = nullptr;      // dots: (b, b)
    double *chunk_sums = nullptr; // long-row chunk sums
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // [0] FP64, [1] FP32
    int graph_chunk[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hbs_solver_s *adj = nullptr;
};

extern "C" void hbs_solve_opts_default(hbs_solve_opts *o)
{
    if (!o) return;
    o->tol = 1e-8;
    o->maxit = 1000;
    o->fixed_iters = 0;
    o->check_every = 4;
    o->use_graph = 1;
    o->inner_maxit = 1000;
    o->inner_tol = 1e-4;
    o->max_sweeps = 8;
}

extern "C" int hbs_solver_destroy(hbs_solver s)
{
    if (!s) return HBS_OK;
    DevGuard g(s->A->ctx->device);
    (void)hipStreamSynchronize(s->A->ctx->stream);
    for (double **p : {&s->X, &s->R, &s->R0, &s->P, &s->V, &s->S, &s->T}) dfree(*p);
    for (float **p : {&s->X32, &s->R32, &s->R032, &s->P32, &s->V32, &s->S32, &s->T32}) dfree(*p);
    for (double **p : {&s->partA, &s->partB, &s->partC, &s->partD, &s->partE, &s->partF, &s->chunk_sums})
        dfree(*p);
    dfree(s->st);
    if (s->hst) (void)hipHostFree(s->hst);
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
    for (double **p : {&s->X, &s->R, &s->R0, &s->P, &s->V, &s->S, &s->T})
        if ((rc = dalloc(p, s->n))) return bail(rc);
    if ((rc = dalloc(&s->st, kNumSlots))) return bail(rc);
    if (hipHostMalloc((void **)&s->hst, 3 * sizeof(SolverState)) != hipSuccess)
        return bail(fail(HBS_ERR_ALLOC, "pinned host allocation failed"));
    std::vector<VecBlock> vb;
    plan_vec(0, s->n, vec_elems_per_block(s->n), 0, vb);
    s->nvb = (int)vb.size();
    const int ng = a->plan.ngroups;
    if ((rc = dalloc(&s->vblocks, vb.size())) || (rc = dalloc(&s->sys, 1))) return bail(rc);
    if ((rc = upload(s->vblocks, vb)) || (rc = upload(s->sys, std::vector<SysRange>{{0, ng, 0, s->nvb}})))
        return bail(rc);
    const size_t ps = (size_t)std::max(ng, 1) * kMaxDots, pv = (size_t)std::max(s->nvb, 1) * kMaxDots;
    if ((rc = dalloc(&s->partA, ps)) || (rc = dalloc(&s->partB, pv)) || (rc = dalloc(&s->partC, ps)) ||
        (rc = dalloc(&s->partD, pv)) || (rc = dalloc(&s->partE, ps)) || (rc = dalloc(&s->partF, pv)) ||
        (rc = dalloc(&s->chunk_sums, a->plan.nchunks)))
        return bail(rc);
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
VecArgs<T> vec_args(hbs_solver s, const Work<T> &w)
{
    VecArgs<T> a{};
    a.blocks = s->vblocks;
    a.sys = s->sys;
    a.nvb = s->nvb;
    a.upb = vec_packs_per_thread(vec_elems_per_block(s->n), (int)sizeof(T));
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
    sa.sys = s->sys;
    sa.chunk_sums = s->chunk_sums;
    return sa;
}

// One BiCGStab iteration, state slots alternating A -> B -> A.
template <typename T>
hipError_t launch_iteration(hbs_solver s, const Work<T> &w)
{
    hipStream_t q = s->A->ctx->stream;
    SolverState *A = s->st + kSlotA, *B = s->st + kSlotB;
    const int ng = s->A->plan.ngroups;
    hipError_t e;
    VecArgs<T> va = vec_args<T>(s, w);
    // P = ((P - V w) b) + R, after FIN_XR of the previous iteration
    va.st_in = A;
    va.st_out = B;
    va.part_in = s->partD;
    va.nparts_in = s->nvb;
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
    sa.nparts_in = s->nvb;
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

template <typename T>
hipError_t launch_chunk(hbs_solver s, const Work<T> &w, int k)
{
    hipError_t e = hipSuccess;
    for (int i = 0; i < k && e == hipSuccess; ++i) e = launch_iteration<T>(s, w);
    if (e != hipSuccess) return e;
    // the state the host polls: FIN_XR of the chunk's last iteration applied
    return launch_apply_xr(s->st + kSlotA, s->sys, 1, s->partD, s->nvb, s->A->ctx->stream);
}

template <typename T>
int run_chunk(hbs_solver s, const Work<T> &w, int k, bool use_graph)
{
    hipStream_t q = s->A->ctx->stream;
    const int gi = sizeof(T) == 8 ? 0 : 1;
    if (!use_graph) {
        HIPCHK(launch_chunk<T>(s, w, k));
        return HBS_OK;
    }
    if (!s->graph[gi] || s->graph_chunk[gi] != k) {
        if (s->graph[gi]) {
            HIPCHK(hipGraphExecDestroy(s->graph[gi]));
            s->graph[gi] = nullptr;
        }
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_chunk<T>(s, w, k);
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

// Queue iteration chunks until the device state says done.  The state of
// chunk c is read after chunk c + 1 is queued, so the GPU never idles on the
// poll; iterations queued after convergence return at once.
template <typename T>
int iterate(hbs_solver s, const Work<T> &w, int maxit, int check_every, bool use_graph)
{
    hipStream_t q = s->A->ctx->stream;
    const int k = check_every > 0 ? check_every : 4;
    long issued = 0;
    int c = 0, pending = -1;
    for (;;) {
        CHK(run_chunk<T>(s, w, k, use_graph));
        issued += k;
        HIPCHK(hipMemcpyAsync(&s->hst[c & 1], s->st + kSlotA, sizeof(SolverState), hipMemcpyDeviceToHost, q));
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

// r = b - A x (FP64) with ||r||^2 into st[slot].resid2.
hipError_t residual(hbs_solver s, const double *x, const double *b, double *r, int slot)
{
    hipStream_t q = s->A->ctx->stream;
    SpmvArgs<double> sa = solver_spmv<double>(s, s->A->val);
    sa.x = x;
    sa.y = r;
    sa.w = b;
    sa.part_out = s->partE;
    hipError_t e = run_spmv(sa, s->A, EPI_RESID, nullptr, q);
    if (e != hipSuccess) return e;
    return launch_finalize(s->st + slot, s->sys, 1, s->partE, s->A->plan.ngroups, 1, FIN_RESID, q);
}

// (b, b) into st[slot].out.
hipError_t bnorm2(hbs_solver s, const double *b, int slot)
{
    hipStream_t q = s->A->ctx->stream;
    VecArgs<double> va = vec_args<double>(s, work64(s));
    va.part_out = s->partF;
    hipError_t e = launch_dot(va, b, b, q);
    if (e != hipSuccess) return e;
    return launch_finalize(s->st + slot, s->sys, 1, s->partF, s->nvb, 0, FIN_STORE, q);
}

}  // namespace

extern "C" int hbs_bicgstab_f64(hbs_solver s, const double *b, double *x, const hbs_solve_opts *opts,
                                hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    if (!b || !x) return fail(HBS_ERR_ARG, "b and x must be device vectors");
    hbs_solve_opts o;
    hbs_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.maxit < 0 || !(o.tol >= 0.0)) return fail(HBS_ERR_ARG, "bad tolerance or iteration cap");
    hbs_csr A = s->A;
    DevGuard g(A->ctx->device);
    hipStream_t q = A->ctx->stream;
    const int n = s->n;
    Work<double> w = work64(s);

    HIPCHK(hipEventRecord(s->e0, q));
    HIPCHK(hipMemcpyAsync(s->X, x, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(launch_state_init(s->st, 2, o.tol, o.maxit, o.fixed_iters ? 1 : 0, q));
    // R = B - A X ; R0 = R ; P = V = 0 ; (R, R0) = ||R||^2 pending in slot A
    // (bicgstab_cpu.c:176-188)
    {
        SpmvArgs<double> sa = solver_spmv<double>(s, A->val);
        sa.x = s->X;
        sa.y = s->R;
        sa.w = b;
        sa.part_out = s->partE;
        HIPCHK(run_spmv(sa, A, EPI_RESID, nullptr, q));
        VecArgs<double> va = vec_args<double>(s, w);
        va.st_out = s->st + kSlotA;
        va.part_out = s->partD;
        HIPCHK(launch_init_vectors(va, q));
    }
    CHK(iterate<double>(s, w, std::max(o.maxit, 1), o.check_every, o.use_graph != 0));
    // final relative residual sqrt(||B - AX||^2 / ||B||^2) (bicgstab_cpu.c:251-255)
    HIPCHK(residual(s, s->X, b, s->V, kSlotA));
    HIPCHK(bnorm2(s, b, kSlotA));
    HIPCHK(hipMemcpyAsync(x, s->X, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipEventRecord(s->e1, q));
    SolverState hs;
    CHK(read_state(s, kSlotA, &hs));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
    const int iters = std::max(hs.iter, 0);
    const int status = status_of(hs, o.fixed_iters != 0);
    if (info) {
        info->iters = iters;
        info->stop = hs.stop;
        info->sweeps = 0;
        info->status = status;
        info->relres = hs.out > 0 ? std::sqrt(hs.resid2 / hs.out) : std::sqrt(hs.resid2);
        info->rel_r0 = hs.rr0 > 0 ? std::sqrt(hs.resid2 / hs.rr0) : 0.0;
        info->time_ms = ms;
    }
    if (status == HBS_ERR_BREAKDOWN) return fail(status, "BiCGStab breakdown after %d iterations", iters);
    if (status == HBS_ERR_NOCONV)
        return fail(status, "BiCGStab reached %d iterations, ||r||/||r0|| = %.3e", iters,
                    hs.rr0 > 0 ? std::sqrt(hs.resid2 / hs.rr0) : 0.0);
    return HBS_OK;
}

extern "C" int hbs_bicgstab_mixed(hbs_solver s, const double *b, double *x, const hbs_solve_opts *opts,
                                  hbs_solve_info *info)
{
    if (!s) return fail(HBS_ERR_ARG, "null solver");
    if (!b || !x) return fail(HBS_ERR_ARG, "b and x must be device vectors");
    hbs_csr A = s->A;
    if (!A->val32) return fail(HBS_ERR_ARG, "mixed precision needs a matrix created with HBS_CSR_F32");
    hbs_solve_opts o;
    hbs_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.maxit < 1 || !(o.tol >= 0.0) || o.max_sweeps < 1) return fail(HBS_ERR_ARG, "bad options");
    DevGuard g(A->ctx->device);
    hipStream_t q = A->ctx->stream;
    const int n = s->n;
    if (!s->X32)
        for (float **p : {&s->X32, &s->R32, &s->R032, &s->P32, &s->V32, &s->S32, &s->T32}) CHK(dalloc(p, n));
    Work<float> w32 = work32(s);

    HIPCHK(hipEventRecord(s->e0, q));
    HIPCHK(hipMemcpyAsync(s->X, x, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(launch_state_init(s->st + kSlotOuter, 1, o.tol, 0, 0, q));
    // FP64 outer residual r = b - A x in R, ||b||^2
    HIPCHK(residual(s, s->X, b, s->R, kSlotOuter));
    HIPCHK(bnorm2(s, b, kSlotOuter));
    SolverState hs;
    CHK(read_state(s, kSlotOuter, &hs));
    const double r0 = std::sqrt(hs.resid2);
    const double bnorm2v = hs.out;
    const double thr = o.tol * r0;
    double rnorm = r0;
    int total_iters = 0, sweeps = 0, stop = 2;
    bool converged = !(r0 > 0.0);
    bool breakdown = false;
    while (!converged && sweeps < o.max_sweeps && total_iters < o.maxit) {
        // inner tolerance: the requested one, or just what is still missing
        double itol = std::max(o.inner_tol, 0.5 * thr / rnorm);
        if (itol > 0.5) itol = 0.5;
        const int imax = std::max(1, std::min(o.inner_maxit, o.maxit - total_iters));
        HIPCHK(launch_state_init(s->st, 2, itol, imax, 0, q));
        VecArgs<float> va = vec_args<float>(s, w32);
        va.st_out = s->st + kSlotA;
        va.part_out = s->partD;
        HIPCHK(launch_ir_start(va, s->R, s->st + kSlotOuter, q));
        CHK(iterate<float>(s, w32, imax, o.check_every, o.use_graph != 0));
        va.st_in = s->st + kSlotA;
        HIPCHK(launch_ir_update(va, s->X, q));
        HIPCHK(residual(s, s->X, b, s->R, kSlotOuter));
        SolverState hi;
        CHK(read_state(s, kSlotA, &hi));
        CHK(read_state(s, kSlotOuter, &hs));
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
    HIPCHK(hipMemcpyAsync(x, s->X, sizeof(double) * n, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipEventRecord(s->e1, q));
    HIPCHK(hipEventSynchronize(s->e1));
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
