// hbs_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the BiCGStab/SpMV engine.
//
// Everything here is HBM-bandwidth bound (no dense contraction, so no MFMA):
// the kernels are written for coalesced 16-byte streams, wave64 reductions and
// XCD-aware workgroup placement.
//
//  * k_spmv       CSR-stream SpMV.  A workgroup owns a row block whose entries
//                 fit in LDS: phase 1 streams col/val coalesced and writes one
//                 product per entry to LDS; phase 2 sums each row.  Short-row
//                 blocks sum serially in CSR order (bit-identical to the
//                 reference CSR_spmxv_seq, hhp_cpu.c:445-455); blocks of longer
//                 rows use 2..64 lanes per row; rows longer than a block are
//                 split into chunks whose sums are combined in chunk order by
//                 the last chunk to finish.  Epilogues fuse the dot products
//                 BiCGStab needs after each SpMV.
//  * k_vec_*      the three BiCGStab vector updates, each fused with the
//                 reductions that follow it (reference k_vecop_P/S/XR,
//                 hhp_dp_kernels.cu:51-74, plus the dots of
//                 bicgstab_cpu.c:203,229,242).
//  * reductions   deterministic: every workgroup stores its partial, the last
//                 one to arrive (atomic counter) sums the partials in a fixed
//                 order and updates the device-resident scalars.
#include <hip/hip_runtime.h>

#include "hbs_internal.h"

namespace hbs {
namespace {

constexpr int kXcds = 8;   // MI355X: 8 XCDs, workgroups dispatched round robin

// Give each XCD a contiguous range of logical blocks so neighbouring row
// blocks (which touch neighbouring x entries) share one XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk)
{
    const int q = nblk / kXcds, r = nblk % kXcds;
    const int x = bid % kXcds, k = bid / kXcds;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ double ld_agent(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-thread dot accumulators.  FP64: plain sequential sum of rounded
// products.  FP32: Dot2 of Ogita, Rump and Oishi (2005) -- TwoProduct by FMA
// and Knuth's TwoSum, the error terms summed in a second float -- handed to
// the FP64 workgroup / grid reduction.
template <typename T>
struct Acc;
template <>
struct Acc<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double a, double b) { s += a * b; }
    __device__ __forceinline__ double get() const { return s; }
};
template <>
struct Acc<float> {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float a, float b)
    {
        const float p = a * b;
        const float e = __builtin_fmaf(a, b, -p);
        const float t = s + p;
        const float z = t - s;
        const float q = (s - (t - z)) + (p - z);
        s = t;
        c += q + e;
    }
    __device__ __forceinline__ double get() const { return (double)s + (double)c; }
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroup sum of ND values; every thread gets the totals.  Fixed order:
// xor butterfly inside each wave, then waves 0..3 in sequence.
template <int ND>
__device__ __forceinline__ void block_sum(double (&v)[ND], double *red)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = wave_sum(v[d]);
    if (l == 0) {
#pragma unroll
        for (int d = 0; d < ND; ++d) red[d * kWaves + w] = v[d];
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double t = red[d * kWaves];
#pragma unroll
        for (int k = 1; k < kWaves; ++k) t += red[d * kWaves + k];
        v[d] = t;
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ T block_sum1(T v, T *red)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    v = wave_sum(v);
    if (l == 0) red[w] = v;
    __syncthreads();
    T t = red[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) t += red[k];
    __syncthreads();
    return t;
}

// Thread 0 has stored this workgroup's partials; count the arrival and, in
// the last workgroup of the system, sum partials [b0, b1) in a fixed order.
template <int ND>
__device__ bool arrive_and_reduce(double (&v)[ND], const double *part, int nslots, unsigned *ctr,
                                  int b0, int b1, double *red, int *s_flag)
{
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(ctr, 1u);
        const int last = prev == (unsigned)(b1 - b0 - 1);
        if (last) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_flag = last;
    }
    __syncthreads();
    const int last = *s_flag;
    __syncthreads();
    if (!last) return false;
    __threadfence();
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double acc = 0.0;
        for (int i = b0 + (int)threadIdx.x; i < b1; i += kThreads) acc += ld_agent(&part[(size_t)d * nslots + i]);
        v[d] = acc;
    }
    block_sum<ND>(v, red);
    return true;
}

__device__ __forceinline__ void mark_done(SolverState *st, int stop, unsigned *ndone)
{
    st->done = 1;
    st->stop = stop;
    if (ndone) atomicAdd(ndone, 1u);
}

// Scalar updates of one system (thread 0 of the last workgroup).
// Order of operations follows bicgstab_cpu.c:203-229.
template <int ND>
__device__ void finalize(SolverState *st, int fin, const double (&v)[ND], unsigned *ndone)
{
    const double d0 = ND > 0 ? v[0] : 0.0;
    const double d1 = ND > 1 ? v[1] : 0.0;
    switch (fin) {
    case FIN_ALPHA:
        st->rv = d0;
        st->alpha = st->rho / d0;
        if (d0 == 0.0 || !isfinite(st->alpha)) mark_done(st, 4, ndone);
        break;
    case FIN_OMEGA:
        st->ts = d0;
        st->tt = d1;
        st->omega = d0 / d1;
        if (d1 == 0.0 || !isfinite(st->omega)) mark_done(st, 4, ndone);
        break;
    case FIN_INIT:
        st->rr0 = d0;
        st->thr = st->tol * sqrt(d0);
        st->rho_next = d0;
        if (st->done) break;
        if (d0 == 0.0) { mark_done(st, 2, ndone); break; }
        if (!isfinite(d0)) { mark_done(st, 4, ndone); break; }
        st->beta = (d0 / st->rho) * (st->alpha / st->omega);
        st->rho = d0;
        break;
    case FIN_RESID:
        st->resid2 = d0;
        break;
    case FIN_S:
        st->ss = d0;
        if (!st->fixed && sqrt(d0) < st->thr) st->s_conv = 1;
        break;
    case FIN_XR:
        st->iter += 1;
        if (st->s_conv) { mark_done(st, 1, ndone); break; }
        st->rho_next = d0;
        st->rr = d1;
        if (!st->fixed && sqrt(d1) < st->thr) { mark_done(st, 2, ndone); break; }
        if (st->iter >= st->maxit) { mark_done(st, 0, ndone); break; }
        if (d0 == 0.0 || !isfinite(d0)) { mark_done(st, 4, ndone); break; }
        st->beta = (d0 / st->rho) * (st->alpha / st->omega);
        st->rho = d0;
        break;
    case FIN_STORE:
        st->out = d0;
        break;
    default:
        break;
    }
}

template <int EPI>
struct EpiDots {
    static constexpr int n = EPI == EPI_DOT2 ? 2 : (EPI == EPI_AXPBY ? 0 : 1);
};

// Row epilogue: write y[r] and feed the fused dots.
template <typename T, int EPI>
__device__ __forceinline__ void row_out(const SpmvArgs<T> &a, int r, T acc, Acc<T> *d)
{
    if constexpr (EPI == EPI_AXPBY) {
        T out = a.alpha * acc;
        if (a.beta != T(0)) out = out + a.beta * a.y[r];
        a.y[r] = out;
    } else if constexpr (EPI == EPI_DOT1) {
        a.y[r] = acc;
        d[0].add(a.w[r], acc);
    } else if constexpr (EPI == EPI_DOT2) {
        a.y[r] = acc;
        d[0].add(acc, a.w[r]);
        d[1].add(acc, acc);
    } else {
        const T out = a.w[r] - acc;
        a.y[r] = out;
        d[0].add(out, out);
    }
}

template <typename T, int EPI>
__global__ __launch_bounds__(kThreads) void k_spmv(SpmvArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *lds = reinterpret_cast<T *>(smem);
    __shared__ double red[kMaxDots * kWaves];
    __shared__ T tred[kWaves];
    __shared__ int s_flag;
    constexpr int ND = EpiDots<EPI>::n;

    const int b = xcd_remap(blockIdx.x, a.nblk);
    const RowBlock rb = a.blocks[b];
    const BlockAux ax = a.aux[b];
    const int sys = aux_sys(ax);
    SolverState *st = a.st ? a.st + sys : nullptr;
    if (st && a.check_done && (st->done || (a.skip_on_sconv && st->s_conv))) return;

    // phase 1: coalesced col/val streams, x gathered, one product per entry
    const int nz = rb.j1 - rb.j0;
    const int *__restrict__ colp = a.col + rb.j0;
    const T *__restrict__ valp = a.val + rb.j0;
    const T *__restrict__ x = a.x;
#pragma unroll 4
    for (int k = threadIdx.x; k < nz; k += kThreads) lds[k] = valp[k] * x[colp[k]];
    __syncthreads();

    Acc<T> d[kMaxDots];
    bool write_slot = true;
    int slot = b;
    if (ax.chunk < 0) {
        const int nr = rb.r1 - rb.r0;
        const int lanes = aux_lanes(ax);
        if (lanes <= 1) {
            // order-exact: res = 0; res += p_j in CSR order
            for (int t = threadIdx.x; t < nr; t += kThreads) {
                const int r = rb.r0 + t;
                const int s = a.rowptr[r] - rb.j0, e = a.rowptr[r + 1] - rb.j0;
                T acc = T(0);
                for (int q = s; q < e; ++q) acc += lds[q];
                row_out<T, EPI>(a, r, acc, d);
            }
        } else {
            const int g = threadIdx.x / lanes, l = threadIdx.x & (lanes - 1);
            const int ng = kThreads / lanes;
            for (int t0 = 0; t0 < nr; t0 += ng) {
                const int t = t0 + g;
                T acc = T(0);
                int r = rb.r0 + t;
                if (t < nr) {
                    const int s = a.rowptr[r] - rb.j0, e = a.rowptr[r + 1] - rb.j0;
                    for (int q = s + l; q < e; q += lanes) acc += lds[q];
                }
                for (int o = lanes >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
                if (t < nr && l == 0) row_out<T, EPI>(a, r, acc, d);
            }
        }
    } else {
        // chunk of a long row: workgroup sum, then the last chunk combines
        T acc = T(0);
        for (int k = threadIdx.x; k < nz; k += kThreads) acc += lds[k];
        acc = block_sum1<T>(acc, tred);
        if (threadIdx.x == 0) {
            st_agent(&a.sc.chunk[b], (double)acc);
            __threadfence();
            const unsigned prev = atomicAdd(&a.sc.ctr_row[ax.first], 1u);
            const int last = prev == (unsigned)(ax.nchunks - 1);
            if (last) __hip_atomic_store(&a.sc.ctr_row[ax.first], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_flag = last;
        }
        __syncthreads();
        const int last = s_flag;
        __syncthreads();
        if (last) {
            __threadfence();
            if (threadIdx.x == 0) {
                T total = T(0);
                for (int c = 0; c < ax.nchunks; ++c) total += (T)ld_agent(&a.sc.chunk[ax.first + c]);
                row_out<T, EPI>(a, rb.r0, total, d);
            }
        }
        // the row's dot contribution lives in the slot of its first chunk,
        // written only by the finalizing chunk; other chunks hold zero
        if (last) {
            slot = ax.first;
            if (ax.chunk != 0 && threadIdx.x == 0 && ND > 0) {
#pragma unroll
                for (int q = 0; q < ND; ++q) st_agent(&a.sc.part[(size_t)q * a.nblk + b], 0.0);
            }
        } else if (ax.chunk == 0) {
            write_slot = false;
        }
    }

    if constexpr (ND > 0) {
        double v[ND];
#pragma unroll
        for (int q = 0; q < ND; ++q) v[q] = d[q].get();
        block_sum<ND>(v, red);
        if (threadIdx.x == 0) {
            if (write_slot) {
#pragma unroll
                for (int q = 0; q < ND; ++q) st_agent(&a.sc.part[(size_t)q * a.nblk + slot], v[q]);
            }
        }
        const SysRange sr = a.sys[sys];
        if (arrive_and_reduce<ND>(v, a.sc.part, a.nblk, &a.sc.ctr_sys[sys], sr.b0, sr.b1, red, &s_flag)) {
            if (threadIdx.x == 0 && st) finalize<ND>(st, a.fin, v, a.sc.ndone);
        }
    }
}

// ---- streaming vector kernels --------------------------------------------

template <typename T>
struct alignas(16) Pk {
    static constexpr int N = 16 / sizeof(T);
    T v[N];
};
template <typename T>
__device__ __forceinline__ Pk<T> ldp(const T *p)
{
    return *reinterpret_cast<const Pk<T> *>(p);
}
template <typename T>
__device__ __forceinline__ void stp(T *p, const Pk<T> &x)
{
    *reinterpret_cast<Pk<T> *>(p) = x;
}

// Visit [e0, e1): scalar head up to a 16-byte boundary, 16-byte packs, tail.
// f(i) handles element i; fv(i) handles the pack starting at i.
template <typename T, typename F, typename FV>
__device__ __forceinline__ void visit(int e0, int e1, F f, FV fv)
{
    constexpr int N = Pk<T>::N;
    const int a0 = min(e1, (e0 + N - 1) / N * N);
    const int a1 = max(a0, e1 / N * N);
    if ((int)threadIdx.x < a0 - e0) f(e0 + threadIdx.x);
    for (int i = a0 + (int)threadIdx.x * N; i < a1; i += kThreads * N) fv(i);
    if ((int)threadIdx.x < e1 - a1) f(a1 + threadIdx.x);
}

template <int ND, typename T>
__device__ __forceinline__ void vec_reduce(const VecArgs<T> &a, const VecBlock &vb, Acc<T> *d,
                                           int fin, double *red, int *s_flag)
{
    double v[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) v[q] = d[q].get();
    block_sum<ND>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < ND; ++q) st_agent(&a.sc.part[(size_t)q * a.nvb + blockIdx.x], v[q]);
    }
    const SysRange sr = a.sys[vb.sys];
    if (arrive_and_reduce<ND>(v, a.sc.part, a.nvb, &a.sc.ctr_sys[vb.sys], sr.v0, sr.v1, red, s_flag)) {
        if (threadIdx.x == 0) finalize<ND>(a.st + vb.sys, fin, v, a.sc.ndone);
    }
}

// P = ((P - V*omega)*beta) + R  (bicgstab_cpu.c:210-213; k_vecop_P)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_vec_P(VecArgs<T> a)
{
    const VecBlock vb = a.blocks[blockIdx.x];
    const SolverState *st = a.st + vb.sys;
    if (st->done) return;
    const T om = (T)st->omega, be = (T)st->beta;
    T *__restrict__ P = a.P;
    const T *__restrict__ V = a.V;
    const T *__restrict__ R = a.R;
    visit<T>(
        vb.e0, vb.e1, [&](int i) { P[i] = ((P[i] - V[i] * om) * be) + R[i]; },
        [&](int i) {
            Pk<T> p = ldp(P + i);
            const Pk<T> v = ldp(V + i), r = ldp(R + i);
#pragma unroll
            for (int u = 0; u < Pk<T>::N; ++u) p.v[u] = ((p.v[u] - v.v[u] * om) * be) + r.v[u];
            stp(P + i, p);
        });
}

// S = R - V*alpha, ||S||^2  (bicgstab_cpu.c:222-223; k_vecop_S; stop test of
// bicgstab_gpu.c:231-236)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_vec_S(VecArgs<T> a)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ int s_flag;
    const VecBlock vb = a.blocks[blockIdx.x];
    const SolverState *st = a.st + vb.sys;
    if (st->done) return;
    const T al = (T)st->alpha;
    T *__restrict__ S = a.S;
    const T *__restrict__ V = a.V;
    const T *__restrict__ R = a.R;
    Acc<T> d[1];
    visit<T>(
        vb.e0, vb.e1,
        [&](int i) {
            const T s = R[i] - V[i] * al;
            S[i] = s;
            d[0].add(s, s);
        },
        [&](int i) {
            const Pk<T> v = ldp(V + i), r = ldp(R + i);
            Pk<T> s;
#pragma unroll
            for (int u = 0; u < Pk<T>::N; ++u) {
                s.v[u] = r.v[u] - v.v[u] * al;
                d[0].add(s.v[u], s.v[u]);
            }
            stp(S + i, s);
        });
    vec_reduce<1, T>(a, vb, d, FIN_S, red, &s_flag);
}

// X = (X + P*alpha) + S*omega ; R = S - T*omega ; (R, R0), ||R||^2
// (bicgstab_cpu.c:232-242 and the next iteration's rho, :203; k_vecop_XR).
// After the ||s|| test fired: X = X + P*alpha only (bicgstab_gpu.c:219-236).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_vec_XR(VecArgs<T> a)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ int s_flag;
    const VecBlock vb = a.blocks[blockIdx.x];
    const SolverState *st = a.st + vb.sys;
    if (st->done) return;
    const T al = (T)st->alpha, om = (T)st->omega;
    T *__restrict__ X = a.X;
    T *__restrict__ R = a.R;
    const T *__restrict__ P = a.P;
    const T *__restrict__ S = a.S;
    const T *__restrict__ Tv = a.T_;
    const T *__restrict__ R0 = a.R0;
    Acc<T> d[2];
    if (st->s_conv) {
        visit<T>(
            vb.e0, vb.e1, [&](int i) { X[i] = X[i] + P[i] * al; },
            [&](int i) {
                Pk<T> x = ldp(X + i);
                const Pk<T> p = ldp(P + i);
#pragma unroll
                for (int u = 0; u < Pk<T>::N; ++u) x.v[u] = x.v[u] + p.v[u] * al;
                stp(X + i, x);
            });
    } else {
        visit<T>(
            vb.e0, vb.e1,
            [&](int i) {
                X[i] = (X[i] + P[i] * al) + S[i] * om;
                const T r = S[i] - Tv[i] * om;
                R[i] = r;
                d[0].add(r, R0[i]);
                d[1].add(r, r);
            },
            [&](int i) {
                Pk<T> x = ldp(X + i);
                const Pk<T> p = ldp(P + i), s = ldp(S + i), t = ldp(Tv + i), r0 = ldp(R0 + i);
                Pk<T> r;
#pragma unroll
                for (int u = 0; u < Pk<T>::N; ++u) {
                    x.v[u] = (x.v[u] + p.v[u] * al) + s.v[u] * om;
                    r.v[u] = s.v[u] - t.v[u] * om;
                    d[0].add(r.v[u], r0.v[u]);
                    d[1].add(r.v[u], r.v[u]);
                }
                stp(X + i, x);
                stp(R + i, r);
            });
    }
    vec_reduce<2, T>(a, vb, d, FIN_XR, red, &s_flag);
}

// R0 = R ; P = V = 0  (bicgstab_cpu.c:180-188)
__global__ __launch_bounds__(kThreads) void k_init_vectors(VecArgs<double> a)
{
    const VecBlock vb = a.blocks[blockIdx.x];
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) {
        const_cast<double *>(a.R0)[i] = a.R[i];
        a.P[i] = 0.0;
        a.V[i] = 0.0;
    }
}

// Mixed sweep start (FP32 inner solve of A d = r / ||r||).
__global__ __launch_bounds__(kThreads) void k_ir_start(VecArgs<float> a, const double *r64,
                                                      const SolverState *outer)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ int s_flag;
    const VecBlock vb = a.blocks[blockIdx.x];
    const double nr = sqrt(outer[vb.sys].resid2);
    const double inv = nr > 0.0 ? 1.0 / nr : 0.0;
    Acc<float> d[1];
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) {
        const float r = (float)(r64[i] * inv);
        a.R[i] = r;
        const_cast<float *>(a.R0)[i] = r;
        a.X[i] = 0.f;
        a.P[i] = 0.f;
        a.V[i] = 0.f;
        d[0].add(r, r);
    }
    if (threadIdx.x == 0 && (int)blockIdx.x == a.sys[vb.sys].v0) a.st[vb.sys].scale = nr;
    vec_reduce<1, float>(a, vb, d, FIN_INIT, red, &s_flag);
}

__global__ __launch_bounds__(kThreads) void k_ir_update(VecArgs<float> a, double *x64)
{
    const VecBlock vb = a.blocks[blockIdx.x];
    const double sc = a.st[vb.sys].scale;
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) x64[i] = x64[i] + (double)a.X[i] * sc;
}

__global__ __launch_bounds__(kThreads) void k_dot(VecArgs<double> a, const double *x, const double *y,
                                                 int fin)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ int s_flag;
    const VecBlock vb = a.blocks[blockIdx.x];
    Acc<double> d[1];
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) d[0].add(x[i], y[i]);
    vec_reduce<1, double>(a, vb, d, fin, red, &s_flag);
}

__global__ void k_state_init(SolverState *st, int nsys, double tol, int maxit, int fixed,
                             unsigned *ndone)
{
    for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
        SolverState z{};
        z.rho = 1.0;
        z.alpha = 1.0;
        z.omega = 1.0;
        z.tol = tol;
        z.maxit = maxit;
        z.fixed = fixed;
        z.done = maxit <= 0 ? 1 : 0;
        st[s] = z;
    }
    if (threadIdx.x == 0 && ndone) *ndone = 0;
}

// ---- BLAS-1 and helpers ---------------------------------------------------

__global__ void k_axpy(int n, double a, const double *x, double *y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] = y[i] + a * x[i];
}
__global__ void k_scal(int n, double a, double *x)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] = x[i] * a;
}
__global__ void k_fill_zero(unsigned *p, size_t nwords, unsigned char *tail, int ntail)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}
__global__ void k_copy(size_t n, const double *src, double *dst)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void k_to_f32(int n, const double *src, float *dst)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}
__global__ void k_gather(int n, const double *in, const int *idx, double *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[idx[i]];
}
__global__ void k_scatter(int n, const double *in, const int *idx, double *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[idx[i]] = in[i];
}

// g[j] = -lambda[row(j)] * x[col[j]] over the CSR pattern; one wave per row
// batch, rows of up to 64 entries per pass.
__global__ __launch_bounds__(kThreads) void k_grad_pattern(int m, const int *rowptr, const int *col,
                                                          const double *lambda, const double *x,
                                                          double *g)
{
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * kThreads + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * kThreads) >> 6;
    for (int r = wave; r < m; r += nwaves) {
        const double lr = -lambda[r];
        for (int j = rowptr[r] + lane; j < rowptr[r + 1]; j += 64) g[j] = lr * x[col[j]];
    }
}

// Y[r, c] = sum_j val[j] * X[col[j], c], row-major with ld = k.  `tile`
// lanes share one row (one RHS column each), in CSR order per column, so
// every column equals the single-RHS order-exact sum.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spmm(int m, const int *__restrict__ rowptr,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ val, int k,
                                                  const T *__restrict__ X, T *__restrict__ Y, int tile)
{
    const int lane = threadIdx.x & (tile - 1);
    const int rows_per_block = kThreads / tile;
    const int r = blockIdx.x * rows_per_block + threadIdx.x / tile;
    const int c = blockIdx.y * tile + lane;
    if (r >= m || c >= k) return;
    T acc = T(0);
    const int e = rowptr[r + 1];
    for (int j = rowptr[r]; j < e; ++j) acc += val[j] * X[(size_t)col[j] * k + c];
    Y[(size_t)r * k + c] = acc;
}

// ---- reference-shaped kernels (hhp_dp_kernels.cu:16-74) -------------------

__global__ void k_dp_vecop_P(double *P, const double *V, const double *R, const double *omega,
                             const double *beta, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) P[i] = (P[i] - (*omega) * V[i]) * (*beta) + R[i];
}
__global__ void k_dp_vecop_S(double *S, const double *R, const double *V, const double *neg_alpha,
                             int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) S[i] = R[i] + (*neg_alpha) * V[i];
}
__global__ void k_dp_vecop_XR(double *X, double *R, const double *P, const double *S, const double *T,
                              const double *alpha, const double *omega, const double *neg_omega, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        X[i] += (*alpha) * P[i] + (*omega) * S[i];
        R[i] = S[i] + (*neg_omega) * T[i];
    }
}
__global__ void k_dp_scalar(int which, double *a, double *b, const double *c, const double *d,
                            const double *e)
{
    if (which == 0) {            // beta = (rho_new/rho_old)*(alpha/omega); rho_old = rho_new
        *a = (*c / *b) * (*d / *e);
        *b = *c;
    } else if (which == 1) {     // alpha = rho/rv; neg = -alpha
        const double v = *c / *d;
        *a = v;
        *b = -v;
    } else if (which == 2) {     // omega = ts/tt; neg = -omega
        const double v = *c / *d;
        *a = v;
        *b = -v;
    } else {                     // total = a + b
        *a = *c + *d;
    }
}

inline int grid_for(size_t n, int per_block = kThreads, int cap = 8192)
{
    size_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

template <typename T>
hipError_t spmv_dispatch(const SpmvArgs<T> &a, int epi, int block_nnz, hipStream_t s)
{
    if (a.nblk <= 0) return hipSuccess;
    const size_t shmem = (size_t)block_nnz * sizeof(T);
    switch (epi) {
    case EPI_AXPBY: k_spmv<T, EPI_AXPBY><<<a.nblk, kThreads, shmem, s>>>(a); break;
    case EPI_DOT1: k_spmv<T, EPI_DOT1><<<a.nblk, kThreads, shmem, s>>>(a); break;
    case EPI_DOT2: k_spmv<T, EPI_DOT2><<<a.nblk, kThreads, shmem, s>>>(a); break;
    case EPI_RESID: k_spmv<T, EPI_RESID><<<a.nblk, kThreads, shmem, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_spmv(const SpmvArgs<double> &a, int epi, int block_nnz, hipStream_t s)
{
    return spmv_dispatch<double>(a, epi, block_nnz, s);
}
hipError_t launch_spmv(const SpmvArgs<float> &a, int epi, int block_nnz, hipStream_t s)
{
    return spmv_dispatch<float>(a, epi, block_nnz, s);
}

#define HBS_VEC_LAUNCH(NAME, KER, T)                                   \
    hipError_t NAME(const VecArgs<T> &a, hipStream_t s)                \
    {                                                                  \
        if (a.nvb <= 0) return hipSuccess;                             \
        KER<T><<<a.nvb, kThreads, 0, s>>>(a);                          \
        return hipGetLastError();                                      \
    }
HBS_VEC_LAUNCH(launch_vec_P, k_vec_P, double)
HBS_VEC_LAUNCH(launch_vec_S, k_vec_S, double)
HBS_VEC_LAUNCH(launch_vec_XR, k_vec_XR, double)
HBS_VEC_LAUNCH(launch_vec_P, k_vec_P, float)
HBS_VEC_LAUNCH(launch_vec_S, k_vec_S, float)
HBS_VEC_LAUNCH(launch_vec_XR, k_vec_XR, float)
#undef HBS_VEC_LAUNCH

hipError_t launch_init_vectors(const VecArgs<double> &a, hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_init_vectors<<<a.nvb, kThreads, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_ir_start(const VecArgs<float> &a, const double *r64, const SolverState *outer,
                           hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_ir_start<<<a.nvb, kThreads, 0, s>>>(a, r64, outer);
    return hipGetLastError();
}

hipError_t launch_ir_update(const VecArgs<float> &a, double *x64, hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_ir_update<<<a.nvb, kThreads, 0, s>>>(a, x64);
    return hipGetLastError();
}

hipError_t launch_dot(const VecArgs<double> &a, const double *x, const double *y, int fin,
                      hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_dot<<<a.nvb, kThreads, 0, s>>>(a, x, y, fin);
    return hipGetLastError();
}

hipError_t launch_state_init(SolverState *st, int nsys, double tol, int maxit, int fixed,
                             unsigned *ndone, hipStream_t s)
{
    k_state_init<<<1, kThreads, 0, s>>>(st, nsys, tol, maxit, fixed, ndone);
    return hipGetLastError();
}

hipError_t launch_axpy_f64(int n, double a, const double *x, double *y, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_axpy<<<grid_for(n), kThreads, 0, s>>>(n, a, x, y);
    return hipGetLastError();
}
hipError_t launch_scal_f64(int n, double a, double *x, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_scal<<<grid_for(n), kThreads, 0, s>>>(n, a, x);
    return hipGetLastError();
}
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return hipSuccess;
    const size_t nwords = bytes / 4;
    const int ntail = (int)(bytes % 4);
    k_fill_zero<<<grid_for(nwords > 0 ? nwords : 1), kThreads, 0, s>>>(
        (unsigned *)p, nwords, (unsigned char *)p + nwords * 4, ntail);
    return hipGetLastError();
}
hipError_t launch_copy_f64(size_t n, const double *src, double *dst, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_copy<<<grid_for(n), kThreads, 0, s>>>(n, src, dst);
    return hipGetLastError();
}
hipError_t launch_to_f32(int n, const double *src, float *dst, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_to_f32<<<grid_for(n), kThreads, 0, s>>>(n, src, dst);
    return hipGetLastError();
}
hipError_t launch_gather_f64(int n, const double *in, const int *idx, double *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_gather<<<grid_for(n), kThreads, 0, s>>>(n, in, idx, out);
    return hipGetLastError();
}
hipError_t launch_scatter_f64(int n, const double *in, const int *idx, double *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_scatter<<<grid_for(n), kThreads, 0, s>>>(n, in, idx, out);
    return hipGetLastError();
}
hipError_t launch_grad_pattern(int m, const int *rowptr, const int *col, const double *lambda,
                               const double *x, double *g, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    k_grad_pattern<<<grid_for((size_t)m * 64), kThreads, 0, s>>>(m, rowptr, col, lambda, x, g);
    return hipGetLastError();
}
hipError_t launch_spmm_f64(int m, const int *rowptr, const int *col, const double *val, int k,
                           const double *X, double *Y, int tile, hipStream_t s)
{
    if (m <= 0 || k <= 0) return hipSuccess;
    const int rpb = kThreads / tile;
    dim3 grid((m + rpb - 1) / rpb, (k + tile - 1) / tile);
    k_spmm<double><<<grid, kThreads, 0, s>>>(m, rowptr, col, val, k, X, Y, tile);
    return hipGetLastError();
}
hipError_t launch_spmm_f32(int m, const int *rowptr, const int *col, const float *val, int k,
                           const float *X, float *Y, int tile, hipStream_t s)
{
    if (m <= 0 || k <= 0) return hipSuccess;
    const int rpb = kThreads / tile;
    dim3 grid((m + rpb - 1) / rpb, (k + tile - 1) / tile);
    k_spmm<float><<<grid, kThreads, 0, s>>>(m, rowptr, col, val, k, X, Y, tile);
    return hipGetLastError();
}

hipError_t launch_dp_vecop_P(double *P, const double *V, const double *R, const double *omega,
                             const double *beta, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_dp_vecop_P<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(P, V, R, omega, beta, n);
    return hipGetLastError();
}
hipError_t launch_dp_vecop_S(double *S, const double *R, const double *V, const double *neg_alpha,
                             int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_dp_vecop_S<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(S, R, V, neg_alpha, n);
    return hipGetLastError();
}
hipError_t launch_dp_vecop_XR(double *X, double *R, const double *P, const double *S,
                              const double *T, const double *alpha, const double *omega,
                              const double *neg_omega, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_dp_vecop_XR<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(X, R, P, S, T, alpha, omega,
                                                                     neg_omega, n);
    return hipGetLastError();
}
hipError_t launch_dp_scalar(int which, double *a, double *b, const double *c, const double *d,
                            const double *e, hipStream_t s)
{
    k_dp_scalar<<<1, 1, 0, s>>>(which, a, b, c, d, e);
    return hipGetLastError();
}

}  // namespace hbs
