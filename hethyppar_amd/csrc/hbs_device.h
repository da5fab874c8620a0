// hbs_device.h -- device helpers shared by the kernel files: XCD-aware
// workgroup placement, deterministic wave / workgroup / grid sums, the FP32
// compensated dot accumulator, and the BiCGStab scalar updates.
#pragma once

#include <hip/hip_runtime.h>

#include "hbs_internal.h"

namespace hbs {
namespace {

constexpr int kXcds = 8;   // MI355X: 8 XCDs, workgroups dispatched round robin

// Give each XCD a contiguous range of logical workgroups so neighbouring row
// blocks (which gather neighbouring x entries) share one XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int n)
{
    const int q = n / kXcds, r = n % kXcds;
    const int x = bid % kXcds, k = bid / kXcds;
    return x * q + (x < r ? x : r) + k;
}

// Per-thread dot accumulators.  FP64: sequential sum of rounded products.
// FP32: Dot2 of Ogita, Rump and Oishi (2005) -- TwoProduct by FMA and Knuth's
// TwoSum, error terms kept in a second float -- handed on as FP64 to the
// workgroup and grid sums.
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

// Workgroup sum of ND values over NW waves; every thread gets the totals.
// Fixed order: xor butterfly inside each wave, then waves 0..NW-1 in
// sequence.  red holds ND * NW doubles.
template <int ND, int NW = kWaves>
__device__ __forceinline__ void block_sum(double (&v)[ND], double *red)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = wave_sum(v[d]);
    if (l == 0) {
#pragma unroll
        for (int d = 0; d < ND; ++d) red[d * NW + w] = v[d];
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double t = red[d * NW];
#pragma unroll
        for (int k = 1; k < NW; ++k) t += red[d * NW + k];
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

// Sum of partials [p0, p1) (stride nparts between the ND rows), fixed order:
// thread t < 256 adds p0 + t, p0 + t + 256, ... in turn, then block_sum.  The
// loads of a span of kSpan partials per thread are all issued before the
// first add, so a consumer's prologue costs one L2 round trip per span rather
// than one per partial.  In workgroups of NW > 4 waves the threads past 256
// add nothing, so every workgroup size derives bit-identical sums.
constexpr int kSpan = 16;
template <int ND, int NW = kWaves>
__device__ __forceinline__ void reduce_parts(const double *part, int nparts, int p0, int p1, double (&v)[ND],
                                             double *red)
{
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = 0.0;
    const int b0 = threadIdx.x < kThreads ? p0 + (int)threadIdx.x : p1;
    for (int b = b0; b < p1; b += kSpan * kThreads) {
        double buf[ND][kSpan];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
#pragma unroll
            for (int k = 0; k < kSpan; ++k) {
                const int i = b + k * kThreads;
                buf[d][k] = i < p1 ? part[(size_t)d * nparts + i] : 0.0;
            }
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) {
#pragma unroll
            for (int k = 0; k < kSpan; ++k) v[d] += buf[d][k];
        }
    }
    block_sum<ND, NW>(v, red);
}

// One row of A times one column of a row-major block: sum over the row's
// entries in CSR order of val[j] * x[col[j] * ld] (x already offset to the
// column).  Indices and values of 8 entries are loaded, then their 8
// gathers, before the ordered adds, so the loads of a row overlap.
template <typename T>
__device__ __forceinline__ T row_dot(const int *__restrict__ rowptr, const int *__restrict__ col,
                                     const T *__restrict__ val, const T *__restrict__ x, int ld, int r)
{
    constexpr int kChunk = 8;
    const int j0 = rowptr[r], j1 = rowptr[r + 1];
    T acc = T(0);
    for (int j = j0; j < j1; j += kChunk) {
        int cc[kChunk];
        T vv[kChunk], xx[kChunk];
#pragma unroll
        for (int i = 0; i < kChunk; ++i) {
            const bool ok = j + i < j1;
            cc[i] = ok ? col[j + i] : 0;
            vv[i] = ok ? val[j + i] : T(0);
        }
#pragma unroll
        for (int i = 0; i < kChunk; ++i) xx[i] = (j + i < j1) ? x[(size_t)cc[i] * ld] : T(0);
#pragma unroll
        for (int i = 0; i < kChunk; ++i)
            if (j + i < j1) acc += vv[i] * xx[i];
    }
    return acc;
}

// RB rows at once (rows r + q * step, q < RB, those below r1): the same
// ordered sums as row_dot, with every row's loads of a chunk in flight
// together.  Rows at or past r1 give 0.
template <typename T, int RB>
__device__ __forceinline__ void rows_dot(const int *__restrict__ rowptr, const int *__restrict__ col,
                                         const T *__restrict__ val, const T *__restrict__ x, int ld, int r,
                                         int step, int r1, T (&acc)[RB])
{
    constexpr int kChunk = 8;
    int jb[RB], je[RB], len = 0;
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        const int rq = r + q * step;
        jb[q] = rq < r1 ? rowptr[rq] : 0;
        je[q] = rq < r1 ? rowptr[rq + 1] : 0;
        len = max(len, je[q] - jb[q]);
        acc[q] = T(0);
    }
    for (int t = 0; t < len; t += kChunk) {
        int cc[RB][kChunk];
        T vv[RB][kChunk], xx[RB][kChunk];
#pragma unroll
        for (int q = 0; q < RB; ++q) {
#pragma unroll
            for (int i = 0; i < kChunk; ++i) {
                const int j = jb[q] + t + i;
                const bool ok = j < je[q];
                cc[q][i] = ok ? col[j] : 0;
                vv[q][i] = ok ? val[j] : T(0);
            }
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
#pragma unroll
            for (int i = 0; i < kChunk; ++i)
                xx[q][i] = (jb[q] + t + i < je[q]) ? x[(size_t)cc[q][i] * ld] : T(0);
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
#pragma unroll
            for (int i = 0; i < kChunk; ++i)
                if (jb[q] + t + i < je[q]) acc[q] += vv[q][i] * xx[q][i];
        }
    }
}

// One step of one wave over 64 consecutive columns of row-major blocks:
// rows r .. r + RB - 1 (those at or past `end` are empty) of X (column 0 of
// row 0 at xb, leading dimension ld), lane l on column lc, each row summed
// over its entries in CSR order, the order of row_dot.  r must be
// wave-uniform.  The row pointers, column indices and values are then the
// same for the whole wave, so they come through the scalar cache in chunks
// of kRowPad entries (the CSR arrays carry kRowPad entries of padding past
// nnz); each entry is one 64-lane gather of a contiguous stretch of an X
// row.  The first kRowPad entries of all RB rows are loaded, then gathered,
// before any is summed: gathers past a row's end re-read the row's first
// column (or column 0), and lanes with lane_ok false read column 0, so
// every load is in bounds and none waits on a branch.  out(q, acc) receives
// row r + q.
template <typename T, int RB, typename Out>
__device__ __forceinline__ void wave_rows64(const int *__restrict__ rowptr, const int *__restrict__ col,
                                            const T *__restrict__ val, const T *__restrict__ xb, int ld, int lc,
                                            int r, int end, bool lane_ok, Out &&out)
{
    constexpr int C = kRowPad;
    const T *__restrict__ xl = xb + (lane_ok ? lc : 0);
    int p[RB + 1];
#pragma unroll
    for (int q = 0; q <= RB; ++q) p[q] = rowptr[min(r + q, end)];
    int cc[RB][C];
    T vv[RB][C];
#pragma unroll
    for (int q = 0; q < RB; ++q) {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            cc[q][i] = col[p[q] + i];
            vv[q][i] = val[p[q] + i];
        }
    }
    T xv[RB][C];
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        const int c0 = p[q] < p[q + 1] ? cc[q][0] : 0;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const int cj = p[q] + i < p[q + 1] ? cc[q][i] : c0;
            xv[q][i] = xl[(size_t)cj * ld];
        }
    }
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        const int c0 = p[q] < p[q + 1] ? cc[q][0] : 0;
        T acc = T(0);
#pragma unroll
        for (int i = 0; i < C; ++i)
            if (p[q] + i < p[q + 1]) acc += vv[q][i] * xv[q][i];
        for (int j0 = p[q] + C; j0 < p[q + 1]; j0 += C) {
            T cv[C], cx[C];
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const int cj = j0 + i < p[q + 1] ? col[j0 + i] : c0;
                cv[i] = val[j0 + i];
                cx[i] = xl[(size_t)cj * ld];
            }
#pragma unroll
            for (int i = 0; i < C; ++i)
                if (j0 + i < p[q + 1]) acc += cv[i] * cx[i];
        }
        out(q, acc);
    }
}

__device__ __forceinline__ void set_done(SolverState &s, int stop)
{
    s.done = 1;
    s.stop = stop;
}

// Scalar updates (identical in every workgroup).  Order of operations of
// bicgstab_cpu.c:203-229; stop tests of bicgstab_gpu.c:148-150,231-236,289-293.
__device__ void apply_fin(SolverState &s, int fin, double d0, double d1)
{
    switch (fin) {
    case FIN_ALPHA:
        s.rv = d0;
        s.alpha = s.rho / d0;
        if (d0 == 0.0 || !isfinite(s.alpha)) set_done(s, 4);
        break;
    case FIN_OMEGA:
        s.ts = d0;
        s.tt = d1;
        s.omega = d0 / d1;
        if (d1 == 0.0 || !isfinite(s.omega)) set_done(s, 4);
        break;
    case FIN_XR:
        // end of iteration `iter` (or, with iter == -1, the initial residual)
        s.pending = 0;
        s.iter += 1;
        if (s.iter == 0) {
            s.rr0 = d1;
            s.thr = s.tol * sqrt(d1);
        }
        if (s.s_conv) { set_done(s, 1); break; }
        s.rho_next = d0;
        s.rr = d1;
        if (s.iter > 0 && !s.fixed && sqrt(d1) < s.thr) { set_done(s, 2); break; }
        if (d1 == 0.0) { set_done(s, 2); break; }
        if (s.iter >= s.maxit) { set_done(s, 0); break; }
        if (d0 == 0.0 || !isfinite(d0)) { set_done(s, 4); break; }
        s.beta = (d0 / s.rho) * (s.alpha / s.omega);
        s.rho = d0;
        break;
    case FIN_S:
        s.ss = d0;
        if (!s.fixed && sqrt(d0) < s.thr) s.s_conv = 1;
        break;
    case FIN_RESID:
        s.resid2 = d0;
        break;
    case FIN_STORE:
        s.out = d0;
        break;
    case FIN_RESIDB:
        s.resid2 = d0;
        s.out = d1;
        break;
    default:
        break;
    }
}

}  // namespace
}  // namespace hbs
