// hbs_device.h -- device helpers shared by the kernel files: XCD-aware
// workgroup placement, deterministic wave / workgroup / grid sums, the FP32
// compensated dot accumulator, and the BiCGStab scalar updates.
#pragma once

#include <hip/hip_runtime.h>

#include "hbs_internal.h"

namespace hbs {
namespace {


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

// Positions [p0, p1) of XCD xcd in the traversal rr of m rows.
__device__ __forceinline__ void xcd_positions(const RowRuns &rr, int m, int xcd, int &p0, int &p1)
{
    if (rr.run) {
        p0 = (int)((long long)rr.nruns * xcd / kXcds) * rr.len;
        p1 = (int)((long long)rr.nruns * (xcd + 1) / kXcds) * rr.len;
    } else {
        p0 = (int)((long long)m * xcd / kXcds);
        p1 = (int)((long long)m * (xcd + 1) / kXcds);
    }
}

// Row at position p of the traversal rr.
__device__ __forceinline__ int pos_row(const RowRuns &rr, int p)
{
    if (!rr.run) return p;
    const int k = p / rr.len;
    return rr.run[k] + (p - k * rr.len);
}

// RB rows at once (the rows at positions p + q * step of the traversal rr,
// q < RB, those below p1): the same ordered sums as row_dot, with every
// row's loads of a chunk in flight together.  Positions at or past p1 give
// 0.
template <typename T, int RB>
__device__ __forceinline__ void rows_dot(const int *__restrict__ rowptr, const int *__restrict__ col,
                                         const T *__restrict__ val, const T *__restrict__ x, int ld,
                                         const RowRuns &rr, int p, int step, int p1, T (&acc)[RB])
{
    constexpr int kChunk = 8;
    int jb[RB], je[RB], len = 0;
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        const int pq = p + q * step;
        const int rq = pq < p1 ? pos_row(rr, pq) : 0;
        jb[q] = pq < p1 ? rowptr[rq] : 0;
        je[q] = pq < p1 ? rowptr[rq + 1] : 0;
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

// A wave's sweep of a 64-column SpMM over steps of RB consecutive rows,
// software-pipelined.  Step k of the wave is at position p0 + t * pstride
// of the traversal rr (below p1; a step never leaves its run), where t = k
// (static dealing) or, with a step counter ctr, the counter's next ticket
// (dynamic dealing, sweep_ctr: all the XCD's waves then work on
// neighbouring steps, so the X rows they share stay in its L2, while
// statically dealt waves drift apart over a long sweep).  While step k
// gathers X, the entries (column, value) of step k + 1, the row pointers
// of step k + 2 and the run of step k + 3 are loading, and the ticket of
// step k + 4 is taken once the gathers are in flight.  Each of these uses
// only what the previous step loaded, so a step waits for one round trip
// (its gathers; lane 0 takes the ticket behind a branch, whose join waits
// for every memory operation in flight -- the gathers' wait, which the sums
// need anyway).  The entries are vector loads, one per lane, and reach
// every lane by readlane; each entry is one 64-lane gather of a contiguous
// stretch of an X row.  Every row is summed over its entries in CSR order
// (the order of row_dot).  Rows of at most kRowPad entries take one round
// of gathers (past a row's end the gathers read the next entries' columns,
// in bounds thanks to the CSR arrays' kRowPad entries of padding, and are
// not added); longer rows continue with chunks of kRowPad entries through
// the scalar cache.  before(ru, nr) runs as the step's gathers are issued;
// out(q, row, acc) receives row row = ru + q.  p0 and pstride are
// wave-uniform.
template <typename T, int RB, typename Before, typename Out>
__device__ __forceinline__ void wave_sweep64(const int *__restrict__ rowptr, const int *__restrict__ col,
                                             const T *__restrict__ val, const T *__restrict__ xb, int ld, int lc,
                                             bool lane_ok, const RowRuns &rr, int p0, int p1, int pstride,
                                             int *ctr, Before &&before, Out &&out)
{
    constexpr int C = kRowPad;
    static_assert(RB * C <= 64, "one entry per lane");
    const int lane = (int)threadIdx.x & 63;
    const T *__restrict__ xl = xb + (lane_ok ? lc : 0);
    int nt = 0;
    // the wave's next ticket: in lane 0 of the result when dealt by ctr
    auto take = [&]() -> int {
        if (ctr) {
            int t = 0;
            if (lane == 0) t = atomicAdd(ctr, 1);
            return t;
        }
        return nt++;
    };
    auto pos_of = [&](int tk) { return p0 + __builtin_amdgcn_readlane(tk, 0) * pstride; };
    // The run of the step at position p (clamped into the sweep): its first
    // row's run start, loaded without a branch -- without runs the load
    // reads rowptr and the value is ignored.
    const int rlen = rr.run ? rr.len : 1;
    const int *rtab = rr.run ? rr.run : rowptr;
    auto run_of = [&](int p) { return rtab[max(min(p, p1 - 1), 0) / rlen]; };
    // rows of the step at position p with run start b: first row and count
    // (0 past the end, where ru is a row of the last position)
    auto rows_at = [&](int p, int b, int &ru, int &nr) {
        const int pc = max(min(p, p1 - 1), 0);
        ru = rr.run ? b + (pc - (pc / rlen) * rlen) : pc;
        nr = p < p1 ? min(RB, p1 - p) : 0;
    };
    // lane l <= RB: rowptr[ru + min(l, nr)]
    auto load_rp = [&](int ru, int nr) { return rowptr[ru + min(lane, nr)]; };
    // lane l < RB * C: entry rp[l / C] + l % C (row pointers in vrp)
    auto load_cv = [&](int vrp, int &vc, T &vv) {
        const int q = min(lane / C, RB - 1);
        const int j = __shfl(vrp, q, 64) + (lane & (C - 1));
        vc = col[j];
        vv = val[j];
    };
    // prologue: steps 0..3 positioned, runs of 0..2, row pointers of 0 and
    // 1, entries of 0; the ticket of step 4 taken
    int p = pos_of(take()), pn = pos_of(take()), pn2 = pos_of(take());
    int tk3 = take();
    int b = run_of(p), bn = run_of(pn), bn2 = run_of(pn2);
    int ru, nr, ru1, nr1;
    rows_at(p, b, ru, nr);
    rows_at(pn, bn, ru1, nr1);
    int vrp = load_rp(ru, nr);
    int vc;
    T vv;
    load_cv(vrp, vc, vv);
    int vrp1 = load_rp(ru1, nr1);
    int tk4 = take();
    while (p < p1) {
        int rp[RB + 1];
#pragma unroll
        for (int q = 0; q <= RB; ++q) rp[q] = __builtin_amdgcn_readlane(vrp, q);
        // entries of step k + 1, row pointers of k + 2, run of k + 3
        int vc1;
        T vv1;
        load_cv(vrp1, vc1, vv1);
        int ru2, nr2;
        rows_at(pn2, bn2, ru2, nr2);
        const int vrp2 = load_rp(ru2, nr2);
        const int pn3 = pos_of(tk3);
        const int bn3 = run_of(pn3);
        before(ru, nr);
        T xv[RB][C], ev[RB][C];
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            // gathers past the row's entries (wave-uniform) are not issued
            const int lq_len = q < nr ? rp[q + 1] - rp[q] : 0;
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const int cj = __builtin_amdgcn_readlane(vc, q * C + i);
                xv[q][i] = i < lq_len ? xl[(size_t)cj * ld] : T(0);
                if constexpr (sizeof(T) == 8) {
                    const long long bits = __builtin_bit_cast(long long, vv);
                    const int lo = __builtin_amdgcn_readlane((int)bits, q * C + i);
                    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), q * C + i);
                    ev[q][i] = __builtin_bit_cast(T, ((long long)hi << 32) | (unsigned)lo);
                } else {
                    ev[q][i] = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vv), q * C + i));
                }
            }
        }
        // the ticket of step k + 5, behind the gathers
        tk3 = tk4;
        tk4 = take();
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int len = q < nr ? rp[q + 1] - rp[q] : 0;
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < C; ++i)
                if (i < len) acc += ev[q][i] * xv[q][i];
            for (int j0 = rp[q] + C; j0 < rp[q] + len; j0 += C) {
                T cv[C], cx[C];
#pragma unroll
                for (int i = 0; i < C; ++i) {
                    cv[i] = val[j0 + i];
                    cx[i] = xl[(size_t)col[j0 + i] * ld];
                }
#pragma unroll
                for (int i = 0; i < C; ++i)
                    if (j0 + i < rp[q] + len) acc += cv[i] * cx[i];
            }
            if (q < nr) out(q, ru + q, acc);
        }
        p = pn;
        pn = pn2;
        pn2 = pn3;
        bn2 = bn3;
        ru = ru1;
        nr = nr1;
        ru1 = ru2;
        nr1 = nr2;
        vrp = vrp1;
        vrp1 = vrp2;
        vc = vc1;
        vv = vv1;
    }
}

// 16 bytes of a row-major block row: 2 FP64 or 4 FP32 columns.
template <typename T>
struct alignas(16) Cols16 {
    static constexpr int N = 16 / sizeof(T);
    T v[N];
};

// wave_sweep64 with 16-byte gathers: a wave's step is G = 64 / LPR rows, LPR
// = 64 * sizeof(T) / 16 lanes per row (32 FP64, 16 FP32), lane l on row
// l / LPR and columns (l % LPR) * N .. + N - 1 of the 64.  One gather
// instruction then covers one entry of every row of the step (G x 512 or
// G x 256 bytes), so a step issues kRowPad gathers instead of G * kRowPad:
// the vector-memory pipeline, not HBM, bounds the scalar-lane form.  Needs
// ld * sizeof(T), ncols * sizeof(T) (the tile's columns, ncols <= 64) and
// xb, the output, 16-byte aligned.  Each column's sums are the ordered sums
// of row_dot.  before(ru, nr) runs ahead of a step's gathers (its first row
// and row count), for loads the step's epilogue uses.  out(q, row, acc, cc)
// receives the lane's N column sums (columns cc .. cc + N - 1 of the tile)
// of row row = ru + q (q = its row in the step).
template <typename T, typename Before, typename Out>
__device__ __forceinline__ void wave_sweep64v(const int *__restrict__ rowptr, const int *__restrict__ col,
                                              const T *__restrict__ val, const T *__restrict__ xb, int ld,
                                              int ncols, const RowRuns &rr, int p0, int p1, int pstride,
                                              int *ctr, Before &&before, Out &&out)
{
    constexpr int C = kRowPad;
    constexpr int N = Cols16<T>::N;
    constexpr int LPR = 64 / N;
    constexpr int RB = 64 / LPR;
    static_assert(RB * C <= 64, "one entry per lane");
    const int lane = (int)threadIdx.x & 63;
    const int lq = lane / LPR;                 // this lane's row of the step
    const int cc = (lane % LPR) * N;           // and its first column
    // lanes past the tile's ncols columns gather its first columns (in bounds)
    const T *__restrict__ xl = xb + (cc < ncols ? cc : 0);
    int nt = 0;
    auto take = [&]() -> int {
        if (ctr) {
            int t = 0;
            if (lane == 0) t = atomicAdd(ctr, 1);
            return t;
        }
        return nt++;
    };
    auto pos_of = [&](int tk) { return p0 + __builtin_amdgcn_readlane(tk, 0) * pstride; };
    const int rlen = rr.run ? rr.len : 1;
    const int *rtab = rr.run ? rr.run : rowptr;
    auto run_of = [&](int p) { return rtab[max(min(p, p1 - 1), 0) / rlen]; };
    auto rows_at = [&](int p, int b, int &ru, int &nr) {
        const int pc = max(min(p, p1 - 1), 0);
        ru = rr.run ? b + (pc - (pc / rlen) * rlen) : pc;
        nr = p < p1 ? min(RB, p1 - p) : 0;
    };
    auto load_rp = [&](int ru, int nr) { return rowptr[ru + min(lane, nr)]; };
    auto load_cv = [&](int vrp, int &vc, T &vv) {
        const int q = min(lane / C, RB - 1);
        const int j = __shfl(vrp, q, 64) + (lane & (C - 1));
        vc = col[j];
        vv = val[j];
    };
    int p = pos_of(take()), pn = pos_of(take()), pn2 = pos_of(take());
    int tk3 = take();
    int b = run_of(p), bn = run_of(pn), bn2 = run_of(pn2);
    int ru, nr, ru1, nr1;
    rows_at(p, b, ru, nr);
    rows_at(pn, bn, ru1, nr1);
    int vrp = load_rp(ru, nr);
    int vc;
    T vv;
    load_cv(vrp, vc, vv);
    int vrp1 = load_rp(ru1, nr1);
    int tk4 = take();
    while (p < p1) {
        // this lane's row: its entries' start and count
        int rpq = __builtin_amdgcn_readlane(vrp, 0), rpe = __builtin_amdgcn_readlane(vrp, 1);
#pragma unroll
        for (int q = 1; q < RB; ++q) {
            const int a0 = __builtin_amdgcn_readlane(vrp, q), a1 = __builtin_amdgcn_readlane(vrp, q + 1);
            rpq = lq == q ? a0 : rpq;
            rpe = lq == q ? a1 : rpe;
        }
        const int len = lq < nr ? rpe - rpq : 0;
        // the step's longest row (wave-uniform): gathers past it are not
        // issued (a 7-point row leaves the 8th of a chunk to padding), since
        // the vector-memory pipeline, not HBM, bounds this sweep
        int lmax = 0;
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int lq_len = __builtin_amdgcn_readlane(vrp, q + 1) - __builtin_amdgcn_readlane(vrp, q);
            lmax = q < nr ? max(lmax, lq_len) : lmax;
        }
        int vc1;
        T vv1;
        load_cv(vrp1, vc1, vv1);
        int ru2, nr2;
        rows_at(pn2, bn2, ru2, nr2);
        const int vrp2 = load_rp(ru2, nr2);
        const int pn3 = pos_of(tk3);
        const int bn3 = run_of(pn3);
        before(ru, nr);
        Cols16<T> xv[C];
        T ev[C];
#pragma unroll
        for (int i = 0; i < C; ++i) {
            int cj = __builtin_amdgcn_readlane(vc, i);
            T e = T(0);
            if constexpr (sizeof(T) == 8) {
                const long long bits = __builtin_bit_cast(long long, vv);
                e = __builtin_bit_cast(T, ((long long)__builtin_amdgcn_readlane((int)(bits >> 32), i) << 32) |
                                              (unsigned)__builtin_amdgcn_readlane((int)bits, i));
            } else {
                e = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vv), i));
            }
#pragma unroll
            for (int q = 1; q < RB; ++q) {
                const int cq = __builtin_amdgcn_readlane(vc, q * C + i);
                T eq;
                if constexpr (sizeof(T) == 8) {
                    const long long bits = __builtin_bit_cast(long long, vv);
                    eq = __builtin_bit_cast(
                        T, ((long long)__builtin_amdgcn_readlane((int)(bits >> 32), q * C + i) << 32) |
                               (unsigned)__builtin_amdgcn_readlane((int)bits, q * C + i));
                } else {
                    eq = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vv), q * C + i));
                }
                cj = lq == q ? cq : cj;
                e = lq == q ? eq : e;
            }
            if (i < lmax) xv[i] = *reinterpret_cast<const Cols16<T> *>(xl + (size_t)cj * ld);
            else
#pragma unroll
                for (int u = 0; u < N; ++u) xv[i].v[u] = T(0);
            ev[i] = e;
        }
        tk3 = tk4;
        tk4 = take();
        Cols16<T> acc;
#pragma unroll
        for (int e = 0; e < N; ++e) acc.v[e] = T(0);
#pragma unroll
        for (int i = 0; i < C; ++i)
            if (i < len) {
#pragma unroll
                for (int e = 0; e < N; ++e) acc.v[e] += ev[i] * xv[i].v[e];
            }
        // rows longer than kRowPad: the rest of this lane's row, in chunks
        for (int j0 = rpq + C; j0 < rpq + len; j0 += C) {
            T cv[C];
            Cols16<T> cx[C];
#pragma unroll
            for (int i = 0; i < C; ++i) {
                cv[i] = val[j0 + i];
                cx[i] = *reinterpret_cast<const Cols16<T> *>(xl + (size_t)col[j0 + i] * ld);
            }
#pragma unroll
            for (int i = 0; i < C; ++i)
                if (j0 + i < rpq + len) {
#pragma unroll
                    for (int e = 0; e < N; ++e) acc.v[e] += cv[i] * cx[i].v[e];
                }
        }
        if (lq < nr) out(lq, ru + lq, acc, cc);
        p = pn;
        pn = pn2;
        pn2 = pn3;
        bn2 = bn3;
        ru = ru1;
        nr = nr1;
        ru1 = ru2;
        nr1 = nr2;
        vrp = vrp1;
        vrp1 = vrp2;
        vc = vc1;
        vv = vv1;
    }
}

// Step counters of one XCD (and column tile) for wave wv of the XCD's nw:
// with k = min(per_xcd, kCtrPerXcd, nw) counters in use, counter j deals steps j,
// j + k, j + 2k, ... of the XCD's share to waves wv % k == j, so the wave
// sweeps positions p0 + j * RB + t * k * RB for its tickets t.  Spreading
// the XCD's steps over several counters (each in its own cache line) keeps
// the same-address atomics from serialising in the L2, while each
// counter's waves still work on neighbouring steps.  Every counter in use
// has a wave (k <= nw).
struct SweepCtr {
    int *set;      // kCtrSet counters, kCtrStride ints apart (null: fixed stride)
    int *mine;     // this wave's step counter
    int k;         // counters in use
};
__device__ __forceinline__ SweepCtr sweep_ctr(int *set, int wv, int nw, int per_xcd)
{
    if (!set) return SweepCtr{nullptr, nullptr, 0};
    const int k = max(1, min(min(kCtrPerXcd, per_xcd), nw));
    return SweepCtr{set, set + (wv % k) * kCtrStride, k};
}
// First position and stride of wave wv's sweep (RB rows per step): dealt by
// sc, or every nw-th step from step wv.
__device__ __forceinline__ void sweep_span(const SweepCtr &sc, int wv, int nw, int x0, int rb, int &p0, int &ps)
{
    p0 = x0 + (sc.set ? wv % sc.k : wv) * rb;
    ps = (sc.set ? sc.k : nw) * rb;
}

// End of a dynamically dealt sweep: every wave of the XCD's launch calls it
// once, after its last ticket; the last of the nw waves to arrive returns
// the step counters and the arrival count to 0 for the next launch.  A wave
// first waits for all its memory operations, so its tickets are taken
// before it arrives.  The counters serve one XCD (and one column tile), so
// their atomics stay in that XCD's L2.
__device__ __forceinline__ void sweep_done(const SweepCtr &sc, int nw)
{
    if (!sc.set) return;
    __builtin_amdgcn_s_waitcnt(0);
    if ((threadIdx.x & 63) == 0) {
        int *arrived = sc.set + kCtrPerXcd * kCtrStride;
        if (atomicAdd(arrived, 1) == nw - 1) {
            for (int k = 0; k < kCtrPerXcd; ++k) atomicExch(sc.set + k * kCtrStride, 0);
            atomicExch(arrived, 0);
        }
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
