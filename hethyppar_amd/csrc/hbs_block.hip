// hbs_block.hip -- BiCGStab on k right-hand sides at once: one independent
// BiCGStab per column, every column sharing each pass over A.
//
// The n x k blocks are row-major, so the kc <= 64 columns of a group sit
// side by side in a wave: lane l works on column l % kp of row l / kp of the
// wave's pass.  The SpMM reads each row's col/val once for all columns and
// gathers whole x rows (kc contiguous values); every column's sum runs over
// the row in CSR order, the same order as the single-vector serial sum.
// Each column keeps its own SolverState; the scalars are reduced by a small
// kernel with one workgroup per column (k_bfin), in a fixed order.
#include <hip/hip_runtime.h>

#include "hbs_device.h"
#include "hbs_internal.h"

namespace hbs {
namespace {

struct LaneMap {
    int c;        // column of the group
    int ro;       // row offset inside the wave's pass
    int rstep;    // rows per workgroup pass
    bool col_ok;  // c < kc
};

__device__ __forceinline__ LaneMap lane_map(const BlockArgs &a)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rpw = 64 / a.kp;
    LaneMap m;
    m.c = l & (a.kp - 1);
    m.ro = w * rpw + l / a.kp;
    m.rstep = kWaves * rpw;
    m.col_ok = m.c < a.kc;
    return m;
}

// Positions of this workgroup in the traversal rr: its XCD's contiguous
// share, dealt in tiles of `tile` positions round robin over the XCD's
// workgroups, so that the workgroups of one XCD work on neighbouring rows
// at the same time and the x rows they gather stay in that XCD's L2.
// ngroups is a multiple of 8.  The streaming kernels take the rows in order
// (rr without runs).
struct RowSweep {
    int first, end, stride;
};
__device__ __forceinline__ RowSweep row_sweep(const RowRuns &rr, int n, int ngroups, int tile)
{
    const int xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds;
    const int per = ngroups / kXcds;
    int p0, p1;
    xcd_positions(rr, n, xcd, p0, p1);
    return RowSweep{p0 + wi * tile, p1, per * tile};
}

// Column sums of the waves' column partials red[d][w][c] -> part[d][c][p],
// waves 0..3 in order.
template <int ND>
__device__ __forceinline__ void waves_to_partials(const BlockArgs &a, const double *red, int p)
{
    __syncthreads();
    if ((int)threadIdx.x < a.kc) {
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double t = red[(d * kWaves) * 64 + threadIdx.x];
#pragma unroll
            for (int k = 1; k < kWaves; ++k) t += red[(d * kWaves + k) * 64 + threadIdx.x];
            a.part[((size_t)d * kBlockCols + threadIdx.x) * a.ngroups + p] = t;
        }
    }
}

// Column sums of ND per-lane values -> part[d][c][p], fixed order: xor
// butterfly over the lanes of one column inside a wave, then waves 0..3.
template <int ND>
__device__ __forceinline__ void column_partials(const BlockArgs &a, double (&v)[ND], int p)
{
    __shared__ double red[ND * kWaves * 64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < ND; ++d)
        for (int o = 32; o >= a.kp; o >>= 1) v[d] += __shfl_xor(v[d], o, 64);
    if (l < a.kp) {
#pragma unroll
        for (int d = 0; d < ND; ++d) red[(d * kWaves + w) * 64 + l] = v[d];
    }
    waves_to_partials<ND>(a, red, p);
}

// column_partials for lanes holding two columns each (2 (l % 32) and
// 2 (l % 32) + 1, the 16-byte form): lanes l and l + 32 first, then waves
// 0..3.
template <int ND>
__device__ __forceinline__ void column_partials2(const BlockArgs &a, double (&v)[ND][2], int p)
{
    __shared__ double red[ND * kWaves * 64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int e = 0; e < 2; ++e) v[d][e] += __shfl_xor(v[d][e], 32, 64);
    if (l < 32) {
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int e = 0; e < 2; ++e) red[(d * kWaves + w) * 64 + 2 * l + e] = v[d][e];
    }
    waves_to_partials<ND>(a, red, p);
}

// Whether column c takes part in the pass.
__device__ __forceinline__ bool col_active(const BlockArgs &a, int c, bool need_not_sconv)
{
    if (c >= a.kc) return false;
    if (!a.skip_done) return true;
    const SolverState &s = a.st[c];
    return !s.done && !(need_not_sconv && s.s_conv);
}

// Whether this lane's column takes part in the pass.
__device__ __forceinline__ bool column_active(const BlockArgs &a, const LaneMap &m, bool need_not_sconv)
{
    return m.col_ok && col_active(a, m.c, need_not_sconv);
}

// WIDE (kp == 64): one row per wave step; otherwise 64 / kp rows per wave.
// The CSR arrays come as their own restrict arguments, so the compiler may
// read them through the scalar cache.
template <int EPI, bool WIDE, int RB>
__global__ __launch_bounds__(kThreads) void k_bspmm(BlockArgs a, const int *__restrict__ rowptr,
                                                   const int *__restrict__ col, const double *__restrict__ val)
{
    constexpr int ND = EPI == BEPI_DOT1 ? 1 : 2;
    const int p = blockIdx.x;
    const LaneMap m = lane_map(a);
    const bool act = column_active(a, m, EPI == BEPI_DOT2);
    double v[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d] = 0.0;
    // epilogue of row rq of this lane's column (w = the row's operand)
    // y is stored non-temporally: it is read back only by the next kernel,
    // and kept out of the L2 it leaves room for the X rows the sweep
    // gathers again
    auto emit = [&](int rq, double acc, double wr) {
        const size_t i = (size_t)rq * a.ld + m.c;
        if constexpr (EPI == BEPI_DOT1) {
            __builtin_nontemporal_store(acc, a.y + i);
            v[0] += wr * acc;
        } else if constexpr (EPI == BEPI_DOT2) {
            __builtin_nontemporal_store(acc, a.y + i);
            v[0] += acc * wr;
            v[1] += acc * acc;
        } else {
            const double o = wr - acc;
            __builtin_nontemporal_store(o, a.y + i);
            v[0] += o * o;
            v[1] += o * o;
        }
    };
    const bool any = __syncthreads_or(act);
    if constexpr (WIDE) {
        // RB consecutive rows per wave step (wave_sweep64), one contiguous
        // 64-column gather per entry; every wave arrives at the XCD's step
        // counter once (sweep_done), also when no column is active
        const int xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds, per = a.ngroups / kXcds;
        const int nw = per * kWaves;
        const int wv = __builtin_amdgcn_readfirstlane(wi * kWaves + (int)(threadIdx.x >> 6));
        const SweepCtr sc =
            sweep_ctr(a.rr.ctr ? a.rr.ctr + xcd * kCtrSet * kCtrStride : nullptr, wv, nw, a.rr.ctr_per_xcd);
        if (!any) {
            sweep_done(sc, nw);
            column_partials<ND>(a, v, p);
            return;
        }
        // positions of a step are consecutive rows of one run (runs have an
        // even length and each XCD's share starts at a run); the epilogue
        // operands load with the step's gathers (every lane, in bounds)
        int x0, x1;
        xcd_positions(a.rr, a.n, xcd, x0, x1);
        const int wc = m.col_ok ? m.c : 0;
        double wr[RB];
        int p0, ps;
        sweep_span(sc, wv, nw, x0, RB, p0, ps);
        wave_sweep64<double, RB>(
            rowptr, col, val, a.x, a.ld, m.c, act, a.rr, p0, x1, ps, sc.mine,
            [&](int ru, int nr) {
#pragma unroll
                for (int q = 0; q < RB; ++q) wr[q] = a.w[(size_t)(ru + min(q, max(nr - 1, 0))) * a.ld + wc];
            },
            [&](int q, int row, double acc) {
                if (act) emit(row, acc, wr[q]);
            });
        sweep_done(sc, nw);
    } else if (any) {
        const RowSweep sw = row_sweep(a.rr, a.n, a.ngroups, RB * m.rstep);
        const double *__restrict__ x = a.x + m.c;
        for (int t0 = sw.first; t0 < sw.end; t0 += sw.stride) {
            if (!act) continue;
            const int p = t0 + m.ro;
            double accs[RB];
            rows_dot<double, RB>(a.rowptr, a.col, a.val, x, a.ld, a.rr, p, m.rstep, sw.end, accs);
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int pq = p + q * m.rstep;
                if (pq >= sw.end) break;
                const int rq = pos_row(a.rr, pq);
                emit(rq, accs[q], a.w[(size_t)rq * a.ld + m.c]);
            }
        }
    }
    column_partials<ND>(a, v, p);
}

// 64-column groups with X gathered in 16-byte pieces (wave_sweep64v): a
// wave step is two consecutive rows, lane l on row l / 32 and columns
// cc = 2 (l % 32), cc + 1, so one gather instruction covers an entry of
// both rows (2 x 512 bytes) and a step issues half the gathers of the
// one-column-per-lane form.  Each column's sum runs over its row in CSR
// order, as there; the epilogue operand's two columns load with the step's
// gathers.  Needs kc, ld and the group's first column even (16-byte
// aligned rows of x, y and w).
template <int EPI>
__global__ __launch_bounds__(kThreads) void k_bspmm2(BlockArgs a, const int *__restrict__ rowptr,
                                                    const int *__restrict__ col, const double *__restrict__ val)
{
    constexpr int ND = EPI == BEPI_DOT1 ? 1 : 2;
    using Pair = Cols16<double>;
    const int lane = threadIdx.x & 63;
    const int cc = (lane & 31) * 2;
    const bool act0 = col_active(a, cc, EPI == BEPI_DOT2), act1 = col_active(a, cc + 1, EPI == BEPI_DOT2);
    double v[ND][2];
#pragma unroll
    for (int d = 0; d < ND; ++d) v[d][0] = v[d][1] = 0.0;
    const bool any = __syncthreads_or(act0 || act1);
    const int xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds, per = a.ngroups / kXcds;
    const int nw = per * kWaves;
    const int wv = __builtin_amdgcn_readfirstlane(wi * kWaves + (int)(threadIdx.x >> 6));
    const SweepCtr sc =
        sweep_ctr(a.rr.ctr ? a.rr.ctr + xcd * kCtrSet * kCtrStride : nullptr, wv, nw, a.rr.ctr_per_xcd);
    if (any) {
        int x0, x1;
        xcd_positions(a.rr, a.n, xcd, x0, x1);
        int p0, ps;
        sweep_span(sc, wv, nw, x0, 2, p0, ps);
        // lanes past the group's columns read its first two (in bounds)
        const int wc = cc < a.kc ? cc : 0;
        Pair wr;
        wave_sweep64v<double>(
            rowptr, col, val, a.x, a.ld, a.kc, a.rr, p0, x1, ps, sc.mine,
            [&](int ru, int nr) {
                const int r = ru + min(lane >> 5, max(nr - 1, 0));
                wr = *reinterpret_cast<const Pair *>(a.w + (size_t)r * a.ld + wc);
            },
            [&](int, int row, const Pair &acc, int) {
                Pair o;
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const bool on = e == 0 ? act0 : act1;
                    if constexpr (EPI == BEPI_DOT1) {
                        o.v[e] = acc.v[e];
                        v[0][e] += on ? wr.v[e] * o.v[e] : 0.0;
                    } else if constexpr (EPI == BEPI_DOT2) {
                        o.v[e] = acc.v[e];
                        v[0][e] += on ? o.v[e] * wr.v[e] : 0.0;
                        v[1][e] += on ? o.v[e] * o.v[e] : 0.0;
                    } else {
                        o.v[e] = wr.v[e] - acc.v[e];
                        v[0][e] += on ? o.v[e] * o.v[e] : 0.0;
                        v[1][e] += on ? o.v[e] * o.v[e] : 0.0;
                    }
                }
                // y stored non-temporally (read back only by the next kernel)
                double *yp = a.y + (size_t)row * a.ld + cc;
                if (act0 && act1) {
                    using V2 = double __attribute__((ext_vector_type(2)));
                    V2 t;
                    t[0] = o.v[0];
                    t[1] = o.v[1];
                    __builtin_nontemporal_store(t, reinterpret_cast<V2 *>(yp));
                } else {
                    if (act0) __builtin_nontemporal_store(o.v[0], yp);
                    if (act1) __builtin_nontemporal_store(o.v[1], yp + 1);
                }
            });
    }
    // every wave arrives at the XCD's step counter once, also when no
    // column is active
    sweep_done(sc, nw);
    column_partials2<ND>(a, v, blockIdx.x);
}

template <int KIND>
__global__ __launch_bounds__(kThreads) void k_bvec(BlockArgs a)
{
    constexpr int ND = (KIND == BVEC_XR) ? 2 : ((KIND == BVEC_S || KIND == BVEC_DOT) ? 1 : 0);
    const int p = blockIdx.x;
    const LaneMap m = lane_map(a);
    const bool act = column_active(a, m, false);
    double v[ND > 0 ? ND : 1] = {};
    if (__syncthreads_or(act)) {
        double al = 0.0, om = 0.0, be = 0.0;
        int sconv = 0;
        if (m.col_ok && (KIND == BVEC_P || KIND == BVEC_S || KIND == BVEC_XR)) {
            const SolverState &s = a.st[m.c];
            al = s.alpha;
            om = s.omega;
            be = s.beta;
            sconv = s.s_conv;
        }
        const RowSweep sw = row_sweep(RowRuns{}, a.n, a.ngroups, m.rstep);
        for (int t0 = sw.first; t0 < sw.end; t0 += sw.stride) {
            const int r = t0 + m.ro;
            if (!act || r >= sw.end) continue;
            const size_t i = (size_t)r * a.ld + m.c;
            if constexpr (KIND == BVEC_P) {
                a.P[i] = ((a.P[i] - a.V[i] * om) * be) + a.R[i];
            } else if constexpr (KIND == BVEC_S) {
                const double s = a.R[i] - a.V[i] * al;
                a.S[i] = s;
                v[0] += s * s;
            } else if constexpr (KIND == BVEC_XR) {
                if (sconv) {
                    a.X[i] = a.X[i] + a.P[i] * al;
                } else {
                    const double s = a.S[i];
                    a.X[i] = (a.X[i] + a.P[i] * al) + s * om;
                    const double rn = s - a.T[i] * om;
                    a.R[i] = rn;
                    v[0] += rn * a.R0[i];
                    v[ND > 1 ? 1 : 0] += rn * rn;
                }
            } else if constexpr (KIND == BVEC_INIT) {
                const_cast<double *>(a.R0)[i] = a.R[i];
                a.P[i] = 0.0;
                a.V[i] = 0.0;
            } else {
                const double b = a.x[i];
                v[0] += b * b;
            }
        }
    }
    if constexpr (ND > 0) column_partials<ND>(a, v, p);
}

// One workgroup per column: ordered sum of its partials, then the scalar
// update of bicgstab_cpu.c / the stop tests (apply_fin).
__global__ __launch_bounds__(kThreads) void k_bfin(SolverState *st, const double *part, int ngroups, int fin)
{
    __shared__ double red[kMaxDots * kWaves];
    const int c = blockIdx.x;
    const int nd = (fin == FIN_OMEGA || fin == FIN_XR) ? 2 : 1;
    double v[2] = {0.0, 0.0};
    for (int d = 0; d < nd; ++d)
        for (int i = threadIdx.x; i < ngroups; i += kThreads) v[d] += part[((size_t)d * kBlockCols + c) * ngroups + i];
    block_sum<2>(v, red);
    if (threadIdx.x == 0) {
        SolverState s = st[c];
        if (fin == FIN_RESID || fin == FIN_STORE) {
            apply_fin(s, fin, v[0], 0.0);
        } else if (!s.done && !(fin == FIN_OMEGA && s.s_conv)) {
            apply_fin(s, fin, v[0], v[1]);
        }
        st[c] = s;
    }
}

template <int EPI, int RB>
void bspmm_launch(const BlockArgs &a, hipStream_t s)
{
    if (a.kp == 64) k_bspmm<EPI, true, RB><<<a.ngroups, kThreads, 0, s>>>(a, a.rowptr, a.col, a.val);
    else k_bspmm<EPI, false, RB><<<a.ngroups, kThreads, 0, s>>>(a, a.rowptr, a.col, a.val);
}
template <int EPI>
hipError_t bspmm_rb(const BlockArgs &a, hipStream_t s)
{
    // the 16-byte form where every column pair it loads or stores is
    // 16-byte aligned (a caller's B need not be), else one column per lane
    const bool v16 = a.v16 && a.kp == 64 && a.kc % 2 == 0 && a.ld % 2 == 0 && (uintptr_t)a.x % 16 == 0 &&
                     (uintptr_t)a.y % 16 == 0 && (uintptr_t)a.w % 16 == 0;
    if (v16) {
        k_bspmm2<EPI><<<a.ngroups, kThreads, 0, s>>>(a, a.rowptr, a.col, a.val);
        return hipGetLastError();
    }
    switch (a.rb) {
    case 1: bspmm_launch<EPI, 1>(a, s); break;
    case 2: bspmm_launch<EPI, 2>(a, s); break;
    case 4: bspmm_launch<EPI, 4>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_block_spmm(const BlockArgs &a, int epi, hipStream_t s)
{
    if (a.ngroups <= 0) return hipSuccess;
    if (a.rr.run && ((long long)a.rr.nruns * a.rr.len != a.n || a.rr.len % 2 != 0 || a.rr.len % a.rb != 0))
        return hipErrorInvalidValue;
    switch (epi) {
    case BEPI_DOT1: return bspmm_rb<BEPI_DOT1>(a, s);
    case BEPI_DOT2: return bspmm_rb<BEPI_DOT2>(a, s);
    case BEPI_RESID: return bspmm_rb<BEPI_RESID>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_block_vec(const BlockArgs &a, int kind, hipStream_t s)
{
    if (a.ngroups <= 0) return hipSuccess;
    switch (kind) {
    case BVEC_P: k_bvec<BVEC_P><<<a.ngroups, kThreads, 0, s>>>(a); break;
    case BVEC_S: k_bvec<BVEC_S><<<a.ngroups, kThreads, 0, s>>>(a); break;
    case BVEC_XR: k_bvec<BVEC_XR><<<a.ngroups, kThreads, 0, s>>>(a); break;
    case BVEC_INIT: k_bvec<BVEC_INIT><<<a.ngroups, kThreads, 0, s>>>(a); break;
    case BVEC_DOT: k_bvec<BVEC_DOT><<<a.ngroups, kThreads, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_block_fin(SolverState *st, int kc, const double *part, int ngroups, int fin, hipStream_t s)
{
    if (kc <= 0) return hipSuccess;
    k_bfin<<<kc, kThreads, 0, s>>>(st, part, ngroups, fin);
    return hipGetLastError();
}

}  // namespace hbs
