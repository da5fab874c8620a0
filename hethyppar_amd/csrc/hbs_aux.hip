// hbs_aux.hip -- auxiliary kernels: BLAS-1, fills and copies, halo
// gather/scatter, the adjoint's pattern gradient, CSR x dense-block SpMM, and
// the reference-shaped device-scalar kernels of hhp_dp_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hbs_device.h"
#include "hbs_internal.h"

namespace hbs {
namespace {

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
__global__ void k_gather_f32(int n, const double *in, const int *idx, float *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (float)in[idx[i]];
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
// lanes share one row (one RHS column each) and sum it in CSR order, so
// every column equals the single-RHS order-exact sum.  Each XCD takes a
// contiguous share of the traversal rr (the rows in order, or runs of
// consecutive rows in y-slab order on a 3-D grid), dealt in tiles round
// robin over its workgroups (gridDim.x, a multiple of 8), so that rows
// which gather the same X rows are worked on together under one L2.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spmm(int m, const int *__restrict__ rowptr,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ val, int k, int ld,
                                                  const T *__restrict__ X, T *__restrict__ Y, int tile,
                                                  RowRuns rr)
{
    constexpr int RB = 2;   // rows per lane in flight together
    const int lane = threadIdx.x & (tile - 1);
    const int rstep = kThreads / tile;
    const int c = blockIdx.y * tile + lane;
    if (c >= k) return;
    const int xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds, per = gridDim.x / kXcds;
    int x0, x1;
    xcd_positions(rr, m, xcd, x0, x1);
    for (int t0 = x0 + wi * RB * rstep; t0 < x1; t0 += per * RB * rstep) {
        const int p = t0 + (int)threadIdx.x / tile;
        T acc[RB];
        rows_dot<T, RB>(rowptr, col, val, X + c, ld, rr, p, rstep, x1, acc);
#pragma unroll
        for (int q = 0; q < RB; ++q)
            if (p + q * rstep < x1) Y[(size_t)pos_row(rr, p + q * rstep) * ld + c] = acc[q];
    }
}

// Tiles of 64 columns: one wave per row step (wave_sweep64), so each entry
// is read once per wave and every gather is a contiguous 64-column stretch
// of an X row.  Each XCD sweeps its share of the positions in steps of RB
// consecutive rows (a run's length is even, so a step never leaves its
// run), taken from the tile's step counter where it has one.
template <typename T, bool V16>
__global__ __launch_bounds__(kThreads) void k_spmm64(int m, const int *__restrict__ rowptr,
                                                    const int *__restrict__ col, const T *__restrict__ val,
                                                    int k, int ld, const T *__restrict__ X, T *__restrict__ Y,
                                                    RowRuns rr)
{
    constexpr int RB = 2;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.y * 64 + lane;
    const bool ok = c < k;
    const int xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds, per = gridDim.x / kXcds;
    int x0, x1;
    xcd_positions(rr, m, xcd, x0, x1);
    const int wv = __builtin_amdgcn_readfirstlane(wi * kWaves + (int)(threadIdx.x >> 6)), nw = per * kWaves;
    const SweepCtr sc = sweep_ctr(
        rr.ctr && (int)blockIdx.y < rr.ctr_tiles ? rr.ctr + ((int)blockIdx.y * kXcds + xcd) * kCtrSet * kCtrStride
                                                 : nullptr,
        wv, nw, rr.ctr_per_xcd);
    int p0, ps;
    if constexpr (V16) {
        // 16-byte gathers: a step of G rows, columns in 16-byte pieces
        constexpr int G = 64 / (64 / Cols16<T>::N);
        sweep_span(sc, wv, nw, x0, G, p0, ps);
        const int cbase = (int)blockIdx.y * 64;
        wave_sweep64v<T>(rowptr, col, val, X + cbase, ld, min(64, k - cbase), rr, p0, x1, ps, sc.mine,
                         [](int, int) {}, [&](int, int row, const Cols16<T> &acc, int cc) {
                             // Y stored non-temporally, out of the L2 the
                             // gathered X rows share (FP64 128^3 x 64: 0.81
                             // -> 0.77 ms)
                             if (cbase + cc < k) {
                                 using V = T __attribute__((ext_vector_type(Cols16<T>::N)));
                                 V vv;
#pragma unroll
                                 for (int e = 0; e < Cols16<T>::N; ++e) vv[e] = acc.v[e];
                                 __builtin_nontemporal_store(vv, reinterpret_cast<V *>(Y + (size_t)row * ld + cbase + cc));
                             }
                         });
    } else {
        sweep_span(sc, wv, nw, x0, RB, p0, ps);
        wave_sweep64<T, RB>(rowptr, col, val, X, ld, c, ok, rr, p0, x1, ps, sc.mine, [](int, int) {},
                            [&](int, int row, T acc) {
                                if (ok) Y[(size_t)row * ld + c] = acc;
                            });
    }
    sweep_done(sc, nw);
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

}  // namespace

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
hipError_t launch_gather_to_f32(int n, const double *in, const int *idx, float *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_gather_f32<<<grid_for(n), kThreads, 0, s>>>(n, in, idx, out);
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
// About 2048 workgroups in all (1024 for the 64-column tiles with 16-byte
// gathers: on the 128^3 operator, 1024 instead of 2048 cut the X re-reads
// from 3.1 to 2.0 GB per FP64 SpMM, since an XCD's rows in flight bound the
// window of X its L2 must hold), the row workgroups a multiple of 8.
template <typename T>
hipError_t spmm_dispatch(int m, const int *rowptr, const int *col, const T *val, int k, const T *X, T *Y,
                         int tile, const RowRuns &rr, hipStream_t s)
{
    if (m <= 0 || k <= 0) return hipSuccess;
    if (tile < 1 || tile > 64 || (tile & (tile - 1))) return hipErrorInvalidValue;
    const int rstep = kThreads / tile;
    const int ytiles = (k + tile - 1) / tile;
    const long long passes = ((long long)m + 2LL * rstep - 1) / (2LL * rstep);
    // 16-byte gathers where every lane's 16 bytes lie in a row
    const bool v16 = g_spmm_v16 && ((size_t)k * sizeof(T)) % 16 == 0 && (uintptr_t)X % 16 == 0 &&
                     (uintptr_t)Y % 16 == 0;
    const long long target = g_spmm_groups > 0 ? g_spmm_groups : (tile == 64 && v16 ? 1024 : 2048);
    long long gx = std::min(std::max(1LL, target / ytiles), passes);
    gx = std::max(8LL, (gx + 7) / 8 * 8);
    dim3 grid((unsigned)gx, (unsigned)ytiles);
    if (rr.run && ((long long)rr.nruns * rr.len != m || rr.len % 2 != 0)) return hipErrorInvalidValue;
    if (tile == 64 && v16) k_spmm64<T, true><<<grid, kThreads, 0, s>>>(m, rowptr, col, val, k, k, X, Y, rr);
    else if (tile == 64) k_spmm64<T, false><<<grid, kThreads, 0, s>>>(m, rowptr, col, val, k, k, X, Y, rr);
    else k_spmm<T><<<grid, kThreads, 0, s>>>(m, rowptr, col, val, k, k, X, Y, tile, rr);
    return hipGetLastError();
}
hipError_t launch_spmm_f64(int m, const int *rowptr, const int *col, const double *val, int k,
                           const double *X, double *Y, int tile, const RowRuns &rr, hipStream_t s)
{
    return spmm_dispatch<double>(m, rowptr, col, val, k, X, Y, tile, rr, s);
}
hipError_t launch_spmm_f32(int m, const int *rowptr, const int *col, const float *val, int k,
                           const float *X, float *Y, int tile, const RowRuns &rr, hipStream_t s)
{
    return spmm_dispatch<float>(m, rowptr, col, val, k, X, Y, tile, rr, s);
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
