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
// contiguous share of the rows, dealt in tiles round robin over its
// workgroups (gridDim.x, a multiple of 8), so that neighbouring rows -- which
// gather the same X rows -- are worked on together under one L2.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spmm(int m, const int *__restrict__ rowptr,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ val, int k, int ld,
                                                  const T *__restrict__ X, T *__restrict__ Y, int tile)
{
    constexpr int RB = 2;   // rows per lane in flight together
    const int lane = threadIdx.x & (tile - 1);
    const int rstep = kThreads / tile;
    const int c = blockIdx.y * tile + lane;
    if (c >= k) return;
    const int xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds, per = gridDim.x / kXcds;
    const int x0 = (int)((long long)m * xcd / kXcds), x1 = (int)((long long)m * (xcd + 1) / kXcds);
    for (int t0 = x0 + wi * RB * rstep; t0 < x1; t0 += per * RB * rstep) {
        const int r = t0 + (int)threadIdx.x / tile;
        T acc[RB];
        rows_dot<T, RB>(rowptr, col, val, X + c, ld, r, rstep, x1, acc);
#pragma unroll
        for (int q = 0; q < RB; ++q)
            if (r + q * rstep < x1) Y[(size_t)(r + q * rstep) * ld + c] = acc[q];
    }
}

// Tiles of 64 columns: one wave per row step (wave_rows64), so the row's
// indices and values are read once per wave through the scalar cache and
// every gather is a contiguous 64-column stretch of an X row.  Rows are
// dealt per XCD as in k_spmm, in steps of RB rows per wave.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spmm64(int m, const int *__restrict__ rowptr,
                                                    const int *__restrict__ col, const T *__restrict__ val,
                                                    int k, int ld, const T *__restrict__ X, T *__restrict__ Y)
{
    constexpr int RB = 2;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.y * 64 + lane;
    const bool ok = c < k;
    const int xcd = blockIdx.x % kXcds, wi = blockIdx.x / kXcds, per = gridDim.x / kXcds;
    const int x0 = (int)((long long)m * xcd / kXcds), x1 = (int)((long long)m * (xcd + 1) / kXcds);
    const int wv = wi * kWaves + (int)(threadIdx.x >> 6), nw = per * kWaves;
    for (int r = x0 + wv * RB; r < x1; r += nw * RB) {
        const int ru = __builtin_amdgcn_readfirstlane(r);
        wave_rows64<T, RB>(rowptr, col, val, X, ld, c, ru, x1, ok, [&](int q, T acc) {
            if (ok && ru + q < x1) Y[(size_t)(ru + q) * ld + c] = acc;
        });
    }
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
// About 2048 workgroups in all, the row workgroups a multiple of 8.
template <typename T>
hipError_t spmm_dispatch(int m, const int *rowptr, const int *col, const T *val, int k, const T *X, T *Y,
                         int tile, hipStream_t s)
{
    if (m <= 0 || k <= 0) return hipSuccess;
    if (tile < 1 || tile > 64 || (tile & (tile - 1))) return hipErrorInvalidValue;
    const int rstep = kThreads / tile;
    const int ytiles = (k + tile - 1) / tile;
    const long long passes = ((long long)m + 2LL * rstep - 1) / (2LL * rstep);
    long long gx = std::min(std::max(1LL, 2048LL / ytiles), passes);
    gx = std::max(8LL, (gx + 7) / 8 * 8);
    dim3 grid((unsigned)gx, (unsigned)ytiles);
    if (tile == 64) k_spmm64<T><<<grid, kThreads, 0, s>>>(m, rowptr, col, val, k, k, X, Y);
    else k_spmm<T><<<grid, kThreads, 0, s>>>(m, rowptr, col, val, k, k, X, Y, tile);
    return hipGetLastError();
}
hipError_t launch_spmm_f64(int m, const int *rowptr, const int *col, const double *val, int k,
                           const double *X, double *Y, int tile, hipStream_t s)
{
    return spmm_dispatch<double>(m, rowptr, col, val, k, X, Y, tile, s);
}
hipError_t launch_spmm_f32(int m, const int *rowptr, const int *col, const float *val, int k,
                           const float *X, float *Y, int tile, hipStream_t s)
{
    return spmm_dispatch<float>(m, rowptr, col, val, k, X, Y, tile, s);
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
