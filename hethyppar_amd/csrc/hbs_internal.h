// hbs_internal.h -- shared definitions of the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "hbs.h"

namespace hbs {

constexpr int kThreads = 256;          // workgroup size of every kernel (4 wave64)
constexpr int kWaves = kThreads / 64;
constexpr int kMaxBlockNnz = 4096;     // upper bound of nnz per SpMV row block
constexpr int kMaxDots = 2;            // reductions fused into one pass

// One SpMV row block.  Normal block: rows [r0, r1) whose entries [j0, j1) fit
// in LDS.  Chunk block: entries [j0, j1) of one long row r0 (r1 == r0 + 1).
struct alignas(16) RowBlock {
    int r0, r1, j0, j1;
};
// chunk < 0: normal block; lanes = lanes per row in the reduction (1 = the
// order-exact serial sum).  chunk >= 0: index of this chunk within its row,
// nchunks chunks starting at block `first`.  sys: system the rows belong to.
struct alignas(16) BlockAux {
    int chunk, nchunks, first;
    int lanes_sys;  // lanes | sys << 8
};
__host__ __device__ inline int aux_lanes(const BlockAux &a) { return a.lanes_sys & 0xff; }
__host__ __device__ inline int aux_sys(const BlockAux &a) { return a.lanes_sys >> 8; }

// A contiguous element range [e0, e1) of one system, processed by one
// workgroup of the streaming vector kernels.
struct alignas(16) VecBlock {
    int sys, e0, e1, pad;
};
// Workgroup ranges of one system: SpMV blocks [b0, b1), vector blocks [v0, v1).
struct alignas(16) SysRange {
    int b0, b1, v0, v1;
};

// Device-resident BiCGStab scalars, one per system.
struct alignas(16) SolverState {
    double rho, alpha, omega, beta;
    double rho_next;     // (r, r0) of the newest r
    double rv, ts, tt;   // (r0, v), (t, s), (t, t)
    double ss, rr;       // ||s||^2, ||r||^2
    double thr;          // tol * ||r0||
    double rr0;          // ||r0||^2
    double tol;
    double scale;        // mixed: ||r|| of the outer residual the sweep solves for
    double resid2;       // ||b - Ax||^2 of the last residual pass
    double out;          // generic reduction result
    int iter, maxit, done, stop;
    int s_conv, fixed, pad0, pad1;
};

enum Epi : int {
    EPI_AXPBY = 0,   // y = alpha*Ax + beta*y
    EPI_DOT1 = 1,    // y = Ax ; d0 = (w, y)
    EPI_DOT2 = 2,    // y = Ax ; d0 = (y, w), d1 = (y, y)
    EPI_RESID = 3,   // y = w - Ax ; d0 = (y, y)
};

enum Fin : int {
    FIN_NONE = 0,
    FIN_ALPHA = 1,   // rv = d0, alpha = rho / rv
    FIN_OMEGA = 2,   // ts = d0, tt = d1, omega = ts / tt
    FIN_INIT = 3,    // rr0 = d0: threshold, first beta
    FIN_RESID = 4,   // resid2 = d0
    FIN_S = 5,       // ss = d0: ||s|| stop test
    FIN_XR = 6,      // rho_next = d0, rr = d1: ||r|| stop test, next beta
    FIN_STORE = 7,   // out = d0
};

// Reduction scratch shared by the kernels of one plan.  Counters reset
// themselves (the last arriving workgroup writes 0), so the scratch needs no
// clearing between launches or graph replays.
struct Scratch {
    double *part = nullptr;       // [nslots * kMaxDots] per-workgroup partials
    double *chunk = nullptr;      // [nslots] long-row chunk sums
    unsigned *ctr_sys = nullptr;  // [nsys] per-system arrival counters
    unsigned *ctr_row = nullptr;  // [nslots] long-row arrival counters
    unsigned *ndone = nullptr;    // [1] systems finished
};

struct SpmvPlan {
    int nblk = 0;
    int nsys = 1;
    int block_nnz = 0;
    int serial_rows = 0;
    int max_row = 0;
    RowBlock *blocks = nullptr;   // device [nblk]
    BlockAux *aux = nullptr;      // device [nblk]
};

struct VecPlan {
    int nvb = 0;
    VecBlock *blocks = nullptr;   // device [nvb]
};

// Launch parameters of the SpMV kernels.
template <typename T>
struct SpmvArgs {
    const RowBlock *blocks;
    const BlockAux *aux;
    const SysRange *sys;
    int nblk;
    const int *rowptr;
    const int *col;
    const T *val;
    const T *x;
    T *y;
    const T *w;
    T alpha, beta;
    Scratch sc;
    SolverState *st;      // [nsys] or null
    int fin;
    int check_done;       // iteration kernels: return if the system is done
    int skip_on_sconv;    // ... or if the ||s|| test fired this iteration
};

// Launch parameters of the vector kernels of one BiCGStab iteration.
template <typename T>
struct VecArgs {
    const VecBlock *blocks;
    const SysRange *sys;
    int nvb;
    T *X, *R, *P, *V, *S, *T_;
    const T *R0;
    SolverState *st;
    Scratch sc;
};

// ---- host-side planning (hbs_plan.cpp) ----
// Row blocks of rows [row0, row0 + m) of one system; appended to blocks/aux.
void plan_rows(int row0, int m, const int *rowptr, int block_nnz, int serial_max_avg, int sys,
               std::vector<RowBlock> &blocks, std::vector<BlockAux> &aux, int *serial_rows,
               int *max_row);
// Vector blocks of elements [e0, e0 + n) of one system.
void plan_vec(int e0, int n, int elems_per_block, int sys, std::vector<VecBlock> &blocks);

// ---- launchers (hbs_kernels.hip) ----
hipError_t launch_spmv(const SpmvArgs<double> &a, int epi, int block_nnz, hipStream_t s);
hipError_t launch_spmv(const SpmvArgs<float> &a, int epi, int block_nnz, hipStream_t s);
hipError_t launch_vec_P(const VecArgs<double> &a, hipStream_t s);
hipError_t launch_vec_S(const VecArgs<double> &a, hipStream_t s);
hipError_t launch_vec_XR(const VecArgs<double> &a, hipStream_t s);
hipError_t launch_vec_P(const VecArgs<float> &a, hipStream_t s);
hipError_t launch_vec_S(const VecArgs<float> &a, hipStream_t s);
hipError_t launch_vec_XR(const VecArgs<float> &a, hipStream_t s);
// R0 = R ; P = V = 0 (FP64, after the residual SpMV wrote R)
hipError_t launch_init_vectors(const VecArgs<double> &a, hipStream_t s);
// Mixed sweep start: R = R0 = (float)(r64 / ||r64||), X = P = V = 0, d0 = ||R||^2
// -> FIN_INIT on the inner state; the inner state's scale = ||r64||.
hipError_t launch_ir_start(const VecArgs<float> &a, const double *r64,
                           const SolverState *outer, hipStream_t s);
// Mixed sweep end: x64 += scale * (double)x32.
hipError_t launch_ir_update(const VecArgs<float> &a, double *x64, hipStream_t s);
// Elementwise helpers over the vector plan (one system or many).
hipError_t launch_dot(const VecArgs<double> &a, const double *x, const double *y, int fin,
                      hipStream_t s);
hipError_t launch_state_init(SolverState *st, int nsys, double tol, int maxit, int fixed,
                             unsigned *ndone, hipStream_t s);

// BLAS-1 helpers on raw device pointers (single system).
hipError_t launch_axpy_f64(int n, double a, const double *x, double *y, hipStream_t s);
hipError_t launch_scal_f64(int n, double a, double *x, hipStream_t s);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_copy_f64(size_t n, const double *src, double *dst, hipStream_t s);
hipError_t launch_to_f32(int n, const double *src, float *dst, hipStream_t s);

// Halo kernels: out[k] = in[idx[k]] and out[idx[k]] = in[k].
hipError_t launch_gather_f64(int n, const double *in, const int *idx, double *out, hipStream_t s);
hipError_t launch_scatter_f64(int n, const double *in, const int *idx, double *out, hipStream_t s);

// Adjoint gradient on the pattern: g[j] = -lambda[row(j)] * x[col[j]].
hipError_t launch_grad_pattern(int m, const int *rowptr, const int *col, const double *lambda,
                               const double *x, double *g, hipStream_t s);

// SpMM (row-major blocks, ld = k).
hipError_t launch_spmm_f64(int m, const int *rowptr, const int *col, const double *val, int k,
                           const double *X, double *Y, int tile, hipStream_t s);
hipError_t launch_spmm_f32(int m, const int *rowptr, const int *col, const float *val, int k,
                           const float *X, float *Y, int tile, hipStream_t s);

// Reference-shaped fused kernels (hhp_dp_kernels.h).
hipError_t launch_dp_vecop_P(double *P, const double *V, const double *R, const double *omega,
                             const double *beta, int n, hipStream_t s);
hipError_t launch_dp_vecop_S(double *S, const double *R, const double *V,
                             const double *neg_alpha, int n, hipStream_t s);
hipError_t launch_dp_vecop_XR(double *X, double *R, const double *P, const double *S,
                              const double *T, const double *alpha, const double *omega,
                              const double *neg_omega, int n, hipStream_t s);
// which: 0 beta, 1 alpha, 2 omega, 3 add
hipError_t launch_dp_scalar(int which, double *a, double *b, const double *c, const double *d,
                            const double *e, hipStream_t s);

}  // namespace hbs
