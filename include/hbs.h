/*
 * hbs.h -- native C-ABI of the MI355X BiCGStab / SpMV engine (libhbs.so).
 *
 * One rank's share of HetHypPar's BiCGStab solve on one AMD Instinct MI355X
 * (gfx950).  Plain C types only: handles are opaque pointers, arrays are plain
 * pointers with int sizes, every call returns an int status (HBS_OK == 0,
 * negative on error; hbs_last_error() gives the message).  Nothing here calls
 * exit().
 *
 * Pointer convention: "host" arguments are host memory; "dev" arguments are
 * device memory (hbs_dev_alloc or any hipMalloc'd buffer).  Device work is
 * queued on the context's stream and is asynchronous unless the call returns a
 * host value.
 *
 * The reference-shaped API (device_csr_create, hhp_dp_vecop_P, ...) that a
 * reference caller links against is in hhp_hip.h and hhp_dp_kernels.h; both
 * are thin layers over this header.
 */
#ifndef HBS_H
#define HBS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define HBS_OK              0
#define HBS_ERR_ARG        -1   /* bad argument (null, size, shape)       */
#define HBS_ERR_HIP        -2   /* HIP runtime error                      */
#define HBS_ERR_ALLOC      -3   /* host or device allocation failed       */
#define HBS_ERR_IO         -4   /* file cannot be opened                  */
#define HBS_ERR_FORMAT     -5   /* malformed or unsupported file          */
#define HBS_ERR_PARTITION  -6   /* partition vector out of range / empty  */
#define HBS_ERR_BREAKDOWN  -7   /* BiCGStab breakdown (rho or (r0,v) == 0, non-finite) */
#define HBS_ERR_NOCONV     -8   /* max iterations reached above tolerance */
#define HBS_ERR_NODEVICE   -9   /* no HIP device visible                  */

const char *hbs_last_error(void);
const char *hbs_version(void);

/* ---- context: one device + one stream -------------------------------- */
typedef struct hbs_ctx_s *hbs_ctx;
/* stream: a hipStream_t, or NULL to create a private non-blocking stream
 * (so the legacy default stream, whose handle is NULL, cannot be chosen:
 * callers on it synchronise it before handing buffers to the library). */
int hbs_ctx_create(int device, void *stream, hbs_ctx *out);
int hbs_ctx_destroy(hbs_ctx ctx);
int hbs_ctx_sync(hbs_ctx ctx);
void *hbs_ctx_stream(hbs_ctx ctx);
int hbs_device_count(int *out);

/* device memory helpers (for C and ctypes callers without their own allocator) */
int hbs_dev_alloc(hbs_ctx ctx, size_t bytes, void **dev);
int hbs_dev_free(hbs_ctx ctx, void *dev);
int hbs_h2d(hbs_ctx ctx, void *dev, const void *host, size_t bytes);   /* synchronous */
int hbs_d2h(hbs_ctx ctx, void *host, const void *dev, size_t bytes);   /* synchronous */
int hbs_d2d(hbs_ctx ctx, void *dst, const void *src, size_t bytes);    /* async */
int hbs_dev_zero(hbs_ctx ctx, void *dev, size_t bytes);                /* async, fill kernel */
/* page-locked host memory (hipHostMalloc), for halo buffers whose copies to
 * the device should run asynchronously at full PCIe rate */
int hbs_pinned_alloc(hbs_ctx ctx, size_t bytes, void **host);
int hbs_pinned_free(hbs_ctx ctx, void *host);

/* ---- CSR matrix on the device ---------------------------------------- */
typedef struct hbs_csr_s *hbs_csr;

/* Deep-copies a host CSR (rowptr[m+1], col[nnz], val[nnz], 0-based int32
 * indices, FP64 values) to HBM and builds the row-block plan of the CSR-stream
 * SpMV.  flags: HBS_CSR_F32 also keeps an FP32 copy of the values (mixed
 * precision); HBS_CSR_TRANSPOSE also keeps A^T as CSR (adjoint).  n is at
 * most 536,870,911 (x is gathered through 32-bit byte offsets); larger is
 * HBS_ERR_ARG.  Replaces device_csr_create (reference
 * src/util/hhp_cuda.c:19-51). */
#define HBS_CSR_F32       1
#define HBS_CSR_TRANSPOSE 2
int hbs_csr_create(hbs_ctx ctx, int m, int n, int nnz, const int *rowptr,
                   const int *col, const double *val, int flags, hbs_csr *out);
int hbs_csr_destroy(hbs_csr a);
/* Many independent systems in one matrix: a block-diagonal CSR of nsys
 * square systems, system q owning rows and columns [sys_rows[q],
 * sys_rows[q+1]) (sys_rows[0] = 0, m = n = sys_rows[nsys]; an entry outside
 * its system's block is an error).  Each system gets its own row blocks and
 * vector workgroups, so hbs_bicgstab_systems_f64 runs all of them in the same
 * launches, each with its own scalars and stop test.  No reference
 * counterpart (the reference solves one system per process). */
int hbs_csr_create_systems(hbs_ctx ctx, int nsys, const int *sys_rows, int nnz,
                           const int *rowptr, const int *col, const double *val,
                           int flags, hbs_csr *out);
int hbs_csr_systems(hbs_csr a, int *nsys);
/* m, n, nnz, number of row blocks, rows handled by serial (order-exact) blocks */
int hbs_csr_info(hbs_csr a, int *m, int *n, int *nnz, int *nblocks, int *serial_rows);
/* Replace the values (same pattern) from a device array val_dev[nnz] in CSR
 * order; the FP32 copy and A^T follow.  Asynchronous on the context stream.
 * For callers that differentiate through the values. */
int hbs_csr_set_values(hbs_csr a, const double *val_dev);
/* Pointers to the device arrays (rowptr, col, val) for callers that need them. */
int hbs_csr_device_arrays(hbs_csr a, const int **rowptr, const int **col, const double **val);

/* y = alpha*A*x + beta*y (beta == 0 never reads y).  Rows of short-row blocks
 * are summed in CSR order with separately rounded products, which reproduces
 * the reference CSR_spmxv_seq (hhp_cpu.c:445-455) bit for bit when alpha == 1
 * and beta == 0.  Replaces device_csr_spmv (hhp_cuda.c:226-236). */
int hbs_csr_spmv_f64(hbs_csr a, double alpha, const double *x_dev, double beta, double *y_dev);
int hbs_csr_spmv_f32(hbs_csr a, float alpha, const float *x_dev, float beta, float *y_dev);
/* y = A^T x (needs HBS_CSR_TRANSPOSE). */
int hbs_csr_spmv_t_f64(hbs_csr a, const double *x_dev, double *y_dev);

/* Batched: Y[m x k] = A * X[n x k], row-major blocks (ld = k).  tile = RHS
 * columns per wavefront pass (0 = autotuned). */
int hbs_csr_spmm_f64(hbs_csr a, int k, const double *X_dev, double *Y_dev, int tile);
int hbs_csr_spmm_f32(hbs_csr a, int k, const float *X_dev, float *Y_dev, int tile);

/* ---- BLAS-1 on device vectors (reference device_vector_* shapes) ------ */
/* dot returns to host (synchronous), like device_vector_dot (hhp_cuda.c:193-195). */
int hbs_dot_f64(hbs_ctx ctx, int n, const double *x_dev, const double *y_dev, double *out_host);
int hbs_axpy_f64(hbs_ctx ctx, int n, double a, const double *x_dev, double *y_dev); /* y += a*x */
int hbs_scal_f64(hbs_ctx ctx, int n, double a, double *x_dev);                       /* x *= a   */
/* hbs_dot_f64 without any device allocation, for callers that cannot
 * allocate (tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same kernels, plan and summation order as
 * hbs_dot_f64, so the same result bit for bit; its plan, partials and
 * result live in one file-scope device array of the library (513 KiB per
 * GPU, held until the process ends).  n is at most 2^26 (larger is
 * HBS_ERR_ARG).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (from two host threads, or in flight on two streams)
 * race. */
int hbs_dot_f64_noalloc(hbs_ctx ctx, int n, const double *x_dev, const double *y_dev, double *out_host);

/* ---- BiCGStab -------------------------------------------------------- */
typedef struct {
    double tol;          /* stop when ||s|| or ||r|| < tol*||r0||            */
    int maxit;           /* iteration cap                                    */
    int fixed_iters;     /* 1: run exactly maxit iterations, no stop test    */
    int check_every;     /* iterations per convergence poll (0 = default 4)  */
    int use_graph;       /* 1: replay hipGraph-captured iteration chunks;    */
                         /* 0: launch the kernels one by one; 2 (default):   */
                         /* graphs below 4M matrix entries, else one by one  */
    int inner_maxit;     /* mixed: FP32 inner iteration cap per sweep        */
    double inner_tol;    /* mixed: FP32 inner relative tolerance per sweep   */
    int max_sweeps;      /* mixed: FP64 refinement sweeps cap                */
} hbs_solve_opts;

typedef struct {
    int iters;           /* BiCGStab iterations executed (mixed: sum over sweeps) */
    int stop;            /* 0 max-iter, 1 ||s|| test, 2 ||r|| test, 4 breakdown   */
    int sweeps;          /* mixed: FP64 refinement sweeps                          */
    int status;          /* HBS_OK, HBS_ERR_NOCONV or HBS_ERR_BREAKDOWN            */
    double relres;       /* sqrt(||b-Ax||^2 / ||b||^2), FP64, computed on device  */
    double rel_r0;       /* ||b-Ax|| / ||r0||                                      */
    double time_ms;      /* device time of the solve (hipEvent)                    */
} hbs_solve_info;

void hbs_solve_opts_default(hbs_solve_opts *o);
/* Longest host wait, in seconds, for one chunk of iterations before a solve
 * gives up with HBS_ERR_HIP (default 600; process-wide).  Work already
 * queued then still runs: synchronise the context before reusing buffers. */
int hbs_set_wait_limit(double seconds);

typedef struct hbs_solver_s *hbs_solver;
/* Workspace for BiCGStab on A (vectors, device scalars, graphs).  Reusable. */
int hbs_solver_create(hbs_csr a, hbs_solver *out);
int hbs_solver_destroy(hbs_solver s);
/* FP64 BiCGStab, the drop-in for the reference loop (bicgstab_cpu.c:172-255,
 * bicgstab_gpu_dp.c:145-208) with the stop test of bicgstab_gpu.c:148-150,
 * 231-236, 289-293.  b_dev, x_dev: device; x holds x0 on entry and may
 * be b itself (x0 = b, b is overwritten by the solution).  On a
 * matrix of several systems every system is solved at once and info gets
 * the most iterations, the largest residuals and the worst status. */
int hbs_bicgstab_f64(hbs_solver s, const double *b_dev, double *x_dev,
                     const hbs_solve_opts *opts, hbs_solve_info *info);
/* Every system of a hbs_csr_create_systems matrix in the same launches;
 * infos[nsys] gets each system's iterations, stop test and residual.  Returns
 * the worst status (HBS_ERR_BREAKDOWN, then HBS_ERR_NOCONV). */
int hbs_bicgstab_systems_f64(hbs_solver s, const double *b_dev, double *x_dev,
                             const hbs_solve_opts *opts, hbs_solve_info *infos);
/* k right-hand sides: B_dev and X_dev are row-major n x k blocks (ld = k),
 * X holds X0 on entry.  One BiCGStab per column, each with its own scalars,
 * iteration count and stop test; columns are processed in groups of up to
 * 64 (the tuned tile) whose SpMM passes read A once for the whole group.  Each column's
 * products are summed in CSR order, as the single-vector solve does.
 * infos[k]; returns the worst status.  No reference counterpart (the
 * reference solves one right-hand side per run). */
int hbs_bicgstab_block_f64(hbs_solver s, int k, const double *B_dev, double *X_dev,
                           const hbs_solve_opts *opts, hbs_solve_info *infos);
/* The block solve's autotuner picks, once per workspace (solver and k), the
 * column-group width -- the SpMM's RHS tile, a power of two up to 64 and
 * k -- and the SpMM rows per lane step (1, 2, 4), timing the (r0, v) SpMM
 * of each candidate on the workspace.  hbs_set_block_tuning forces either
 * (0: tuned) for workspaces created afterwards; hbs_block_tuning reports the
 * choice for k: tile, rows per lane step, and the rows a wavefront works on
 * at once ((64 / tile) * rows per step).  64-column tiles with k even, wide
 * loads on (hbs_set_spmm_tuning) and the rows per step not forced gather X
 * in 16-byte pieces, two columns per lane and two rows per step (reported
 * as 2, 2).  No reference counterpart. */
int hbs_set_block_tuning(int tile, int rows_per_step);
int hbs_block_tuning(hbs_solver s, int k, int *tile, int *rows_per_step, int *rows_per_wave);
/* The tuner also times the SpMMs with the streaming kernels' workgroup
 * count, a half and a quarter of it (fewer rows in flight per XCD keep
 * more of the gathered X rows in its L2).  hbs_set_block_spmm_groups forces
 * the count for workspaces created afterwards (0: tuned; else a multiple of
 * 8, capped at the streaming kernels'); hbs_block_spmm_groups reports the
 * workgroups the SpMMs of the block solve for k use.  No reference
 * counterpart. */
int hbs_set_block_spmm_groups(int groups);
int hbs_block_spmm_groups(hbs_solver s, int k, int *groups);
/* Mixed precision: FP32 BiCGStab sweeps (FP32 values and vectors, int32
 * indices, compensated dots) inside FP64 iterative refinement; FP64 for
 * r = b - A x, the scalar reductions and the convergence norm.  Needs
 * HBS_CSR_F32. */
int hbs_bicgstab_mixed(hbs_solver s, const double *b_dev, double *x_dev,
                       const hbs_solve_opts *opts, hbs_solve_info *info);
/* Adjoint: solve A^T lambda = g (g = dL/dx) with FP64 BiCGStab on the
 * transpose (HBS_CSR_TRANSPOSE), then dL/db = lambda and
 * dL/dA_ij = -lambda_i * x_j on A's pattern (grad_val[nnz], CSR order). */
int hbs_adjoint_f64(hbs_solver s, const double *g_dev, const double *x_dev,
                    double *lambda_dev, double *grad_val_dev,
                    const hbs_solve_opts *opts, hbs_solve_info *info);

/* ---- host-side layout (reference L2, src/util/hhp_matrix.c) ----------- */
/* Host structures with the reference's field layout, so a reference caller's
 * structs can be passed as they are (hhp_hip.h gives them the reference's
 * names).  Arrays returned by the calls below are malloc'd by the library and
 * released with hbs_host_free / hbs_host_shard_free. */
typedef struct hbs_host_csr {      /* COO/CSR/CSC, hhp_common.h:43-58 */
    int *I;                        /* CSR: row pointers [m+1]; SHARD CSC: column pointers [n+1] */
    int *J;                        /* CSR: column indices;     SHARD CSC: row indices          */
    double *val;
    int nnz;
    int m;
    int n;
} hbs_host_csr;
typedef struct hbs_host_ivec {     /* iVector, hhp_common.h:12-15 */
    int *vals;
    unsigned int nvals;
} hbs_host_ivec;
typedef struct hbs_host_comm {     /* COMM, hhp_common.h:88-134 */
    double *val;                   /* staging buffer [I[num]]                */
    int *J;                        /* local indices [I[num]]                 */
    int *I;                        /* per-peer offsets [num+1]               */
    int *ranks;                    /* peer ranks, ascending [num]            */
    int num;
} hbs_host_comm;
typedef struct hbs_host_shard {    /* SHARD_CSC, hhp_common.h:165-171 */
    hbs_host_csr loc;              /* m_loc x m_loc, SHARD CSC               */
    hbs_host_csr shr;              /* m_loc x n_shr, SHARD CSC               */
    hbs_host_ivec gind;            /* global index of [loc cols ; shr cols]  */
    hbs_host_comm send;
    hbs_host_comm recv;
} hbs_host_shard;
void hbs_host_free(void *p);

/* Matrix Market reader (reference readMatrixMarketFromFile,
 * hhp_wrap_bora.c:13-102): per-row file order kept, duplicates kept. */
int hbs_mm_read(const char *fname, int *m, int *n, int *nnz,
                int **rowptr, int **col, double **val);
/* whitespace separated doubles (reference vector_read, hhp_cpu.c:411-428) */
int hbs_vector_read(const char *fname, int size, double *out);
/* "%lf\n" per value (reference vector_write, hhp_cpu.c:430-442) */
int hbs_vector_write(const char *fname, int size, const double *v);
/* Rank `rank`'s slice of a global vector big[size]: out[k] = big[i_k] for
 * the rows i_0 < i_1 < ... with partvec[i] == rank, which is the order of
 * the shard's owned columns (gind[0..loc.n)).  Replaces
 * MPI_vector_read_parted (hhp_cpu.c:252-334) minus the Scatterv.  out may be
 * NULL to count; *nloc (may be NULL) receives the count. */
int hbs_vector_part(int size, const double *big, const int *partvec, int nparts, int rank, double *out,
                    int *nloc);
/* The same from a vector file (vector_read format); *out is malloc'd
 * (hbs_host_free). */
int hbs_vector_read_parted(const char *fname, int size, const int *partvec, int nparts, int rank,
                           double **out, int *nloc);
/* One rank's part of the gather back: big[gind[k]] = loc[k], k < nloc.
 * Replaces MPI_vector_gather (hhp_cpu.c:199-250), whose root Gathervs every
 * rank's values and gind and scatters them this way; called for every
 * rank's share it assembles the global vector. */
int hbs_vector_gather(int nloc, const double *loc, const int *gind, int size, double *big);

/* ---- synthetic workloads (SplitMix64 streams, one per row) ----------- */
/* 2D 5-point convection-diffusion on nx*ny: diag 4+conv, west -1-conv,
 * east/south/north -1 (Dirichlet neighbours dropped). */
int hbs_gen_convdiff2d(int nx, int ny, double conv, int *m, int *nnz,
                       int **rowptr, int **col, double **val);
/* 3D 7-point: diag 6+conv, -x neighbour -1-conv, others -1. */
int hbs_gen_convdiff3d(int nx, int ny, int nz, double conv, int *m, int *nnz,
                       int **rowptr, int **col, double **val);
/* Markov-like: per row k~U{kmin..kmax} distinct columns in a band of
 * +-band around i (i excluded) plus one uniform column outside the band,
 * values U[0,1], diagonal = dom * row sum. */
int hbs_gen_markov(int n, int kmin, int kmax, int band, double dom, uint64_t seed,
                   int *nnz, int **rowptr, int **col, double **val);
/* Circuit-like: truncated-Zipf(zipf_s) off-diagonal counts calibrated to
 * mean_nnz per row, nheavy rows with log-uniform [heavy_lo, heavy_hi] nnz,
 * uniform columns, values N(0,1)*10^U(-3,3), diagonal = dom * sum|a_ij|;
 * equilibrate != 0 divides each row by its diagonal. */
int hbs_gen_circuit(int n, double mean_nnz, double zipf_s, int nheavy, int heavy_lo,
                    int heavy_hi, double dom, int equilibrate, uint64_t seed,
                    int *nnz, int **rowptr, int **col, double **val);
/* FEM-like: 3x3 blocks, nblk_per_row blocks per block row (diagonal, band,
 * one random coupling), values N(0,1), diagonal shifted to dom * sum|a_ij|. */
int hbs_gen_fem(int nblockrows, int nblk_per_row, double dom, uint64_t seed,
                int *nnz, int **rowptr, int **col, double **val);
/* x ~ U[lo,hi] from seed */
int hbs_gen_uniform(int n, double lo, double hi, uint64_t seed, double *out);
/* y = A x on the host in CSR order (the reference CSR_spmxv_seq order); used
 * to form b = A*x_target for synthetic workloads. */
int hbs_host_csr_spmv(int m, const int *rowptr, const int *col, const double *val,
                      const double *x, double *y);

/* ---- row partition --------------------------------------------------- */
/* Partition vector for the shard build: partvec[r] in [0, nparts) for each
 * row of a square CSR (n x n), replacing CalcPartVec (hhp_matrix.c:225-300),
 * which calls the closed PaToH library.  Same model: the rows are nets over
 * their columns, cell r (column and row r) weighs row r's entries, and the
 * metric is connectivity - 1 (PATOH_CONPART), the x values the rows' owners
 * must receive.  Part weights follow target[nparts] (relative; NULL = equal)
 * within a factor 1 + final_imbal; `seed` fixes the result.  The algorithm
 * is this library's own multilevel recursive bisection, not PaToH's, so
 * partitions differ from PaToH's.  *cut (may be NULL) receives the
 * connectivity - 1 of the result. */
int hbs_partition(int n, const int *rowptr, const int *col, int nparts, const double *target,
                  double final_imbal, int seed, int *partvec, long long *cut);

/* ---- one rank's share: shard build, halo pack/unpack, shard SpMV ------ */
/* Builds rank `rank`'s SHARD_CSC (loc, shr, gind, send, recv) from the global
 * CSR and partition vector in one process; bit-identical to the reference's
 * MPI_CSR_split_row (hhp_matrix.c:1249-1338). */
int hbs_shard_build_host(int n, const int *rowptr, const int *col, const double *val,
                         const int *partvec, int nparts, int rank,
                         hbs_host_shard *out);
void hbs_host_shard_free(hbs_host_shard *s);

typedef struct hbs_shard_s *hbs_shard;
/* Uploads one rank's share: loc and shr merged into one CSR over the extended
 * vector [x_loc (loc.n) ; x_shr (shr.n)], plus send.J and recv.J. */
int hbs_shard_create(hbs_ctx ctx, const hbs_host_shard *host, hbs_shard *out);
int hbs_shard_destroy(hbs_shard s);
int hbs_shard_info(hbs_shard s, int *m_loc, int *n_shr, int *n_send, int *nnz);
/* Halo pack: send_dev[i] = x_loc_dev[send.J[i]] (reference hhp_cpu.c:533-538). */
int hbs_shard_pack(hbs_shard s, const double *x_loc_dev, double *send_dev);
/* Halo unpack: x_ext_dev[loc.n + recv.J[p]] = recv_dev[p] (reference CPU
 * semantics, hhp_cpu.c:574-584). */
int hbs_shard_unpack(hbs_shard s, const double *recv_dev, double *x_ext_dev);
/* y_loc = loc*x_loc + shr*x_shr.  x_ext_dev holds x_loc in its first loc.n
 * entries; recv_host (shr.n values in recv order, e.g. straight from MPI) is
 * copied to the device and unpacked into x_ext_dev[loc.n..] first.  Pass
 * recv_host = NULL when x_ext_dev already holds the halo.  Replaces
 * MPI_device_SHARD_CSC_mpi_spmxv (hhp_cuda.c:259-352) minus the MPI transport. */
int hbs_shard_spmv(hbs_shard s, double *x_ext_dev, const double *recv_host, double *y_dev);
/* Partial dot of one rank (the reference then MPI_Allreduces it,
 * hhp_cuda.c:184-191). */
int hbs_shard_dot_partial(hbs_shard s, const double *a_dev, const double *b_dev, double *out_host);
/* The merged local+halo CSR (m_loc x (m_loc + n_shr)) for SpMV benchmarks. */
int hbs_shard_csr(hbs_shard s, hbs_csr *out);
/* Pack straight to host memory (what the reference then MPI_Isends,
 * hhp_cpu.c:533-547); synchronous. */
int hbs_shard_pack_host(hbs_shard s, const double *x_loc_dev, double *send_host);
/* Shard SpMV for callers that keep x_loc and x_shr in separate device
 * buffers (the reference's dX / dX_shr): y = alpha*(loc*x_loc + shr*x_shr) +
 * beta*y.  With recv_host the halo is copied in, unpacked through recv.J and
 * also written to x_shr_dev (if not NULL); without it x_shr_dev is the halo. */
int hbs_shard_spmv2(hbs_shard s, const double *x_loc_dev, const double *recv_host,
                    double *x_shr_dev, double alpha, double beta, double *y_dev);

/* ---- timing helpers for benchmarks (hipEvents on the context stream) -- */
int hbs_timer_start(hbs_ctx ctx);
int hbs_timer_stop(hbs_ctx ctx, double *ms);
/* Average device time per launch of `reps` back-to-back launches of the
 * SpMV (after one warm-up), from one event pair around all of them. */
int hbs_bench_spmv_f64(hbs_csr a, const double *x_dev, double *y_dev, int reps,
                       double *avg_ms);
int hbs_bench_spmv_f32(hbs_csr a, const float *x_dev, float *y_dev, int reps,
                       double *avg_ms);
int hbs_bench_copy(hbs_ctx ctx, const double *src_dev, double *dst_dev, size_t n,
                   int reps, double *avg_ms);

/* ---- tuning knobs (process-wide, for sweeps) -------------------------- */
/* nnz per row block (a multiple of 256, <= 16384; 0 = by the matrix: 1024
 * under 10 entries per row on average, else within [2048, 6144], sized so
 * the resident SpMV workgroups get equal numbers of blocks) and the average
 * row length above which a block uses several lanes per row instead of the
 * order-exact serial sum. */
int hbs_set_spmv_plan(int block_nnz, int serial_max_avg);
/* Cap on the SpMV workgroups of matrices created afterwards (1..65536; 0,
 * the default: 4096, or for row blocks above 2048 entries as many as are
 * resident on the device at once); each takes a contiguous run of row
 * blocks, and its dot partial is one of those a consumer sums. */
int hbs_set_spmv_groups(int max_groups);
/* Row blocks of matrices created afterwards whose entries, taken in column
 * order, touch at most max_line_pct percent as many 128-byte lines of x as
 * they have entries form their products in column order (a column-sorted
 * copy of their entries), so that a wave's gathers share cache lines; the
 * row sums keep CSR order.  0 disables; default 70. */
int hbs_set_spmv_sort(int max_line_pct);
/* Vector workgroups a solver created afterwards aims at (1..65536, default
 * 512): each takes 512, 1024, 2048 or 4096 elements, the fewest that keep
 * the count at or below the target.  Fewer workgroups mean fewer partials
 * for each consumer to re-sum. */
int hbs_set_vec_blocks(int target);
/* 1: load the matrix streams (col, val) non-temporally (default 0: measured
 * slower on MI355X for every workload tried). */
int hbs_set_spmv_stream_nt(int on);
/* Entries per thread per pass of the 512-thread SpMV workgroups (row blocks
 * above 2048 entries) of matrices created afterwards: 4 or 8, 0 for the
 * default 4. */
int hbs_set_spmv_pass(int entries_per_thread);
/* Plan of a device CSR: SpMV workgroups, long-row chunks, nnz per row
 * block, longest row. */
int hbs_csr_plan_info(hbs_csr a, int *ngroups, int *nchunks, int *block_nnz, int *max_row);
/* x-segmented SpMV for matrices created afterwards.  Its first kernel
 * stages x in segments of 8192 columns in LDS and forms every product there;
 * its second sums them per row, in the same order as the CSR-stream SpMV (so
 * the serial rows stay bit-identical to the sequential CSR sum).  The cost is
 * one more pass over the entries (the products), against gathers that each
 * miss L2 when columns spread over all of x.  mode 0: never; 1: every matrix
 * of one system; 2 (default): one-system matrices with x above 8 MiB (FP64)
 * whose columns touch a new 128-byte line of x for more than half the
 * entries. */
int hbs_set_spmv_xseg(int mode);
/* Workgroups per x segment in the product pass of x-segmented matrices
 * created afterwards (1..16, default 1): each stages the segment's x and
 * forms the products of a contiguous share of its entries, so more
 * workgroups share the GPU at the cost of staging x once per share. */
int hbs_set_xseg_split(int split);
/* Segments of an x-segmented device CSR (0: CSR-stream SpMV). */
int hbs_csr_xseg(hbs_csr a, int *nseg);

/* FP32 SpMVs (hbs_csr_spmv_f32 and the FP32 sweeps of the mixed solve) of
 * square x-segmented device CSRs created afterwards with HBS_CSR_F32: 1
 * (default) keeps the diagonal apart -- the off-diagonal entries form an
 * x-segmented matrix of their own and each row adds d_i x_i (x_i alone while
 * every diagonal entry is 1) to its off-diagonal sum, so the product pass
 * and the row sums move no diagonal entry; 0 keeps it among the entries.
 * FP64 SpMVs always sum every row in CSR order.  hbs_csr_diag_split: 0 not
 * split, 1 split with a unit diagonal, 2 split with diagonal values (also
 * after hbs_csr_set_values on a matrix split with a unit diagonal). */
int hbs_set_spmv_diag_split(int mode);
/* BiCGStab on an x-segmented system (solvers created afterwards): 1
 * (default) updates P and S inside the product passes of the SpMVs that
 * read them -- each x segment's workgroup derives the iteration's scalars
 * from the partials, updates its segment, writes it and stages it -- in
 * place of the separate P and S vector passes; 0 keeps those passes.  The
 * iterates are the same either way; only ||s||^2 (the ||s|| stop test) is
 * summed over other partitions. */
int hbs_set_xseg_fuse(int mode);
int hbs_csr_diag_split(hbs_csr a, int *mode);

/* 3x3-block matrices (every block row's three rows share one list of
 * aligned column triples, at most 256 entries per row; one system, CSR-stream
 * SpMV) of device CSRs created afterwards: 1 (default) reads the values
 * block-major with one column index per 3x3 block, 0 the plain CSR
 * streams.  Each row is summed in CSR order either way. */
int hbs_set_spmv_block3(int mode);
/* Whether a device CSR uses the 3x3-block streams (0 no; 1 yes, every row
 * block's block rows of one length; 2 yes, some row blocks mix block rows
 * of different lengths); whether a host CSR (m x n) is such a 3x3-block
 * matrix. */
int hbs_csr_block3(hbs_csr a, int *on);
int hbs_detect_block3(int m, int n, const int *rowptr, const int *col, int *yes);

/* Row traversal of the multi-column SpMM (hbs_csr_spmm_*, the block solve)
 * for device CSRs created afterwards: 1 (default) sweeps a square matrix
 * that is a 7-point operator on a 3-D grid -- offsets 0, +-1, +-nx, +-nx*ny
 * only, a plane of X larger than an XCD's L2 -- as runs of x-lines in
 * y-slabs, plane after plane, so the X rows a slab gathers stay in L2; 0
 * takes the rows in order.  Each row's sum is the same either way.  No
 * counterpart in the reference (its solver has one right-hand side). */
int hbs_set_spmm_order(int mode);
/* The runs that mode 1 plans for a host CSR (m x m): *nruns run starts in
 * runs[0 .. cap) (error if more), each a run of *len consecutive rows;
 * *nruns == 0 when the matrix is not such a grid. */
int hbs_plan_grid_runs(int m, const int *rowptr, const int *col, int *runs, int cap, int *nruns, int *len);
/* Tuning of the multi-column SpMM: y-lines per slab of the grid runs of
 * device CSRs created afterwards (0: from the L2 size), workgroups per SpMM
 * launch (0: 1024 for 64-column tiles gathered in 16-byte pieces, else 2048;
 * else a multiple of 8), how the 64-column sweeps deal
 * their row steps for device CSRs created afterwards (> 0: each XCD's
 * waves take the next step from one of `dynamic` counters (1..16, default
 * 16), so they work on neighbouring rows; 0: a fixed stride per wave) and whether hbs_csr_spmm_* gathers X in
 * 16-byte pieces where k * element size and X, Y are 16-byte aligned (1,
 * default) or one element per lane (0); the block solve's 64-column SpMMs
 * follow the same switch for workspaces created afterwards.  The multi-column SpMMs of one
 * matrix (hbs_csr_spmm_*, the block solve) must not overlap in time. */
int hbs_set_spmm_tuning(int slab_lines, int groups, int dynamic, int wide_loads);
/* Runs of a device CSR's SpMM traversal (*nruns == 0: rows in order). */
int hbs_csr_spmm_runs(hbs_csr a, int *nruns, int *len);

#ifdef __cplusplus
}
#endif
#endif
