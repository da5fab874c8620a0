/*
 * hhp_oracle.h -- CPU oracle for the BiCGStab / SpMV hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU algorithms (HetHypPar, src/util/hhp_cpu.c, src/util/hhp_matrix.c,
 * src/util/hhp_wrap_bora.c, src/entry/bicgstab_cpu.c, src/entry/bicgstab_gpu.c)
 * used to check the HIP path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only to check results: it is
 * never the thing measured as the GPU product and never a fallback for it.
 *
 * Every function cites the reference file:line it follows.  Arithmetic is
 * FP64 in the reference's operation order; the build compiles this file with
 * -ffp-contract=off so no multiply-add is fused (the reference is built for
 * plain x86-64, which has no FMA instruction).
 *
 * Pinning (what each function is checked against, see tests/test_oracle.py):
 *   orc_mm_read            data/sample{,320}/in/A.mtx: nnz, banner, per-row file order
 *   orc_vector_read        data/sample{,320}/in/B.txt, X_mul.txt
 *   orc_csr_spmv_seq/omp   KAT A*X_mul=[1..16], A*[1..n]=B (sample, sample320),
 *                          dense brute force (numpy) on small random inputs
 *   orc_csr_to_csc /
 *   orc_csc_to_csr         scipy.sparse conversions, round trip
 *   orc_bicgstab           sample320 reaches X=[1..320] (reference outputs
 *                          out/X_seq.txt, X1.txt, X_mpi_gpu.txt to 2e-6),
 *                          scipy.sparse.linalg.bicgstab cross-check on 64x64
 *                          convection-diffusion, x_target recovery
 *   orc_bicgstab_blocked   reduces to orc_bicgstab bit for bit when one block
 *                          holds the vector (an order-sensitivity probe only)
 *   orc_split_row /
 *   orc_loc_split /
 *   orc_setup_comm         invariants: recv.J permutation grouped by ascending
 *                          owner; send(p->q) == locmap_p(gind_q[shr part of
 *                          q's recv block for p]); shard SpMV == global SpMV;
 *                          planted-mistake checks.  No golden halo files exist
 *                          in the reference, so halo lists rest on these.
 *   orc_shard_spmv         equals global CSR SpMV rows within rounding; exact
 *                          when a row has no shared columns
 *   orc_vector_read_parted /
 *   orc_vector_gather      brute force on small inputs; slice then gather is
 *                          the identity over every rank of the reference
 *                          partitions; a planted swap of two gind entries
 *                          must break the round trip
 */
#ifndef HHP_ORACLE_H
#define HHP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same field layout as the reference's COO/CSR/CSC (src/include/hhp_common.h:43-58).
 * CSR: I = row pointers [m+1], J = column indices [nnz].
 * SHARD-convention CSC: I = column pointers [n+1], J = row indices [nnz]. */
typedef struct {
    int *I;
    int *J;
    double *val;
    int nnz;
    int m;
    int n;
} orc_csr;

/* Reference COMM (src/include/hhp_common.h:88-134). */
typedef struct {
    double *val;
    int *J;
    int *I;
    int *ranks;
    int num;
} orc_comm;

/* Reference SHARD_CSC (src/include/hhp_common.h:165-171); gind is iVector. */
typedef struct {
    orc_csr loc;
    orc_csr shr;
    int *gind;
    int ngind;
    orc_comm send;
    orc_comm recv;
} orc_shard;

/* status codes */
#define ORC_OK 0
#define ORC_ERR_IO -1
#define ORC_ERR_FORMAT -2
#define ORC_ERR_ARG -3
#define ORC_ERR_ALLOC -4

int  orc_mm_read(const char *fname, orc_csr *out);
int  orc_vector_read(const char *fname, int size, double *out);
void orc_csr_free(orc_csr *a);

void   orc_csr_spmv_seq(const orc_csr *a, const double *x, double *y);
void   orc_csr_spmv_omp(const orc_csr *a, const double *x, double *y);
void   orc_csr_spmv_seq_acc(const orc_csr *a, const double *x, double *y);
void   orc_csc_spmv_seq_acc(const orc_csr *a, const double *x, double *y);
double orc_vector_dot_seq(const double *a, const double *b, int n);
void   orc_vector_scale_seq(const double *a, double s, double *out, int n);
void   orc_vector_add_seq(const double *a, const double *b, double *out, int n);
void   orc_vector_sub_seq(const double *a, const double *b, double *out, int n);

/* BiCGStab.  fixed_iters != 0 runs exactly maxit iterations with no exit
 * (bicgstab_cpu.c); otherwise the stop test of bicgstab_gpu.c applies with
 * threshold = tol * ||r0||.  x holds x0 on entry and the solution on exit.
 * iters_out = iterations executed (the breaking iteration included).
 * relres_out = sqrt(||b-Ax||^2/||b||^2) as bicgstab_cpu.c:251-255.
 * stop_out: 0 = max iterations, 1 = ||s|| test, 2 = ||r|| test. */
int orc_bicgstab(const orc_csr *a, const double *b, double *x, double tol,
                 int maxit, int fixed_iters, int *iters_out, double *relres_out,
                 int *stop_out);

/* Same loop with every dot re-associated: sums of consecutive blocks of blk
 * elements added in order (blk <= 0: the reference's sequential dot).  Not
 * a restatement of the reference: tests use it only to measure how far an
 * iteration count moves when the dot sums are re-associated, as any
 * parallel reduction does. */
int orc_bicgstab_blocked(const orc_csr *a, const double *b, double *x, double tol,
                         int maxit, int fixed_iters, int blk, int *iters_out,
                         double *relres_out, int *stop_out);

int  orc_csr_to_csc(const orc_csr *csr, orc_csr *csc);
int  orc_shard_csc_to_csr(const orc_csr *csc, orc_csr *csr);

int  orc_split_row(const orc_csr *big, const int *partvec, int nparts, int part,
                   orc_csr *out);
int  orc_loc_split(const orc_csr *part_csr, const int *partvec, int nvals,
                   int myrank, orc_shard *out);
int  orc_build_shards(const orc_csr *big, const int *partvec, int nparts,
                      orc_shard *shards /* [nparts] */);
void orc_shard_free(orc_shard *s);

void orc_vector_read_parted(const double *big, const int *partvec, int n,
                            int rank, double *out);
void orc_vector_gather(int nparts, const int *counts, const double *const *locs, const int *const *ginds,
                       int size, double *big);
void orc_shard_spmv(const orc_shard *s, const double *x_loc,
                    const double *recv_buf, double *y);
void orc_shard_pack(const orc_shard *s, const double *x_loc, double *send_buf);

#ifdef __cplusplus
}
#endif
#endif
