/*
 * hhp_hip.h -- the reference's device API (src/include/hhp_cuda.h:46-65 and
 * the types of src/include/hhp_common.h:7-180) on HIP, over libhbs.so.
 *
 * A reference caller keeps its code: the structs have the reference's field
 * layout, the functions keep their names, argument order and meaning, and
 * return 0 on success like the cuSPARSE/cuBLAS status they replace (negative
 * HBS_ERR_* codes on failure, never exit()).  Differences, by design:
 *   - the cusparseHandle_t / cublasHandle_t arguments become hhp_handle_t,
 *     an hbs_ctx (device + stream); functions without a handle in the
 *     reference use a process-wide context on device 0 (hhp_default_handle);
 *   - SpMV needs no external buffer: device_buffer_spmv_create succeeds and
 *     returns NULL;
 *   - the MPI rank functions lose their MPI_Comm: device_vector_dot returns
 *     the rank's partial (the caller Allreduces), and device_SHARD_CSC_spmxv
 *     takes the halo already received in A.recv.val (recv order) and leaves
 *     the packed send values in A.send.val.
 */
#ifndef HHP_HIP_H
#define HHP_HIP_H

#include "hbs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {            /* hhp_common.h:7-10 */
    double *vals;
    unsigned int nvals;
} Vector;
typedef hbs_host_ivec iVector;   /* hhp_common.h:12-15 */
typedef hbs_host_csr COO;        /* hhp_common.h:43-50 */
typedef COO CSR;                 /* hhp_common.h:54 */
typedef COO CSC;                 /* hhp_common.h:58 */
typedef hbs_host_comm COMM;      /* hhp_common.h:88-134 */
typedef hbs_host_shard SHARD_CSC;  /* hhp_common.h:165-171 */

typedef hbs_ctx hhp_handle_t;

typedef struct {            /* hhp_common.h:17-21; desc unused (NULL) */
    void *desc;
    double *vals;
    unsigned int nvals;
} Device_Vector;

typedef void *Device_Buffer_SpMV;   /* hhp_common.h:23 */

typedef struct {            /* hhp_common.h:25-32 */
    double send_fill;
    double local_spmv;
    double comm_wait;
    double shared_spmv;
    double send_wait;
} SpMV_Profile;

/* hhp_common.h:60-68.  desc holds the library handle; data holds the device
 * arrays of the matrix in CSR form (for a Device_CSC: the CSR of the same
 * matrix, built from the SHARD-convention CSC at creation) and its m, n, nnz. */
typedef struct {
    hbs_csr desc;
    CSR data;
} Device_CSR;
typedef Device_CSR Device_CSC;

/* hhp_common.h:174-180, plus the library's shard handle at the end. */
typedef struct {
    Device_CSC loc;
    Device_CSC shr;
    iVector gind;
    COMM send;
    COMM recv;
    hbs_shard dev;
} Device_SHARD_CSC;

/* Process-wide context on device 0 (created on first use). */
hhp_handle_t hhp_default_handle(void);

/* hhp_cuda.h:46-49 */
int device_csr_create(CSR in, Device_CSR *out);
int device_csr_destroy(Device_CSR *in);
int device_csc_create(CSC in, Device_CSC *out);
int device_csc_destroy(Device_CSC *in);
/* hhp_cuda.h:50-51 */
int device_buffer_spmv_create(hhp_handle_t handle, hbs_csr mat, Device_Vector X, Device_Vector Y,
                              const double *alpha, const double *beta, Device_Buffer_SpMV *out);
int device_buffer_spmv_destroy(Device_Buffer_SpMV *in);
/* hhp_cuda.h:52-54 */
int device_vector_init(size_t size, Device_Vector *out);
int device_vector_destroy(Device_Vector *in);
int device_vector_zero(Device_Vector vec);
/* hhp_cuda.h:55 (device_vector_dot); the rank's partial of hhp_cuda.h:56 */
int device_vector_dot(hhp_handle_t handle, Device_Vector v1, Device_Vector v2, double *out);
/* hhp_cuda.h:57-58: out += a*v1 ; v *= alpha */
int device_vector_axpy(hhp_handle_t handle, Device_Vector v1, const double a, Device_Vector out);
int device_vector_scale(hhp_handle_t handle, const double alpha, Device_Vector v);
/* hhp_cuda.h:59-61 */
int device_vector_toGPU(Vector src, Device_Vector dest);
int device_vector_toCPU(Device_Vector src, Vector dest);
int device_vector_GPUtoGPU(Device_Vector src, Device_Vector dst);
/* hhp_cuda.h:62-64: Y = alpha*A*X + beta*Y; the synchronous forms wait for
 * the handle's stream like the reference's cudaDeviceSynchronize */
int device_csr_spmv(hhp_handle_t handle, Device_CSR mat, Device_Vector X, Device_Vector Y,
                    const double alpha, const double beta, Device_Buffer_SpMV buf);
int device_csr_spmv_async(hhp_handle_t handle, Device_CSR mat, Device_Vector X, Device_Vector Y,
                          const double alpha, const double beta, Device_Buffer_SpMV buf);
int device_csc_spmv(hhp_handle_t handle, Device_CSC mat, Device_Vector X, Device_Vector Y,
                    const double alpha, const double beta, Device_Buffer_SpMV buf);

/* Device copy of one rank's SHARD_CSC (the reference builds it inline at
 * bicgstab_mpi_gpu.c:248-257); gind/send/recv are borrowed from A. */
int device_SHARD_CSC_create(SHARD_CSC A, Device_SHARD_CSC *out);
int device_SHARD_CSC_destroy(Device_SHARD_CSC *A);
/* MPI_device_SHARD_CSC_mpi_spmxv (hhp_cuda.c:259-352) without the MPI
 * transport: A.send.val <- dX[send.J] (for the caller to send); the halo is
 * A.recv.val (recv order, as MPI delivered it), placed by x_shr[recv.J[p]] =
 * A.recv.val[p] (the CPU semantics of hhp_cpu.c:574-584); dX_shr receives
 * x_shr.  X is unused (kept for the reference signature); prof may be NULL.
 *   - A from device_SHARD_CSC_create (A.dev set): one pass over the merged
 *     local+halo rows, Y = alpha*(loc*dX + shr*x_shr) + beta*Y, each row
 *     summed loc entries first, then shr entries in recv order -- bit for
 *     bit the CPU MPI_SHARD_CSC_mpi_spmxv_seq (hhp_cpu.c:514-601) when
 *     alpha = 1 and beta = 0, the coefficients every reference driver
 *     passes.
 *   - A built as the reference driver builds it (bicgstab_mpi_gpu.c:248-257:
 *     device_csc_create of A.loc and, if it has columns, A.shr; A.dev NULL):
 *     the reference's two SpMVs, Y = 0, Y = alpha*loc*dX + beta*Y, then
 *     Y = alpha*shr*x_shr + alpha*Y (so alpha reaches the loc part twice,
 *     as in the reference). */
int device_SHARD_CSC_spmxv(Device_SHARD_CSC A, Vector X, Device_Vector dX, Device_Vector dX_shr,
                           Device_Vector Y, hhp_handle_t handle, const double alpha,
                           const double beta, Device_Buffer_SpMV locbuf, Device_Buffer_SpMV shrbuf,
                           SpMV_Profile *prof);

#ifdef __cplusplus
}
#endif

#endif /* HHP_HIP_H */
