/*
 * A plain-C caller of the reference-shaped device API (include/hhp_hip.h,
 * include/hhp_dp_kernels.h), in the order of the reference SpMV harness
 * (spmv_gpu.c:90-136): read A and x, device_csr_create, vectors to the
 * device, Y zeroed, the SpMV buffer, one warm-up SpMV, n timed SpMVs, Y back
 * to the host, and the largest |Y_i - (i + 1)| printed as the reference
 * prints it.  It also exercises device_csr_spmv_async, device_vector_dot /
 * axpy / scale / GPUtoGPU and hhp_dp_add on the handle's stream.
 *
 * usage: spmv_caller A.mtx x.txt n_iters     (exit status 0 on success)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "hhp_dp_kernels.h"
#include "hhp_hip.h"

#define CHECK(call)                                                                  \
    do {                                                                             \
        int rc_ = (call);                                                            \
        if (rc_ != 0) {                                                              \
            fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #call, rc_, \
                    hbs_last_error());                                               \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s A.mtx x.txt n_iters\n", argv[0]);
        return 2;
    }
    const int niters = atoi(argv[3]);
    const double mx_alpha = 1.0, mx_beta = 0.0;
    hhp_handle_t h = hhp_default_handle();
    if (!h) {
        fprintf(stderr, "no device\n");
        return 1;
    }

    CSR A;
    CHECK(hbs_mm_read(argv[1], &A.m, &A.n, &A.nnz, &A.I, &A.J, &A.val));
    Device_CSR dA;
    CHECK(device_csr_create(A, &dA));

    Vector Xt = {malloc(sizeof(double) * (size_t)A.n), (unsigned)A.n};
    CHECK(hbs_vector_read(argv[2], A.n, Xt.vals));
    Device_Vector X, Y;
    CHECK(device_vector_init(Xt.nvals, &X));
    CHECK(device_vector_toGPU(Xt, X));
    CHECK(device_vector_init((size_t)A.m, &Y));
    CHECK(device_vector_zero(Y));

    Device_Buffer_SpMV buf;
    CHECK(device_buffer_spmv_create(h, dA.desc, X, Y, &mx_alpha, &mx_beta, &buf));
    CHECK(device_csr_spmv(h, dA, X, Y, mx_alpha, mx_beta, buf));   /* warm-up */
    CHECK(hbs_timer_start(h));
    for (int i = 0; i < niters; i++) CHECK(device_csr_spmv(h, dA, X, Y, mx_alpha, mx_beta, buf));
    double ms = 0.0;
    CHECK(hbs_timer_stop(h, &ms));

    Vector Yh = {malloc(sizeof(double) * (size_t)A.m), (unsigned)A.m};
    CHECK(device_vector_toCPU(Y, Yh));
    double max_abs_err = 0.0;
    for (unsigned i = 0; i < Yh.nvals; i++) {
        double err = fabs(Yh.vals[i] - (double)(i + 1));
        if (err > max_abs_err) max_abs_err = err;
    }

    /* asynchronous form, Y2 = 2*A*X + 0.5*Y; the caller synchronises */
    Device_Vector Y2;
    CHECK(device_vector_init((size_t)A.m, &Y2));
    CHECK(device_vector_GPUtoGPU(Y, Y2));
    CHECK(device_csr_spmv_async(h, dA, X, Y2, 2.0, 0.5, buf));
    CHECK(hbs_ctx_sync(h));
    Vector Y2h = {malloc(sizeof(double) * (size_t)A.m), (unsigned)A.m};
    CHECK(device_vector_toCPU(Y2, Y2h));
    double async_err = 0.0;
    for (unsigned i = 0; i < Y2h.nvals; i++) {
        double err = fabs(Y2h.vals[i] - (2.0 * Yh.vals[i] + 0.5 * Yh.vals[i]));
        if (err > async_err) async_err = err;
    }

    /* BLAS-1 on the handle: dot, Y2 += -1*Y, Y2 *= 2 */
    double dot = 0.0;
    CHECK(device_vector_dot(h, Y, Y, &dot));
    CHECK(device_vector_axpy(h, Y, -1.0, Y2));
    CHECK(device_vector_scale(h, 2.0, Y2));
    CHECK(device_vector_toCPU(Y2, Y2h));
    double blas_err = 0.0, dot_ref = 0.0;
    for (unsigned i = 0; i < Y2h.nvals; i++) {
        double err = fabs(Y2h.vals[i] - 2.0 * (2.5 * Yh.vals[i] - Yh.vals[i]));
        if (err > blas_err) blas_err = err;
        dot_ref += Yh.vals[i] * Yh.vals[i];
    }

    /* device-scalar kernel C-ABI on the handle's stream: total = a + b */
    Device_Vector sc;
    CHECK(device_vector_init(3, &sc));
    double s3[3] = {1.25, -0.5, 0.0};
    Vector s3h = {s3, 3};
    CHECK(device_vector_toGPU(s3h, sc));
    hhp_dp_add(sc.vals + 2, sc.vals, sc.vals + 1, (hipStream_t)hbs_ctx_stream(h));
    CHECK(hbs_ctx_sync(h));
    CHECK(device_vector_toCPU(sc, s3h));

    printf("Matrix name : %s\n", argv[1]);
    printf("n_iters : %d \n", niters);
    printf("spmv_time_per_iter : %.9f\n", ms * 1e-3 / (niters > 0 ? niters : 1));
    printf("max_abs_err_vs_1ton : %.6e\n", max_abs_err);
    printf("async_err : %.6e\n", async_err);
    printf("blas_err : %.6e\n", blas_err);
    printf("dot_rel_err : %.6e\n", fabs(dot - dot_ref) / (dot_ref > 0 ? dot_ref : 1.0));
    printf("dp_add : %.17g\n", s3[2]);

    CHECK(device_vector_destroy(&sc));
    CHECK(device_vector_destroy(&Y2));
    CHECK(device_vector_destroy(&X));
    CHECK(device_vector_destroy(&Y));
    CHECK(device_buffer_spmv_destroy(&buf));
    CHECK(device_csr_destroy(&dA));
    hbs_host_free(A.I);
    hbs_host_free(A.J);
    hbs_host_free(A.val);
    free(Xt.vals);
    free(Yh.vals);
    free(Y2h.vals);
    return 0;
}
