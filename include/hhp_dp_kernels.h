/*
 * hhp_dp_kernels.h -- the reference's device-scalar kernel C-ABI on HIP.
 *
 * Same names, argument order and meaning as the reference header
 * src/include/hhp_dp_kernels.h:14-59 (HetHypPar), with hipStream_t in place
 * of cudaStream_t.  Every pointer is a device pointer, work is queued on `s`,
 * n <= 0 is a no-op.  Implemented by libhbs.so (hand-written gfx950
 * kernels).  The hybrid CPU+GPU helpers of that header (hhp_dp_set_flag,
 * hhp_dp_cu_*) belong to the out-of-scope hybrid solvers and are not
 * provided; hhp_dp_add is.
 *
 * The fused BiCGStab path of this library (hbs_bicgstab_f64) goes further:
 * each of these updates is fused with the dot product that follows it.
 */
#ifndef HHP_DP_KERNELS_H
#define HHP_DP_KERNELS_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* total = a + b  (hhp_dp_kernels.h:14-15) */
void hhp_dp_add(double *total, const double *a, const double *b, hipStream_t s);

/* out[k] = in[idx[k]] for k in [0, nh)  (hhp_dp_kernels.h:20-21) */
void hhp_dp_gather(double *out, const double *in, const int *idx, int nh, hipStream_t s);

/* P = (P - omega*V)*beta + R  (hhp_dp_kernels.h:38-40) */
void hhp_dp_vecop_P(double *P, const double *V, const double *R,
                    const double *omega, const double *beta, int n, hipStream_t s);
/* S = R + neg_alpha*V  (hhp_dp_kernels.h:41-43) */
void hhp_dp_vecop_S(double *S, const double *R, const double *V,
                    const double *neg_alpha, int n, hipStream_t s);
/* X += alpha*P + omega*S ; R = S + neg_omega*T  (hhp_dp_kernels.h:44-47) */
void hhp_dp_vecop_XR(double *X, double *R, const double *P, const double *S,
                     const double *T, const double *alpha, const double *omega,
                     const double *neg_omega, int n, hipStream_t s);

/* beta = (rho_new / rho_old) * (alpha / omega); rho_old <- rho_new
 * (hhp_dp_kernels.h:49-51) */
void hhp_dp_update_beta(double *beta, double *rho_old, const double *rho_new,
                        const double *alpha, const double *omega, hipStream_t s);
/* alpha = rho / rv; neg_alpha = -alpha  (hhp_dp_kernels.h:53-55) */
void hhp_dp_update_alpha(double *alpha, double *neg_alpha,
                         const double *rho, const double *rv, hipStream_t s);
/* omega = ts / tt; neg_omega = -omega  (hhp_dp_kernels.h:57-59) */
void hhp_dp_update_omega(double *omega, double *neg_omega,
                         const double *ts, const double *tt, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif /* HHP_DP_KERNELS_H */
