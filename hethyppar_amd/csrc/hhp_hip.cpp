// hhp_hip.cpp -- the reference-shaped device API (hhp_hip.h, hhp_dp_kernels.h)
// as thin wrappers over the native C-ABI (hbs.h) and the HIP launchers.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "hbs_err.h"
#include "hbs_internal.h"
#include "hhp_dp_kernels.h"
#include "hhp_hip.h"

namespace {

std::mutex g_mu;
hbs_ctx g_default = nullptr;

int dv_check(const Device_Vector &v) { return v.vals || v.nvals == 0 ? HBS_OK : HBS_ERR_ARG; }

// Host CSR copy of a SHARD-convention CSC (column pointers in I, row indices
// in J): columns visited in ascending order (as SHARD_loc_CSC_to_CSR,
// hhp_matrix.c:1441-1474).
void csc_to_csr(const CSC &in, std::vector<int> &rp, std::vector<int> &col, std::vector<double> &val)
{
    rp.assign((size_t)in.m + 1, 0);
    col.resize((size_t)in.nnz);
    val.resize((size_t)in.nnz);
    for (int c = 0; c < in.n; ++c)
        for (int j = in.I[c]; j < in.I[c + 1]; ++j) rp[(size_t)in.J[j] + 1]++;
    for (int i = 0; i < in.m; ++i) rp[(size_t)i + 1] += rp[(size_t)i];
    std::vector<int> fill(rp.begin(), rp.end() - 1);
    for (int c = 0; c < in.n; ++c)
        for (int j = in.I[c]; j < in.I[c + 1]; ++j) {
            const int d = fill[(size_t)in.J[j]]++;
            col[(size_t)d] = c;
            val[(size_t)d] = in.val[j];
        }
}

int wrap_csr(hbs_csr h, Device_CSR *out)
{
    const int *rp = nullptr, *c = nullptr;
    const double *v = nullptr;
    int m = 0, n = 0, nnz = 0;
    hbs_csr_device_arrays(h, &rp, &c, &v);
    hbs_csr_info(h, &m, &n, &nnz, nullptr, nullptr);
    out->desc = h;
    out->data.I = const_cast<int *>(rp);
    out->data.J = const_cast<int *>(c);
    out->data.val = const_cast<double *>(v);
    out->data.m = m;
    out->data.n = n;
    out->data.nnz = nnz;
    return HBS_OK;
}

}  // namespace

extern "C" hhp_handle_t hhp_default_handle(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_default && hbs_ctx_create(0, nullptr, &g_default) != HBS_OK) g_default = nullptr;
    return g_default;
}

extern "C" int device_csr_create(CSR in, Device_CSR *out)
{
    if (!out) return hbs::set_error(HBS_ERR_ARG, "null out");
    hbs_ctx ctx = hhp_default_handle();
    if (!ctx) return HBS_ERR_NODEVICE;
    hbs_csr h = nullptr;
    int rc = hbs_csr_create(ctx, in.m, in.n, in.nnz, in.I, in.J, in.val, 0, &h);
    if (rc) return rc;
    return wrap_csr(h, out);
}

extern "C" int device_csr_destroy(Device_CSR *in)
{
    if (!in || !in->desc) return hbs::set_error(HBS_ERR_ARG, "Attempted to free NULLED device CSR");
    int rc = hbs_csr_destroy(in->desc);
    std::memset(in, 0, sizeof *in);
    return rc;
}

extern "C" int device_csc_create(CSC in, Device_CSC *out)
{
    if (!out || (in.nnz > 0 && (!in.I || !in.J || !in.val)) || in.m < 0 || in.n < 0)
        return hbs::set_error(HBS_ERR_ARG, "bad CSC");
    hbs_ctx ctx = hhp_default_handle();
    if (!ctx) return HBS_ERR_NODEVICE;
    std::vector<int> rp, col;
    std::vector<double> val;
    csc_to_csr(in, rp, col, val);
    hbs_csr h = nullptr;
    int rc = hbs_csr_create(ctx, in.m, in.n, in.nnz, rp.data(), col.data(), val.data(), 0, &h);
    if (rc) return rc;
    return wrap_csr(h, out);
}

extern "C" int device_csc_destroy(Device_CSC *in) { return device_csr_destroy(in); }

extern "C" int device_buffer_spmv_create(hhp_handle_t, hbs_csr, Device_Vector, Device_Vector, const double *,
                                         const double *, Device_Buffer_SpMV *out)
{
    if (out) *out = nullptr;   // the CSR-stream SpMV needs no external buffer
    return HBS_OK;
}

extern "C" int device_buffer_spmv_destroy(Device_Buffer_SpMV *in)
{
    if (in) *in = nullptr;
    return HBS_OK;
}

extern "C" int device_vector_init(size_t size, Device_Vector *out)
{
    if (!out) return hbs::set_error(HBS_ERR_ARG, "null out");
    hbs_ctx ctx = hhp_default_handle();
    if (!ctx) return HBS_ERR_NODEVICE;
    void *p = nullptr;
    int rc = hbs_dev_alloc(ctx, size * sizeof(double), &p);
    if (rc) return rc;
    out->desc = nullptr;
    out->vals = static_cast<double *>(p);
    out->nvals = (unsigned)size;
    return HBS_OK;
}

extern "C" int device_vector_destroy(Device_Vector *in)
{
    if (!in || !in->vals) return hbs::set_error(HBS_ERR_ARG, "Attempted to free NULL Device vector");
    int rc = hbs_dev_free(hhp_default_handle(), in->vals);
    in->vals = nullptr;
    in->nvals = 0;
    return rc;
}

extern "C" int device_vector_zero(Device_Vector vec)
{
    if (dv_check(vec)) return hbs::set_error(HBS_ERR_ARG, "bad vector");
    hbs_ctx ctx = hhp_default_handle();
    int rc = hbs_dev_zero(ctx, vec.vals, sizeof(double) * vec.nvals);
    return rc ? rc : hbs_ctx_sync(ctx);
}

extern "C" int device_vector_dot(hhp_handle_t h, Device_Vector v1, Device_Vector v2, double *out)
{
    if (!h) h = hhp_default_handle();
    return hbs_dot_f64(h, (int)v1.nvals, v1.vals, v2.vals, out);
}

extern "C" int device_vector_axpy(hhp_handle_t h, Device_Vector v1, const double a, Device_Vector out)
{
    if (!h) h = hhp_default_handle();
    return hbs_axpy_f64(h, (int)v1.nvals, a, v1.vals, out.vals);
}

extern "C" int device_vector_scale(hhp_handle_t h, const double alpha, Device_Vector v)
{
    if (!h) h = hhp_default_handle();
    return hbs_scal_f64(h, (int)v.nvals, alpha, v.vals);
}

extern "C" int device_vector_toGPU(Vector src, Device_Vector dest)
{
    return hbs_h2d(hhp_default_handle(), dest.vals, src.vals, sizeof(double) * src.nvals);
}

extern "C" int device_vector_toCPU(Device_Vector src, Vector dest)
{
    return hbs_d2h(hhp_default_handle(), dest.vals, src.vals, sizeof(double) * src.nvals);
}

extern "C" int device_vector_GPUtoGPU(Device_Vector src, Device_Vector dst)
{
    hbs_ctx ctx = hhp_default_handle();
    int rc = hbs_d2d(ctx, dst.vals, src.vals, sizeof(double) * src.nvals);
    return rc ? rc : hbs_ctx_sync(ctx);
}

extern "C" int device_csr_spmv_async(hhp_handle_t, Device_CSR mat, Device_Vector X, Device_Vector Y,
                                     const double alpha, const double beta, Device_Buffer_SpMV)
{
    return hbs_csr_spmv_f64(mat.desc, alpha, X.vals, beta, Y.vals);
}

extern "C" int device_csr_spmv(hhp_handle_t h, Device_CSR mat, Device_Vector X, Device_Vector Y,
                               const double alpha, const double beta, Device_Buffer_SpMV buf)
{
    int rc = device_csr_spmv_async(h, mat, X, Y, alpha, beta, buf);
    return rc ? rc : hbs_ctx_sync(hhp_default_handle());
}

extern "C" int device_csc_spmv(hhp_handle_t h, Device_CSC mat, Device_Vector X, Device_Vector Y,
                               const double alpha, const double beta, Device_Buffer_SpMV buf)
{
    return device_csr_spmv(h, mat, X, Y, alpha, beta, buf);
}

extern "C" int device_SHARD_CSC_create(SHARD_CSC A, Device_SHARD_CSC *out)
{
    if (!out) return hbs::set_error(HBS_ERR_ARG, "null out");
    std::memset(out, 0, sizeof *out);
    hbs_ctx ctx = hhp_default_handle();
    if (!ctx) return HBS_ERR_NODEVICE;
    int rc = hbs_shard_create(ctx, &A, &out->dev);
    if (rc) return rc;
    out->gind = A.gind;
    out->send = A.send;
    out->recv = A.recv;
    hbs_csr merged = nullptr;
    hbs_shard_csr(out->dev, &merged);
    wrap_csr(merged, &out->loc);
    out->loc.data.n = A.loc.n;
    out->shr.data.m = A.shr.m;
    out->shr.data.n = A.shr.n;
    out->shr.data.nnz = A.shr.nnz;
    return HBS_OK;
}

extern "C" int device_SHARD_CSC_destroy(Device_SHARD_CSC *A)
{
    if (!A) return HBS_OK;
    int rc = hbs_shard_destroy(A->dev);
    std::memset(A, 0, sizeof *A);
    return rc;
}

namespace {

// The shard built the reference driver's way (bicgstab_mpi_gpu.c:248-257:
// device_csc_create of A.loc and A.shr, gind/send/recv borrowed, no merged
// shard): the two SpMVs of MPI_device_SHARD_CSC_mpi_spmxv (hhp_cuda.c:
// 259-352), Y = 0; Y = alpha*loc*dX + beta*Y; Y = alpha*shr*x_shr + alpha*Y,
// with the halo placed by recv.J (x_shr[recv.J[p]] = recv.val[p], the CPU
// semantics of hhp_cpu.c:574-584) and A.send.val = X[send.J] packed on the
// host from a copy of dX, as the reference does.
int shard_spmxv_two_pass(const Device_SHARD_CSC &A, Device_Vector dX, Device_Vector dX_shr, Device_Vector Y,
                         double alpha, double beta, SpMV_Profile *prof)
{
    if (!A.loc.desc) return hbs::set_error(HBS_ERR_ARG, "shard has neither a merged form nor a loc block");
    hbs_ctx ctx = hhp_default_handle();
    if (!ctx) return HBS_ERR_NODEVICE;
    const int nloc = A.loc.data.n, m = A.loc.data.m, nshr = A.shr.desc ? A.shr.data.n : 0;
    if ((int)dX.nvals < nloc || (int)Y.nvals < m || (nshr > 0 && (int)dX_shr.nvals < nshr))
        return hbs::set_error(HBS_ERR_ARG, "vector shorter than the shard");
    const int nsend = A.send.num > 0 ? A.send.I[A.send.num] : 0;
    const int nrecv = A.recv.num > 0 ? A.recv.I[A.recv.num] : 0;
    if (nrecv != nshr) return hbs::set_error(HBS_ERR_ARG, "recv list has %d entries for %d halo columns", nrecv, nshr);
    double ms = 0.0;
    int rc = hbs_dev_zero(ctx, Y.vals, sizeof(double) * (size_t)m);
    if (rc) return rc;
    if (nsend > 0) {
        if (prof) hbs_timer_start(ctx);
        std::vector<double> xh((size_t)nloc);
        if ((rc = hbs_d2h(ctx, xh.data(), dX.vals, sizeof(double) * (size_t)nloc))) return rc;
        for (int i = 0; i < nsend; ++i) {
            const int k = A.send.J[i];
            if (k < 0 || k >= nloc) return hbs::set_error(HBS_ERR_ARG, "send.J[%d] = %d out of range", i, k);
            A.send.val[i] = xh[(size_t)k];
        }
        if (prof) {
            hbs_timer_stop(ctx, &ms);
            prof->send_fill += ms * 1e-3;
        }
    }
    if (prof) hbs_timer_start(ctx);
    if ((rc = hbs_csr_spmv_f64(A.loc.desc, alpha, dX.vals, beta, Y.vals))) return rc;
    if (prof) {
        hbs_timer_stop(ctx, &ms);
        prof->local_spmv += ms * 1e-3;
    }
    if (nshr > 0) {
        if (prof) hbs_timer_start(ctx);
        std::vector<double> xs((size_t)nshr);
        for (int p = 0; p < nrecv; ++p) {
            const int k = A.recv.J[p];
            if (k < 0 || k >= nshr) return hbs::set_error(HBS_ERR_ARG, "recv.J[%d] = %d out of range", p, k);
            xs[(size_t)k] = A.recv.val[p];
        }
        if ((rc = hbs_h2d(ctx, dX_shr.vals, xs.data(), sizeof(double) * (size_t)nshr))) return rc;
        if ((rc = hbs_csr_spmv_f64(A.shr.desc, alpha, dX_shr.vals, alpha, Y.vals))) return rc;
        if (prof) {
            hbs_timer_stop(ctx, &ms);
            prof->shared_spmv += ms * 1e-3;
        }
    }
    return hbs_ctx_sync(ctx);
}

}  // namespace

extern "C" int device_SHARD_CSC_spmxv(Device_SHARD_CSC A, Vector, Device_Vector dX, Device_Vector dX_shr,
                                      Device_Vector Y, hhp_handle_t, const double alpha, const double beta,
                                      Device_Buffer_SpMV, Device_Buffer_SpMV, SpMV_Profile *prof)
{
    if (!A.dev) return shard_spmxv_two_pass(A, dX, dX_shr, Y, alpha, beta, prof);
    hbs_ctx ctx = hhp_default_handle();
    int nsend = A.send.num > 0 ? A.send.I[A.send.num] : 0;
    if (prof) hbs_timer_start(ctx);
    int rc = nsend ? hbs_shard_pack_host(A.dev, dX.vals, A.send.val) : HBS_OK;
    if (rc) return rc;
    double ms = 0.0;
    if (prof) {
        hbs_timer_stop(ctx, &ms);
        prof->send_fill += ms * 1e-3;
        hbs_timer_start(ctx);
    }
    rc = hbs_shard_spmv2(A.dev, dX.vals, A.recv.val, dX_shr.vals, alpha, beta, Y.vals);
    if (rc) return rc;
    if (prof) {
        hbs_timer_stop(ctx, &ms);
        prof->local_spmv += ms * 1e-3;
    }
    return hbs_ctx_sync(ctx);
}

// ---- hhp_dp_kernels.h ----------------------------------------------------

extern "C" void hhp_dp_add(double *total, const double *a, const double *b, hipStream_t s)
{
    (void)hbs::launch_dp_scalar(3, total, nullptr, a, b, nullptr, s);
}

extern "C" void hhp_dp_gather(double *out, const double *in, const int *idx, int nh, hipStream_t s)
{
    (void)hbs::launch_gather_f64(nh, in, idx, out, s);
}

extern "C" void hhp_dp_vecop_P(double *P, const double *V, const double *R, const double *omega,
                               const double *beta, int n, hipStream_t s)
{
    (void)hbs::launch_dp_vecop_P(P, V, R, omega, beta, n, s);
}

extern "C" void hhp_dp_vecop_S(double *S, const double *R, const double *V, const double *neg_alpha, int n,
                               hipStream_t s)
{
    (void)hbs::launch_dp_vecop_S(S, R, V, neg_alpha, n, s);
}

extern "C" void hhp_dp_vecop_XR(double *X, double *R, const double *P, const double *S, const double *T,
                                const double *alpha, const double *omega, const double *neg_omega, int n,
                                hipStream_t s)
{
    (void)hbs::launch_dp_vecop_XR(X, R, P, S, T, alpha, omega, neg_omega, n, s);
}

extern "C" void hhp_dp_update_beta(double *beta, double *rho_old, const double *rho_new, const double *alpha,
                                   const double *omega, hipStream_t s)
{
    (void)hbs::launch_dp_scalar(0, beta, rho_old, rho_new, alpha, omega, s);
}

extern "C" void hhp_dp_update_alpha(double *alpha, double *neg_alpha, const double *rho, const double *rv,
                                    hipStream_t s)
{
    (void)hbs::launch_dp_scalar(1, alpha, neg_alpha, rho, rv, nullptr, s);
}

extern "C" void hhp_dp_update_omega(double *omega, double *neg_omega, const double *ts, const double *tt,
                                    hipStream_t s)
{
    (void)hbs::launch_dp_scalar(2, omega, neg_omega, ts, tt, nullptr, s);
}
