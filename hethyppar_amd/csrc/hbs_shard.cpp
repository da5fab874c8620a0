// hbs_shard.cpp -- one rank's share on the device: the local and halo blocks
// merged into one CSR, halo pack/unpack through the reference's send.J /
// recv.J lists, and the share's SpMV with the halo from a host buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <cstring>
#include <vector>

#include "hbs_objects.h"

using namespace hbs;

#define fail ::hbs::set_error

// ---------------------------------------------------------------------------
// one rank's share

struct hbs_shard_s {
    hbs_ctx ctx = nullptr;
    int m_loc = 0, n_shr = 0, n_send = 0;
    hbs_csr merged = nullptr;     // m_loc x (m_loc + n_shr)
    int *send_J = nullptr;        // device [n_send]
    int *recv_src = nullptr;      // device [n_shr]: recv position of shr column c
    double *stage = nullptr;      // device [n_shr]: halo as received
    double *x_ext = nullptr;      // device [m_loc + n_shr], for split x_loc / x_shr callers
    double *d_send = nullptr;     // device [n_send], packed halo before D2H
};

extern "C" int hbs_shard_destroy(hbs_shard s)
{
    if (!s) return HBS_OK;
    DevGuard g(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->merged) hbs_csr_destroy(s->merged);
    dfree(s->send_J);
    dfree(s->recv_src);
    dfree(s->stage);
    dfree(s->x_ext);
    dfree(s->d_send);
    delete s;
    return HBS_OK;
}

// Upload of a SHARD_CSC.  loc and shr become one CSR over the extended
// vector [x_loc ; x_shr]: row i holds its loc entries in ascending local
// column, then its shr entries in the order the reference's CPU halo SpMV
// adds them (MPI_SHARD_CSC_mpi_spmxv_seq, hhp_cpu.c:563-584: the loc column
// scatter first, then recv blocks in rank order, entries in recv order), so
// the order-exact row sum reproduces that function when blocks arrive in
// rank order.
extern "C" int hbs_shard_create(hbs_ctx ctx, const hbs_host_shard *h, hbs_shard *out)
{
    if (!ctx || !h || !out) return fail(HBS_ERR_ARG, "null argument");
    *out = nullptr;
    const int m = h->loc.m;
    const int ns = h->shr.n;
    if (m <= 0 || h->loc.n != m || h->shr.m != m || ns < 0) return fail(HBS_ERR_ARG, "inconsistent shard shape");
    if (!h->loc.I || (ns > 0 && (!h->shr.I || !h->recv.J || !h->recv.I))) return fail(HBS_ERR_ARG, "null shard array");
    const int nrecv = h->recv.num > 0 ? h->recv.I[h->recv.num] : 0;
    if (nrecv != ns) return fail(HBS_ERR_ARG, "recv list covers %d of %d shr columns", nrecv, ns);
    const int nsend = h->send.num > 0 ? h->send.I[h->send.num] : 0;
    std::vector<int> cnt((size_t)m + 1, 0);
    for (int c = 0; c < m; ++c)
        for (int j = h->loc.I[c]; j < h->loc.I[c + 1]; ++j) {
            const int r = h->loc.J[j];
            if (r < 0 || r >= m) return fail(HBS_ERR_ARG, "loc row index out of range");
            cnt[(size_t)r + 1]++;
        }
    std::vector<int> pos((size_t)ns, -1);
    for (int p = 0; p < ns; ++p) {
        const int c = h->recv.J[p];
        if (c < 0 || c >= ns || pos[(size_t)c] >= 0) return fail(HBS_ERR_ARG, "recv.J is not a permutation of the shr columns");
        pos[(size_t)c] = p;
    }
    for (int c = 0; c < ns; ++c)
        for (int j = h->shr.I[c]; j < h->shr.I[c + 1]; ++j) {
            const int r = h->shr.J[j];
            if (r < 0 || r >= m) return fail(HBS_ERR_ARG, "shr row index out of range");
            cnt[(size_t)r + 1]++;
        }
    for (int i = 0; i < m; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
    const int nnz = cnt[(size_t)m];
    std::vector<int> col((size_t)std::max(nnz, 1));
    std::vector<double> val((size_t)std::max(nnz, 1));
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int c = 0; c < m; ++c)
        for (int j = h->loc.I[c]; j < h->loc.I[c + 1]; ++j) {
            const int d = fill[(size_t)h->loc.J[j]]++;
            col[(size_t)d] = c;
            val[(size_t)d] = h->loc.val[j];
        }
    for (int p = 0; p < ns; ++p) {
        const int c = h->recv.J[p];
        for (int j = h->shr.I[c]; j < h->shr.I[c + 1]; ++j) {
            const int d = fill[(size_t)h->shr.J[j]]++;
            col[(size_t)d] = m + c;
            val[(size_t)d] = h->shr.val[j];
        }
    }
    for (int i = 0; i < nsend; ++i)
        if (h->send.J[i] < 0 || h->send.J[i] >= m) return fail(HBS_ERR_ARG, "send.J entry out of range");
    DevGuard g(ctx->device);
    hbs_shard s = new hbs_shard_s();
    s->ctx = ctx;
    s->m_loc = m;
    s->n_shr = ns;
    s->n_send = nsend;
    auto bail = [&](int rc) {
        std::string e = hbs_last_error();
        hbs_shard_destroy(s);
        set_error(rc, "%s", e.c_str());
        return rc;
    };
    int rc = hbs_csr_create(ctx, m, m + ns, nnz, cnt.data(), col.data(), val.data(), 0, &s->merged);
    if (rc) return bail(rc);
    if ((rc = dalloc(&s->send_J, nsend)) || (rc = dalloc(&s->recv_src, ns)) || (rc = dalloc(&s->stage, ns)))
        return bail(rc);
    if ((nsend && hipMemcpy(s->send_J, h->send.J, sizeof(int) * nsend, hipMemcpyHostToDevice) != hipSuccess) ||
        (ns && hipMemcpy(s->recv_src, pos.data(), sizeof(int) * ns, hipMemcpyHostToDevice) != hipSuccess))
        return bail(fail(HBS_ERR_HIP, "shard list upload failed"));
    *out = s;
    return HBS_OK;
}

extern "C" int hbs_shard_info(hbs_shard s, int *m_loc, int *n_shr, int *n_send, int *nnz)
{
    if (!s) return fail(HBS_ERR_ARG, "null shard");
    if (m_loc) *m_loc = s->m_loc;
    if (n_shr) *n_shr = s->n_shr;
    if (n_send) *n_send = s->n_send;
    if (nnz) *nnz = s->merged->nnz;
    return HBS_OK;
}

extern "C" int hbs_shard_pack(hbs_shard s, const double *x_loc, double *send)
{
    if (!s || (s->n_send > 0 && (!x_loc || !send))) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    HIPCHK(launch_gather_f64(s->n_send, x_loc, s->send_J, send, s->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_shard_unpack(hbs_shard s, const double *recv, double *x_ext)
{
    if (!s || (s->n_shr > 0 && (!recv || !x_ext))) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    HIPCHK(launch_gather_f64(s->n_shr, recv, s->recv_src, x_ext + s->m_loc, s->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_shard_spmv(hbs_shard s, double *x_ext, const double *recv_host, double *y)
{
    if (!s || !x_ext || !y) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    if (recv_host && s->n_shr > 0) {
        HIPCHK(hipMemcpyAsync(s->stage, recv_host, sizeof(double) * s->n_shr, hipMemcpyHostToDevice, s->ctx->stream));
        CHK(hbs_shard_unpack(s, s->stage, x_ext));
    }
    return hbs_csr_spmv_f64(s->merged, 1.0, x_ext, 0.0, y);
}

extern "C" int hbs_shard_dot_partial(hbs_shard s, const double *a, const double *b, double *out)
{
    if (!s) return fail(HBS_ERR_ARG, "null shard");
    return hbs_dot_f64(s->ctx, s->m_loc, a, b, out);
}

extern "C" int hbs_shard_csr(hbs_shard s, hbs_csr *out)
{
    if (!s || !out) return fail(HBS_ERR_ARG, "null argument");
    *out = s->merged;
    return HBS_OK;
}

namespace {
int shard_scratch(hbs_shard s)
{
    if (!s->x_ext) CHK(dalloc(&s->x_ext, (size_t)s->m_loc + s->n_shr));
    if (!s->d_send) CHK(dalloc(&s->d_send, (size_t)std::max(s->n_send, 1)));
    return HBS_OK;
}
}  // namespace

extern "C" int hbs_shard_pack_host(hbs_shard s, const double *x_loc, double *send_host)
{
    if (!s || (s->n_send > 0 && (!x_loc || !send_host))) return fail(HBS_ERR_ARG, "bad argument");
    if (s->n_send == 0) return HBS_OK;
    DevGuard g(s->ctx->device);
    CHK(shard_scratch(s));
    CHK(hbs_shard_pack(s, x_loc, s->d_send));
    HIPCHK(hipMemcpyAsync(send_host, s->d_send, sizeof(double) * s->n_send, hipMemcpyDeviceToHost,
                          s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return HBS_OK;
}

extern "C" int hbs_shard_spmv2(hbs_shard s, const double *x_loc, const double *recv_host, double *x_shr,
                               double alpha, double beta, double *y)
{
    if (!s || !x_loc || !y) return fail(HBS_ERR_ARG, "bad argument");
    DevGuard g(s->ctx->device);
    CHK(shard_scratch(s));
    hipStream_t q = s->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->x_ext, x_loc, sizeof(double) * s->m_loc, hipMemcpyDeviceToDevice, q));
    if (s->n_shr > 0) {
        if (recv_host) {
            HIPCHK(hipMemcpyAsync(s->stage, recv_host, sizeof(double) * s->n_shr, hipMemcpyHostToDevice, q));
            CHK(hbs_shard_unpack(s, s->stage, s->x_ext));
            if (x_shr)
                HIPCHK(hipMemcpyAsync(x_shr, s->x_ext + s->m_loc, sizeof(double) * s->n_shr,
                                      hipMemcpyDeviceToDevice, q));
        } else {
            if (!x_shr) return fail(HBS_ERR_ARG, "no halo: pass recv_host or x_shr");
            HIPCHK(hipMemcpyAsync(s->x_ext + s->m_loc, x_shr, sizeof(double) * s->n_shr,
                                  hipMemcpyDeviceToDevice, q));
        }
    }
    return hbs_csr_spmv_f64(s->merged, alpha, s->x_ext, beta, y);
}
