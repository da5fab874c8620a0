/*
 * hhp_oracle.c -- CPU oracle (test infrastructure only; see hhp_oracle.h).
 *
 * Plain C restatement of the reference's CPU path, in the reference's
 * operation order.  Compiled with -ffp-contract=off (no fused multiply-add).
 */
#include "hhp_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Matrix Market reader                                                */
/* ------------------------------------------------------------------ */

/* Banner parse: src/dependencies/mmio/mmio.c:96-177 (mm_read_banner).
 * Returns 0 and sets the five lower-cased fields, or ORC_ERR_FORMAT. */
static int orc_mm_banner(FILE *f, char *mtx, char *crd, char *dtype, char *storage)
{
    char line[1025], banner[64];
    if (!fgets(line, sizeof line, f)) return ORC_ERR_FORMAT;
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype, storage) != 5)
        return ORC_ERR_FORMAT;
    for (char *p = mtx; *p; p++) *p = (char)tolower(*p);
    for (char *p = crd; *p; p++) *p = (char)tolower(*p);
    for (char *p = dtype; *p; p++) *p = (char)tolower(*p);
    for (char *p = storage; *p; p++) *p = (char)tolower(*p);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0) return ORC_ERR_FORMAT;
    if (strcmp(mtx, "matrix") != 0) return ORC_ERR_FORMAT;
    if (strcmp(crd, "coordinate") != 0 && strcmp(crd, "array") != 0) return ORC_ERR_FORMAT;
    if (strcmp(dtype, "real") != 0 && strcmp(dtype, "complex") != 0 &&
        strcmp(dtype, "pattern") != 0 && strcmp(dtype, "integer") != 0)
        return ORC_ERR_FORMAT;
    if (strcmp(storage, "general") != 0 && strcmp(storage, "symmetric") != 0 &&
        strcmp(storage, "hermitian") != 0 && strcmp(storage, "skew-symmetric") != 0)
        return ORC_ERR_FORMAT;
    return 0;
}

/* Follows src/util/hhp_wrap_bora.c:13-102 (readMatrixMarketFromFile) with the
 * size line of mmio.c:189-218 (mm_read_mtx_crd_size).
 * Accepted: coordinate, real/integer, general (complex, array and symmetric are
 * rejected as at :36-41; pattern is rejected too because the reference's
 * "%d %d %lf" scan cannot read it).  Entries of one row keep FILE order because
 * the fill walks the triplets backwards with --ia[r] (:78-84); duplicates are
 * kept, not summed. */
int orc_mm_read(const char *fname, orc_csr *out)
{
    memset(out, 0, sizeof *out);
    FILE *f = fopen(fname, "r");
    if (!f) return ORC_ERR_IO;
    char mtx[64], crd[64], dtype[64], storage[64];
    if (orc_mm_banner(f, mtx, crd, dtype, storage) != 0) { fclose(f); return ORC_ERR_FORMAT; }
    if (strcmp(dtype, "complex") == 0 || strcmp(crd, "coordinate") != 0 ||
        strcmp(storage, "symmetric") == 0 || strcmp(dtype, "pattern") == 0) {
        fclose(f);
        return ORC_ERR_FORMAT;
    }
    char line[1025];
    int M = 0, N = 0, nz = 0;
    do {
        if (!fgets(line, sizeof line, f)) { fclose(f); return ORC_ERR_FORMAT; }
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %d", &M, &N, &nz) != 3) { fclose(f); return ORC_ERR_FORMAT; }
    if (M < 0 || N < 0 || nz < 0) { fclose(f); return ORC_ERR_FORMAT; }

    int *ri = malloc(sizeof(int) * ((size_t)nz + 1));
    int *ci = malloc(sizeof(int) * ((size_t)nz + 1));
    double *vv = malloc(sizeof(double) * ((size_t)nz + 1));
    int *ia = calloc((size_t)M + 1, sizeof(int));
    int *ja = malloc(sizeof(int) * ((size_t)nz + 1));
    double *va = malloc(sizeof(double) * ((size_t)nz + 1));
    if (!ri || !ci || !vv || !ia || !ja || !va) {
        free(ri); free(ci); free(vv); free(ia); free(ja); free(va); fclose(f);
        return ORC_ERR_ALLOC;
    }
    int rc = ORC_OK;
    for (int i = 0; i < nz; i++) {
        if (fscanf(f, "%d %d %lf\n", &ri[i], &ci[i], &vv[i]) != 3) { rc = ORC_ERR_FORMAT; break; }
        ri[i]--;
        ci[i]--;
        if (ri[i] < 0 || ri[i] >= M || ci[i] < 0 || ci[i] >= N) { rc = ORC_ERR_FORMAT; break; }
        ia[ri[i]]++;
    }
    fclose(f);
    if (rc != ORC_OK) {
        free(ri); free(ci); free(vv); free(ia); free(ja); free(va);
        return rc;
    }
    /* running sum: ia[r] = end of row r (hhp_wrap_bora.c:75-76) */
    for (int i = 1; i <= M; i++) ia[i] += ia[i - 1];
    /* backward fill (hhp_wrap_bora.c:78-84) */
    for (int i = nz - 1; i >= 0; i--) {
        int off = --ia[ri[i]];
        ja[off] = ci[i];
        va[off] = vv[i];
    }
    free(ri); free(ci); free(vv);
    out->I = ia; out->J = ja; out->val = va;
    out->m = M; out->n = N; out->nnz = nz;
    return ORC_OK;
}

/* src/util/hhp_cpu.c:411-428 (vector_read): whitespace separated "%lf",
 * no header.  Returns the count read; a short file is ORC_ERR_FORMAT. */
int orc_vector_read(const char *fname, int size, double *out)
{
    FILE *f = fopen(fname, "r");
    if (!f) return ORC_ERR_IO;
    int i = 0;
    while (i < size && fscanf(f, "%lf", &out[i]) == 1) i++;
    fclose(f);
    return i < size ? ORC_ERR_FORMAT : i;
}

void orc_csr_free(orc_csr *a)
{
    if (!a) return;
    free(a->I); free(a->J); free(a->val);
    memset(a, 0, sizeof *a);
}

/* ------------------------------------------------------------------ */
/* SpMV and vector operations                                          */
/* ------------------------------------------------------------------ */

/* src/util/hhp_cpu.c:445-455 (CSR_spmxv_seq): per row, res = 0.0 and
 * res += val[j]*x[col[j]] in CSR order; y overwritten. */
void orc_csr_spmv_seq(const orc_csr *a, const double *x, double *y)
{
    for (int i = 0; i < a->m; i++) {
        double res = 0.0;
        for (int j = a->I[i]; j < a->I[i + 1]; j++) res += a->val[j] * x[a->J[j]];
        y[i] = res;
    }
}

/* src/util/hhp_cpu.c:459-468 (CSR_spmxv_omp): same per-row arithmetic,
 * rows shared over OpenMP threads with schedule(guided). */
void orc_csr_spmv_omp(const orc_csr *a, const double *x, double *y)
{
    const int m = a->m;
#pragma omp parallel for schedule(guided)
    for (int i = 0; i < m; i++) {
        double res = 0.0;
        for (int j = a->I[i]; j < a->I[i + 1]; j++) res += a->val[j] * x[a->J[j]];
        y[i] = res;
    }
}

/* src/util/hhp_cpu.c:471-480 (CSR_spmxv_seq_acc): y[i] += row sum. */
void orc_csr_spmv_seq_acc(const orc_csr *a, const double *x, double *y)
{
    for (int i = 0; i < a->m; i++) {
        double res = 0.0;
        for (int j = a->I[i]; j < a->I[i + 1]; j++) res += a->val[j] * x[a->J[j]];
        y[i] += res;
    }
}

/* src/util/hhp_cpu.c:483-491 (CSC_spmxv_seq_acc): column scatter,
 * y[row] += val*x[col] column by column. */
void orc_csc_spmv_seq_acc(const orc_csr *a, const double *x, double *y)
{
    for (int c = 0; c < a->n; c++)
        for (int j = a->I[c]; j < a->I[c + 1]; j++) y[a->J[j]] += a->val[j] * x[c];
}

/* src/util/hhp_cpu.c:84-90 */
double orc_vector_dot_seq(const double *a, const double *b, int n)
{
    double res = 0.0;
    for (int i = 0; i < n; i++) res += a[i] * b[i];
    return res;
}

/* src/util/hhp_cpu.c:93-97 */
void orc_vector_scale_seq(const double *a, double s, double *out, int n)
{
    for (int i = 0; i < n; i++) out[i] = a[i] * s;
}

/* src/util/hhp_cpu.c:100-105 */
void orc_vector_add_seq(const double *a, const double *b, double *out, int n)
{
    for (int i = 0; i < n; i++) out[i] = a[i] + b[i];
}

/* src/util/hhp_cpu.c:108-113 */
void orc_vector_sub_seq(const double *a, const double *b, double *out, int n)
{
    for (int i = 0; i < n; i++) out[i] = a[i] - b[i];
}

/* ------------------------------------------------------------------ */
/* BiCGStab                                                            */
/* ------------------------------------------------------------------ */

/* Loop body and init: src/entry/bicgstab_cpu.c:172-244; final relative
 * residual: :251-255.  Stop test (fixed_iters == 0) from the reference's
 * gpu_BiCGStab, src/entry/bicgstab_gpu.c: threshold = tol*||r0|| (:148-150),
 * break when ||s|| < threshold after the alpha step with x already advanced by
 * alpha*p (:219-236), or when ||r|| < threshold after the omega step
 * (:289-293).  Norms are sqrt of the sequential dot. */
/* Dot product as sums of consecutive blocks of blk elements (each block
 * sequential), the block sums added in order; blk < 0: -blk interleaved
 * running sums; blk == 0 is the reference's sequential dot. */
static double dot_blocked(const double *x, const double *y, int n, int blk)
{
    if (blk < 0) {
        /* interleaved: -blk running sums, entry i into sum i mod -blk (the
         * way lanes of a parallel reduction stride through a vector), the
         * sums then added in order */
        const int k = -blk;
        double *acc = calloc((size_t)k, sizeof(double));
        if (!acc) return orc_vector_dot_seq(x, y, n);
        for (int i = 0; i < n; i++) acc[i % k] += x[i] * y[i];
        double total = 0.0;
        for (int j = 0; j < k; j++) total += acc[j];
        free(acc);
        return total;
    }
    if (blk == 0) return orc_vector_dot_seq(x, y, n);
    double total = 0.0;
    for (int i0 = 0; i0 < n; i0 += blk) {
        const int len = n - i0 < blk ? n - i0 : blk;
        total += orc_vector_dot_seq(x + i0, y + i0, len);
    }
    return total;
}

static int bicgstab_impl(const orc_csr *a, const double *b, double *x, double tol,
                         int maxit, int fixed_iters, int blk, int *iters_out,
                         double *relres_out, int *stop_out);

int orc_bicgstab(const orc_csr *a, const double *b, double *x, double tol,
                 int maxit, int fixed_iters, int *iters_out, double *relres_out,
                 int *stop_out)
{
    return bicgstab_impl(a, b, x, tol, maxit, fixed_iters, 0, iters_out, relres_out, stop_out);
}

int orc_bicgstab_blocked(const orc_csr *a, const double *b, double *x, double tol,
                         int maxit, int fixed_iters, int blk, int *iters_out,
                         double *relres_out, int *stop_out)
{
    return bicgstab_impl(a, b, x, tol, maxit, fixed_iters, blk, iters_out, relres_out, stop_out);
}

static int bicgstab_impl(const orc_csr *a, const double *b, double *x, double tol,
                         int maxit, int fixed_iters, int blk, int *iters_out,
                         double *relres_out, int *stop_out)
{
    const int n = a->m;
    if (n <= 0 || a->n != n) return ORC_ERR_ARG;
    double *Y = calloc(n, sizeof(double));
    double *V = calloc(n, sizeof(double));
    double *P = calloc(n, sizeof(double));
    double *R = malloc(sizeof(double) * n);
    double *R0 = malloc(sizeof(double) * n);
    double *S = malloc(sizeof(double) * n);
    double *T = malloc(sizeof(double) * n);
    if (!Y || !V || !P || !R || !R0 || !S || !T) {
        free(Y); free(V); free(P); free(R); free(R0); free(S); free(T);
        return ORC_ERR_ALLOC;
    }
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    memcpy(R, b, sizeof(double) * n);
    orc_csr_spmv_seq(a, x, Y);
    orc_vector_sub_seq(R, Y, R, n);
    memcpy(R0, R, sizeof(double) * n);
    const double thr = tol * sqrt(dot_blocked(R0, R0, n, blk));

    int it = 0, stop = 0;
    for (it = 0; it < maxit; it++) {
        double temp_rho = dot_blocked(R, R0, n, blk);
        double beta = (temp_rho / rho) * (alpha / omega);
        rho = temp_rho;
        orc_vector_scale_seq(V, omega, V, n);
        orc_vector_sub_seq(P, V, P, n);
        orc_vector_scale_seq(P, beta, P, n);
        orc_vector_add_seq(P, R, P, n);
        orc_csr_spmv_seq(a, P, V);
        alpha = rho / dot_blocked(R0, V, n, blk);
        orc_vector_scale_seq(V, alpha, S, n);
        orc_vector_sub_seq(R, S, S, n);
        if (!fixed_iters && sqrt(dot_blocked(S, S, n, blk)) < thr) {
            orc_vector_scale_seq(P, alpha, Y, n);
            orc_vector_add_seq(x, Y, x, n);
            stop = 1;
            it++;
            break;
        }
        orc_csr_spmv_seq(a, S, T);
        omega = dot_blocked(T, S, n, blk) / dot_blocked(T, T, n, blk);
        orc_vector_scale_seq(P, alpha, Y, n);
        orc_vector_add_seq(x, Y, x, n);
        orc_vector_scale_seq(S, omega, Y, n);
        orc_vector_add_seq(x, Y, x, n);
        orc_vector_scale_seq(T, omega, R, n);
        orc_vector_sub_seq(S, R, R, n);
        if (!fixed_iters && sqrt(dot_blocked(R, R, n, blk)) < thr) {
            stop = 2;
            it++;
            break;
        }
    }
    orc_csr_spmv_seq(a, x, Y);
    orc_vector_sub_seq(b, Y, Y, n);
    double sy = orc_vector_dot_seq(Y, Y, n);
    double sb = orc_vector_dot_seq(b, b, n);
    if (iters_out) *iters_out = it;
    if (relres_out) *relres_out = sqrt(sy / sb);
    if (stop_out) *stop_out = stop;
    free(Y); free(V); free(P); free(R); free(R0); free(S); free(T);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Format conversions                                                  */
/* ------------------------------------------------------------------ */

/* src/util/hhp_matrix.c:1476-1535 (CSR_to_CSC): column counts, prefix sum,
 * rows visited in ascending order so row indices ascend within a column.
 * Output uses the SHARD convention (I = column pointers, J = row indices).
 * The reference aborts on nnz == 0 (:1484-1486); this returns ORC_ERR_ARG. */
int orc_csr_to_csc(const orc_csr *csr, orc_csr *csc)
{
    memset(csc, 0, sizeof *csc);
    if (csr->nnz == 0) return ORC_ERR_ARG;
    csc->nnz = csr->nnz; csc->m = csr->m; csc->n = csr->n;
    csc->I = calloc((size_t)csr->n + 1, sizeof(int));
    csc->J = malloc(sizeof(int) * csr->nnz);
    csc->val = malloc(sizeof(double) * csr->nnz);
    int *cnt = calloc((size_t)csr->n + 1, sizeof(int));
    if (!csc->I || !csc->J || !csc->val || !cnt) { free(cnt); orc_csr_free(csc); return ORC_ERR_ALLOC; }
    for (int i = 0; i < csr->nnz; i++) csc->I[csr->J[i] + 1]++;
    for (int i = 1; i <= csr->n; i++) csc->I[i] += csc->I[i - 1];
    for (int i = 0; i < csr->m; i++) {
        for (int j = csr->I[i]; j < csr->I[i + 1]; j++) {
            int c = csr->J[j];
            int d = csc->I[c] + cnt[c]++;
            csc->J[d] = i;
            csc->val[d] = csr->val[j];
        }
    }
    free(cnt);
    return ORC_OK;
}

/* src/util/hhp_matrix.c:1441-1474 (SHARD_loc_CSC_to_CSR): columns visited in
 * ascending order, so column indices ascend within each CSR row. */
int orc_shard_csc_to_csr(const orc_csr *csc, orc_csr *csr)
{
    memset(csr, 0, sizeof *csr);
    csr->nnz = csc->nnz; csr->m = csc->m; csr->n = csc->n;
    csr->I = calloc((size_t)csc->m + 1, sizeof(int));
    csr->J = malloc(sizeof(int) * (csc->nnz > 0 ? csc->nnz : 1));
    csr->val = malloc(sizeof(double) * (csc->nnz > 0 ? csc->nnz : 1));
    int *cnt = calloc((size_t)csc->m + 1, sizeof(int));
    if (!csr->I || !csr->J || !csr->val || !cnt) { free(cnt); orc_csr_free(csr); return ORC_ERR_ALLOC; }
    for (int c = 0; c < csc->n; c++)
        for (int j = csc->I[c]; j < csc->I[c + 1]; j++) cnt[csc->J[j]]++;
    csr->I[0] = 0;
    for (int i = 0; i < csc->m; i++) { csr->I[i + 1] = csr->I[i] + cnt[i]; cnt[i] = 0; }
    for (int c = 0; c < csc->n; c++) {
        for (int j = csc->I[c]; j < csc->I[c + 1]; j++) {
            int r = csc->J[j];
            int d = csr->I[r] + cnt[r]++;
            csr->J[d] = c;
            csr->val[d] = csc->val[j];
        }
    }
    free(cnt);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Row partition, loc/shr split, communication lists                   */
/* ------------------------------------------------------------------ */

/* src/util/hhp_matrix.c:820-894 (internal_CSR_split_row), for one part:
 * rows of `part` in ascending global row, entries copied in row order,
 * columns stay global; requires one partvec entry per row (== big->n). */
int orc_split_row(const orc_csr *big, const int *partvec, int nparts, int part,
                  orc_csr *out)
{
    memset(out, 0, sizeof *out);
    out->n = big->n;
    for (int i = 0; i < big->n; i++) {
        int p = partvec[i];
        if (p < 0 || p >= nparts) return ORC_ERR_ARG;   /* :840-842 aborts */
        if (p == part) { out->m++; out->nnz += big->I[i + 1] - big->I[i]; }
    }
    out->I = malloc(sizeof(int) * ((size_t)out->m + 1));
    out->J = malloc(sizeof(int) * (out->nnz > 0 ? out->nnz : 1));
    out->val = malloc(sizeof(double) * (out->nnz > 0 ? out->nnz : 1));
    if (!out->I || !out->J || !out->val) { orc_csr_free(out); return ORC_ERR_ALLOC; }
    out->I[0] = 0;
    int row = 0, pos = 0;
    for (int i = 0; i < big->n; i++) {
        if (partvec[i] != part) continue;
        int size = big->I[i + 1] - big->I[i];
        out->I[row + 1] = pos + size;
        memcpy(out->J + pos, big->J + big->I[i], sizeof(int) * size);
        memcpy(out->val + pos, big->val + big->I[i], sizeof(double) * size);
        row++;
        pos += size;
    }
    return ORC_OK;
}

/* src/util/hhp_matrix.c:896-981 (internal_CSR_loc_split): CSC over global
 * columns; owned columns (ascending) -> loc, non-owned columns with at least
 * one entry (ascending) -> shr; gind = [owned globals, shr globals]. */
int orc_loc_split(const orc_csr *in, const int *partvec, int nvals, int myrank,
                  orc_shard *res)
{
    memset(res, 0, sizeof *res);
    res->loc.m = in->m;
    res->shr.m = in->m;
    orc_csr t;
    int rc = orc_csr_to_csc(in, &t);
    if (rc != ORC_OK) return rc;
    for (int i = 0; i < nvals; i++) {
        int cnt = t.I[i + 1] - t.I[i];
        if (partvec[i] == myrank) { res->loc.nnz += cnt; res->loc.n++; continue; }
        if (cnt == 0) continue;
        res->shr.nnz += cnt;
        res->shr.n++;
    }
    if (res->loc.n != res->loc.m) { orc_csr_free(&t); return ORC_ERR_ARG; }  /* :923-924 */
    res->loc.I = calloc((size_t)res->loc.n + 1, sizeof(int));
    res->loc.J = malloc(sizeof(int) * (res->loc.nnz > 0 ? res->loc.nnz : 1));
    res->loc.val = malloc(sizeof(double) * (res->loc.nnz > 0 ? res->loc.nnz : 1));
    res->shr.I = calloc((size_t)res->shr.n + 1, sizeof(int));
    res->shr.J = malloc(sizeof(int) * (res->shr.nnz > 0 ? res->shr.nnz : 1));
    res->shr.val = malloc(sizeof(double) * (res->shr.nnz > 0 ? res->shr.nnz : 1));
    res->ngind = res->loc.n + res->shr.n;
    res->gind = malloc(sizeof(int) * (res->ngind > 0 ? res->ngind : 1));
    int nl = 0, ns = 0;
    for (int i = 0; i < nvals; i++) {
        int cnt = t.I[i + 1] - t.I[i];
        int isloc = partvec[i] == myrank;
        if (isloc) res->gind[nl++] = i;
        else if (cnt != 0) res->gind[res->loc.n + ns++] = i;
        else continue;
        orc_csr *mm = isloc ? &res->loc : &res->shr;
        int col = isloc ? nl : ns;
        int off = mm->I[col - 1];
        mm->I[col] = off + cnt;
        for (int j = 0; j < cnt; j++) {
            mm->J[off + j] = t.J[t.I[i] + j];
            mm->val[off + j] = t.val[t.I[i] + j];
        }
    }
    orc_csr_free(&t);
    return ORC_OK;
}

/* recv side of src/util/hhp_matrix.c:1003-1071: owners with a non-zero count
 * in ascending rank; recv.J[block(owner)+k] = index of the k-th shr column
 * (ascending) owned by that owner. */
static int orc_setup_recv(orc_shard *s, const int *partvec, int nparts, int myrank)
{
    int *cnt = calloc((size_t)nparts + 1, sizeof(int));
    if (!cnt) return ORC_ERR_ALLOC;
    for (int i = 0; i < s->shr.n; i++) {
        int owner = partvec[s->gind[s->loc.n + i]];
        if (owner == myrank) { free(cnt); return ORC_ERR_ARG; }
        cnt[owner + 1]++;
    }
    orc_comm r = {0};
    for (int i = 1; i <= nparts; i++) {
        if (cnt[i] != 0) r.num++;
        cnt[i] += cnt[i - 1];
    }
    r.I = calloc((size_t)r.num + 1, sizeof(int));
    r.ranks = calloc(r.num > 0 ? r.num : 1, sizeof(int));
    r.J = calloc(s->shr.n > 0 ? s->shr.n : 1, sizeof(int));
    r.val = calloc(s->shr.n > 0 ? s->shr.n : 1, sizeof(double));
    int num = 0;
    for (int i = 0; i < nparts; i++) {
        if (cnt[i + 1] - cnt[i] == 0) continue;
        r.ranks[num] = i;
        r.I[num + 1] = cnt[i + 1];
        num++;
    }
    int *cnt2 = calloc((size_t)nparts, sizeof(int));
    for (int i = 0; i < s->shr.n; i++) {
        int owner = partvec[s->gind[s->loc.n + i]];
        r.J[cnt[owner] + cnt2[owner]] = i;
        cnt2[owner]++;
    }
    free(cnt2);
    free(cnt);
    s->recv = r;
    return ORC_OK;
}

/* send side of src/util/hhp_matrix.c:1077-1239, with the MPI exchange done
 * in one process: the peers q with a non-zero count, ascending; the block for
 * q holds the global indices of q's recv block for `me` (q's recv_gJ,
 * :1167-1177), converted to local indices through locmap (:1181-1185,
 * :1230-1232). */
static int orc_setup_send(orc_shard *shards, int nparts, int me, int nglobal)
{
    orc_shard *s = &shards[me];
    int *count = calloc((size_t)nparts, sizeof(int));
    int *first = calloc((size_t)nparts, sizeof(int));
    if (!count || !first) { free(count); free(first); return ORC_ERR_ALLOC; }
    for (int q = 0; q < nparts; q++) {
        if (q == me) continue;
        const orc_comm *rq = &shards[q].recv;
        for (int k = 0; k < rq->num; k++) {
            if (rq->ranks[k] == me) {
                count[q] = rq->I[k + 1] - rq->I[k];
                first[q] = rq->I[k];
            }
        }
    }
    orc_comm sc = {0};
    int stotal = 0;
    for (int q = 0; q < nparts; q++) {
        stotal += count[q];
        if (count[q] != 0) sc.num++;
    }
    sc.I = calloc((size_t)sc.num + 1, sizeof(int));
    sc.ranks = calloc(sc.num > 0 ? sc.num : 1, sizeof(int));
    sc.J = calloc(stotal > 0 ? stotal : 1, sizeof(int));
    sc.val = calloc(stotal > 0 ? stotal : 1, sizeof(double));
    int *locmap = malloc(sizeof(int) * (nglobal > 0 ? nglobal : 1));
    for (int i = 0; i < nglobal; i++) locmap[i] = -1;
    for (int i = 0; i < s->loc.n; i++) locmap[s->gind[i]] = i;
    int num = 0;
    for (int q = 0; q < nparts; q++) {
        if (count[q] == 0) continue;
        sc.ranks[num] = q;
        sc.I[num + 1] = sc.I[num] + count[q];
        const orc_shard *sq = &shards[q];
        for (int k = 0; k < count[q]; k++) {
            int lind = sq->recv.J[first[q] + k];
            int g = sq->gind[sq->loc.n + lind];
            sc.J[sc.I[num] + k] = locmap[g];
        }
        num++;
    }
    free(locmap); free(count); free(first);
    s->send = sc;
    return ORC_OK;
}

/* All ranks of src/util/hhp_matrix.c:1249-1338 (MPI_CSR_split_row) in one
 * process. */
int orc_build_shards(const orc_csr *big, const int *partvec, int nparts, orc_shard *shards)
{
    memset(shards, 0, sizeof(orc_shard) * nparts);
    for (int p = 0; p < nparts; p++) {
        orc_csr part;
        int rc = orc_split_row(big, partvec, nparts, p, &part);
        if (rc != ORC_OK) return rc;
        rc = orc_loc_split(&part, partvec, big->n, p, &shards[p]);
        orc_csr_free(&part);
        if (rc != ORC_OK) return rc;
        rc = orc_setup_recv(&shards[p], partvec, nparts, p);
        if (rc != ORC_OK) return rc;
    }
    for (int p = 0; p < nparts; p++) {
        int rc = orc_setup_send(shards, nparts, p, big->n);
        if (rc != ORC_OK) return rc;
    }
    return ORC_OK;
}

static void orc_comm_free(orc_comm *c)
{
    free(c->val); free(c->J); free(c->I); free(c->ranks);
    memset(c, 0, sizeof *c);
}

void orc_shard_free(orc_shard *s)
{
    if (!s) return;
    orc_csr_free(&s->loc);
    orc_csr_free(&s->shr);
    free(s->gind);
    orc_comm_free(&s->send);
    orc_comm_free(&s->recv);
    memset(s, 0, sizeof *s);
}

/* src/util/hhp_cpu.c:252-334 (MPI_vector_read_parted): the local slice holds
 * the entries of `rank` in ascending global index. */
void orc_vector_read_parted(const double *big, const int *partvec, int n, int rank, double *out)
{
    int k = 0;
    for (int i = 0; i < n; i++)
        if (partvec[i] == rank) out[k++] = big[i];
}

/* src/util/hhp_cpu.c:199-250 (MPI_vector_gather), every rank in one call:
 * the root's two Gathervs concatenate the ranks' local values and their gind
 * prefixes in rank order (counts[r] entries each, displacements the running
 * sums), then big[gmap[i]] = temp_big[i] for i in order. */
void orc_vector_gather(int nparts, const int *counts, const double *const *locs, const int *const *ginds,
                       int size, double *big)
{
    int total = 0;
    for (int r = 0; r < nparts; r++) total += counts[r];
    double *temp_big = malloc(sizeof(double) * (total > 0 ? total : 1));
    int *gmap = malloc(sizeof(int) * (total > 0 ? total : 1));
    int d = 0;
    for (int r = 0; r < nparts; r++) {
        for (int k = 0; k < counts[r]; k++) {
            temp_big[d + k] = locs[r][k];
            gmap[d + k] = ginds[r][k];
        }
        d += counts[r];
    }
    for (int i = 0; i < total && i < size; i++) big[gmap[i]] = temp_big[i];
    free(temp_big);
    free(gmap);
}

/* src/util/hhp_cpu.c:514-601 (MPI_SHARD_CSC_mpi_spmxv_seq) with the halo
 * already received: y = 0; y += loc*x (column scatter, :563); then every recv
 * block in rank order (the MPI_Waitany arrival order taken as ascending), for
 * each entry i: ii = recv.J[i], y[shr.J[j]] += shr.val[j]*recv_buf[i]
 * (:574-584).  x_shr is therefore addressed through recv.J, never by identity
 * (the GPU path of the reference, hhp_cuda.c:323-329, skips that mapping). */
void orc_shard_spmv(const orc_shard *s, const double *x_loc, const double *recv_buf, double *y)
{
    memset(y, 0, sizeof(double) * (s->loc.m > 0 ? s->loc.m : 0));
    orc_csc_spmv_seq_acc(&s->loc, x_loc, y);
    for (int b = 0; b < s->recv.num; b++) {
        for (int i = s->recv.I[b]; i < s->recv.I[b + 1]; i++) {
            int ii = s->recv.J[i];
            double v = recv_buf[i];
            for (int j = s->shr.I[ii]; j < s->shr.I[ii + 1]; j++)
                y[s->shr.J[j]] += s->shr.val[j] * v;
        }
    }
}

/* src/util/hhp_cpu.c:533-538: send.val[i] = X[send.J[i]]. */
void orc_shard_pack(const orc_shard *s, const double *x_loc, double *send_buf)
{
    int ns = s->send.num > 0 ? s->send.I[s->send.num] : 0;
    for (int i = 0; i < ns; i++) send_buf[i] = x_loc[s->send.J[i]];
}
