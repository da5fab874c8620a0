// hbs_layout.cpp -- host-side layout: Matrix Market reader, vector files, and
// one rank's shard (row split, loc/shr split, halo send/recv lists).
//
// Behaviour follows the reference's L2 layer (src/util/hhp_wrap_bora.c,
// src/util/hhp_matrix.c, src/util/hhp_cpu.c); the algorithms are this
// library's own: a single in-memory parse for the reader, and for the shard a
// direct construction of one rank's lists without the MPI round trip.
#include <algorithm>
#include <cctype>
#include <climits>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "hbs.h"
#include "hbs_err.h"

namespace {

int lfail(int code)
{
    const char *what = code == HBS_ERR_IO       ? "cannot open file"
                       : code == HBS_ERR_FORMAT ? "malformed or unsupported file"
                       : code == HBS_ERR_ALLOC  ? "host allocation failed"
                       : code == HBS_ERR_PARTITION
                           ? "partition vector out of range, or a rank without rows or entries"
                           : "bad argument";
    return hbs::set_error(code, "%s", what);
}

bool read_file(const char *fname, std::string &buf)
{
    FILE *f = std::fopen(fname, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
        std::fclose(f);
        return false;
    }
    buf.resize((size_t)sz);
    size_t got = sz ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    return got == (size_t)sz;
}

template <typename T>
T *hmalloc(size_t n)
{
    return static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
}

constexpr int kMaxThreads = 64;   // host threads of the parallel builders

// Split items [0, n) into T contiguous runs of about equal weight, given the
// weight prefix pre[0..n] (pre[i] = weight of the items before i).  Returns
// the T + 1 run bounds.
template <typename W>
std::vector<int> runs(const W *pre, int n, int T)
{
    std::vector<int> b((size_t)T + 1, n);
    b[0] = 0;
    const double total = (double)(pre[n] - pre[0]);
    for (int t = 1; t < T; ++t) {
        const double want = (double)pre[0] + total * t / T;
        b[(size_t)t] = (int)(std::lower_bound(pre, pre + n + 1, want,
                                              [](W a, double w) { return (double)a < w; }) -
                             pre);
        b[(size_t)t] = std::min(std::max(b[(size_t)t], b[(size_t)t - 1]), n);
    }
    return b;
}

// Parse one line "r c v" of [p, e) into (r-1, c-1, v); false unless the line
// holds exactly those three numbers within range.
bool parse_line(const char *p, const char *e, long M, long N, int &r0, int &c0, double &v)
{
    char *q = nullptr;
    const long r = std::strtol(p, &q, 10);
    if (q == p || q > e) return false;
    p = q;
    const long c = std::strtol(p, &q, 10);
    if (q == p || q > e) return false;
    p = q;
    v = std::strtod(p, &q);
    if (q == p || q > e) return false;
    for (p = q; p < e; ++p)
        if (!std::isspace((unsigned char)*p)) return false;
    if (r < 1 || r > M || c < 1 || c > N) return false;
    r0 = (int)(r - 1);
    c0 = (int)(c - 1);
    return true;
}

// The triplet section [p, end) parsed by several threads, each a run of whole
// lines, the runs concatenated in file order -- the same values (strtol /
// strtod, as the reference's fscanf) in the same order as one pass.  Taken
// only for large files laid out one entry per line (blank lines allowed);
// false otherwise, and the caller parses the token stream in one pass.
bool parse_triplets_parallel(const char *p, const char *end, long M, long N, long NZ, std::vector<int> &ri,
                             std::vector<int> &ci, std::vector<double> &vv)
{
    const size_t len = (size_t)(end - p);
    int T = omp_get_max_threads();
    if (T > kMaxThreads) T = kMaxThreads;
    if (T < 2 || NZ < 100000 || len < ((size_t)1 << 20)) return false;
    std::vector<const char *> cut((size_t)T + 1);
    cut[0] = p;
    cut[(size_t)T] = end;
    for (int t = 1; t < T; ++t) {
        const char *c = p + len * (size_t)t / (size_t)T;
        const char *nl = static_cast<const char *>(std::memchr(c, '\n', (size_t)(end - c)));
        cut[(size_t)t] = nl ? nl + 1 : end;
        if (cut[(size_t)t] < cut[(size_t)t - 1]) cut[(size_t)t] = cut[(size_t)t - 1];
    }
    std::vector<std::vector<int>> lr((size_t)T), lc((size_t)T);
    std::vector<std::vector<double>> lv((size_t)T);
    std::vector<int> ok((size_t)T, 1);
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        const char *a = cut[(size_t)t], *b = cut[(size_t)t + 1];
        const size_t guess = (size_t)(b - a) / 16 + 16;
        lr[(size_t)t].reserve(guess);
        lc[(size_t)t].reserve(guess);
        lv[(size_t)t].reserve(guess);
        while (a < b) {
            const char *nl = static_cast<const char *>(std::memchr(a, '\n', (size_t)(b - a)));
            const char *e = nl ? nl : b;
            const char *f = a;
            while (f < e && std::isspace((unsigned char)*f)) ++f;
            if (f < e) {
                int r0, c0;
                double v;
                if (!parse_line(f, e, M, N, r0, c0, v)) {
                    ok[(size_t)t] = 0;
                    break;
                }
                lr[(size_t)t].push_back(r0);
                lc[(size_t)t].push_back(c0);
                lv[(size_t)t].push_back(v);
            }
            a = nl ? nl + 1 : b;
        }
    }
    size_t total = 0;
    for (int t = 0; t < T; ++t) {
        if (!ok[(size_t)t]) return false;
        total += lr[(size_t)t].size();
    }
    if (total < (size_t)NZ) return false;
    size_t k = 0;
    for (int t = 0; t < T && k < (size_t)NZ; ++t)
        for (size_t i = 0; i < lr[(size_t)t].size() && k < (size_t)NZ; ++i, ++k) {
            ri[k] = lr[(size_t)t][i];
            ci[k] = lc[(size_t)t][i];
            vv[k] = lv[(size_t)t][i];
        }
    return true;
}

}  // namespace

extern "C" void hbs_host_free(void *p) { std::free(p); }

// Matrix Market coordinate reader.  Semantics of the reference
// readMatrixMarketFromFile (hhp_wrap_bora.c:13-102): banner checked as
// mm_read_banner (mmio.c:96-177); complex, array and symmetric are refused
// (:36-41), and pattern too since its entries carry no value; comment lines
// skipped before the size line (mmio.c:189-218); 1-based triplets; within a
// row the entries keep FILE order and duplicates are kept.  Here the file is
// read in one piece, large line-structured files parsed by several threads
// (runs of whole lines, concatenated in file order), and the triplets
// bucketed by a stable counting sort.
extern "C" int hbs_mm_read(const char *fname, int *m, int *n, int *nnz, int **rowptr, int **col,
                           double **val)
{
    if (!fname || !m || !n || !nnz || !rowptr || !col || !val) return lfail(HBS_ERR_ARG);
    std::string buf;
    if (!read_file(fname, buf)) return lfail(HBS_ERR_IO);
    const char *p = buf.c_str();
    const char *end = p + buf.size();
    // banner line
    const char *eol = std::strchr(p, '\n');
    std::string line(p, eol ? eol : end);
    char banner[64] = {0}, mtx[64] = {0}, crd[64] = {0}, dtype[64] = {0}, storage[64] = {0};
    if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype, storage) != 5)
        return lfail(HBS_ERR_FORMAT);
    auto lower = [](char *s) {
        for (; *s; ++s) *s = (char)std::tolower((unsigned char)*s);
    };
    lower(mtx);
    lower(crd);
    lower(dtype);
    lower(storage);
    if (std::strncmp(banner, "%%MatrixMarket", 14) != 0) return lfail(HBS_ERR_FORMAT);
    if (std::strcmp(mtx, "matrix") != 0 || std::strcmp(crd, "coordinate") != 0) return lfail(HBS_ERR_FORMAT);
    if (std::strcmp(dtype, "real") != 0 && std::strcmp(dtype, "integer") != 0) return lfail(HBS_ERR_FORMAT);
    if (std::strcmp(storage, "general") != 0 && std::strcmp(storage, "hermitian") != 0 &&
        std::strcmp(storage, "skew-symmetric") != 0)
        return lfail(HBS_ERR_FORMAT);
    p = eol ? eol + 1 : end;
    // skip comment lines, then the size line
    while (p < end && *p == '%') {
        const char *e = std::strchr(p, '\n');
        p = e ? e + 1 : end;
    }
    char *q = nullptr;
    long M = std::strtol(p, &q, 10);
    if (q == p) return lfail(HBS_ERR_FORMAT);
    p = q;
    long N = std::strtol(p, &q, 10);
    if (q == p) return lfail(HBS_ERR_FORMAT);
    p = q;
    long NZ = std::strtol(p, &q, 10);
    if (q == p) return lfail(HBS_ERR_FORMAT);
    p = q;
    if (M < 0 || N < 0 || NZ < 0 || M > 0x7fffffffL || N > 0x7fffffffL || NZ > 0x7fffffffL)
        return lfail(HBS_ERR_FORMAT);
    std::vector<int> ri((size_t)NZ), ci((size_t)NZ);
    std::vector<double> vv((size_t)NZ);
    if (!parse_triplets_parallel(p, end, M, N, NZ, ri, ci, vv)) {
        // one token stream, whatever the line structure (the reference's fscanf loop)
        for (long k = 0; k < NZ; ++k) {
            long r = std::strtol(p, &q, 10);
            if (q == p) return lfail(HBS_ERR_FORMAT);
            p = q;
            long c = std::strtol(p, &q, 10);
            if (q == p) return lfail(HBS_ERR_FORMAT);
            p = q;
            double v = std::strtod(p, &q);
            if (q == p) return lfail(HBS_ERR_FORMAT);
            p = q;
            if (r < 1 || r > M || c < 1 || c > N) return lfail(HBS_ERR_FORMAT);
            ri[(size_t)k] = (int)(r - 1);
            ci[(size_t)k] = (int)(c - 1);
            vv[(size_t)k] = v;
        }
    }
    int *rp = hmalloc<int>((size_t)M + 1);
    int *cc = hmalloc<int>((size_t)NZ);
    double *va = hmalloc<double>((size_t)NZ);
    if (!rp || !cc || !va) {
        std::free(rp);
        std::free(cc);
        std::free(va);
        return lfail(HBS_ERR_ALLOC);
    }
    std::fill(rp, rp + M + 1, 0);
    for (long k = 0; k < NZ; ++k) rp[ri[(size_t)k] + 1]++;
    for (long i = 0; i < M; ++i) rp[i + 1] += rp[i];
    std::vector<int> fill(rp, rp + M);
    for (long k = 0; k < NZ; ++k) {   // stable: file order within a row
        const int d = fill[(size_t)ri[(size_t)k]]++;
        cc[d] = ci[(size_t)k];
        va[d] = vv[(size_t)k];
    }
    *m = (int)M;
    *n = (int)N;
    *nnz = (int)NZ;
    *rowptr = rp;
    *col = cc;
    *val = va;
    return HBS_OK;
}

// Whitespace separated doubles, no header (reference vector_read,
// hhp_cpu.c:411-428).  A short file is an error here instead of a printed
// warning with uninitialised tail.
extern "C" int hbs_vector_read(const char *fname, int size, double *out)
{
    if (!fname || size < 0 || (size > 0 && !out)) return lfail(HBS_ERR_ARG);
    std::string buf;
    if (!read_file(fname, buf)) return lfail(HBS_ERR_IO);
    const char *p = buf.c_str();
    char *q = nullptr;
    for (int i = 0; i < size; ++i) {
        double v = std::strtod(p, &q);
        if (q == p) return lfail(HBS_ERR_FORMAT);
        out[i] = v;
        p = q;
    }
    return HBS_OK;
}

// "%lf\n" per value (reference vector_write, hhp_cpu.c:430-442).
extern "C" int hbs_vector_write(const char *fname, int size, const double *v)
{
    if (!fname || size < 0 || (size > 0 && !v)) return lfail(HBS_ERR_ARG);
    FILE *f = std::fopen(fname, "w");
    if (!f) return lfail(HBS_ERR_IO);
    for (int i = 0; i < size; ++i)
        if (std::fprintf(f, "%lf\n", v[i]) < 0) {
            std::fclose(f);
            return lfail(HBS_ERR_IO);
        }
    std::fclose(f);
    return HBS_OK;
}

// Rank `rank`'s slice of a global vector: the entries of the rows the
// partition gives it, in ascending row order (reference
// MPI_vector_read_parted, hhp_cpu.c:252-334: rank 0 orders the vector by
// part, then by row, and Scatterv hands each rank its run).
extern "C" int hbs_vector_part(int size, const double *big, const int *partvec, int nparts, int rank, double *out,
                               int *nloc)
{
    if (size < 0 || nparts < 1 || rank < 0 || rank >= nparts || (size > 0 && (!big || !partvec)))
        return lfail(HBS_ERR_ARG);
    int k = 0;
    for (int i = 0; i < size; ++i) {
        const int p = partvec[i];
        if (p < 0 || p >= nparts) return lfail(HBS_ERR_PARTITION);
        if (p == rank) {
            if (out) out[k] = big[i];
            ++k;
        }
    }
    if (nloc) *nloc = k;
    return HBS_OK;
}

extern "C" int hbs_vector_read_parted(const char *fname, int size, const int *partvec, int nparts, int rank,
                                      double **out, int *nloc)
{
    if (!fname || !out || !nloc || size < 0) return lfail(HBS_ERR_ARG);
    *out = nullptr;
    *nloc = 0;
    std::vector<double> big((size_t)size);
    int rc = hbs_vector_read(fname, size, big.data());
    if (rc) return rc;
    int k = 0;
    if ((rc = hbs_vector_part(size, big.data(), partvec, nparts, rank, nullptr, &k))) return rc;
    double *o = static_cast<double *>(std::malloc(sizeof(double) * (size_t)std::max(k, 1)));
    if (!o) return lfail(HBS_ERR_ALLOC);
    hbs_vector_part(size, big.data(), partvec, nparts, rank, o, &k);
    *out = o;
    *nloc = k;
    return HBS_OK;
}

// One rank's part of the gather back to a global vector: big[gind[k]] =
// loc[k] for its nloc owned entries (reference MPI_vector_gather,
// hhp_cpu.c:199-250, whose root Gathervs every rank's loc values and gind
// and scatters them the same way).  Calling it for every rank's share
// assembles the whole vector.
extern "C" int hbs_vector_gather(int nloc, const double *loc, const int *gind, int size, double *big)
{
    if (nloc < 0 || size < 0 || (nloc > 0 && (!loc || !gind || !big))) return lfail(HBS_ERR_ARG);
    for (int k = 0; k < nloc; ++k)
        if (gind[k] < 0 || gind[k] >= size) return lfail(HBS_ERR_ARG);
    for (int k = 0; k < nloc; ++k) big[gind[k]] = loc[k];
    return HBS_OK;
}

// y = A x with each row summed from 0.0 in CSR order (the reference
// CSR_spmxv_seq order, hhp_cpu.c:445-455).
extern "C" int hbs_host_csr_spmv(int m, const int *rowptr, const int *col, const double *val,
                                 const double *x, double *y)
{
    if (m < 0 || !rowptr || (m > 0 && (!x || !y))) return lfail(HBS_ERR_ARG);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i) {
        double res = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) res += val[j] * x[col[j]];
        y[i] = res;
    }
    return HBS_OK;
}

extern "C" void hbs_host_shard_free(hbs_host_shard *s)
{
    if (!s) return;
    for (hbs_host_csr *c : {&s->loc, &s->shr}) {
        std::free(c->I);
        std::free(c->J);
        std::free(c->val);
    }
    std::free(s->gind.vals);
    for (hbs_host_comm *c : {&s->send, &s->recv}) {
        std::free(c->val);
        std::free(c->J);
        std::free(c->I);
        std::free(c->ranks);
    }
    std::memset(s, 0, sizeof *s);
}

// One rank's SHARD_CSC, with the contents the reference's MPI_CSR_split_row
// (hhp_matrix.c:1249-1338) gives that rank:
//   rows      owned rows in ascending global order (internal_CSR_split_row,
//             :820-894);
//   loc       owned columns, ascending global = local index, SHARD CSC with
//             row indices ascending inside a column (internal_CSR_loc_split
//             :896-981 through CSR_to_CSC :1476-1535);
//   shr       non-owned columns with at least one entry in the owned rows,
//             ascending global;
//   gind      [owned globals ; shr globals];
//   recv      owners ascending, per owner its shr columns ascending
//             (:1003-1071);
//   send      peers q ascending; for q the local indices of the owned
//             columns that q's rows reference, ascending -- the list the
//             reference obtains by receiving q's recv block and mapping it
//             through locmap (:1167-1232).
// Built directly, without the MPI round trip, by the host's threads: each
// thread takes a contiguous run of rows with about equal entries.  The CSC is
// a counting sort whose per-thread column offsets follow the thread order, so
// rows stay ascending inside every column whatever the thread count; the send
// lists come from per-peer marks over the owned columns, read back in column
// order.  Errors
// the reference aborts on come back as codes: a partition id out of range
// (:840-842) or a rank without rows or entries (hhp_cpu.c:19-20,
// hhp_matrix.c:1484-1486) -> HBS_ERR_PARTITION.
extern "C" int hbs_shard_build_host(int n, const int *rowptr, const int *col, const double *val,
                                    const int *partvec, int nparts, int rank, hbs_host_shard *out)
{
    if (!out) return lfail(HBS_ERR_ARG);
    std::memset(out, 0, sizeof *out);
    if (n <= 0 || !rowptr || !col || !val || !partvec || nparts <= 0 || rank < 0 || rank >= nparts)
        return lfail(HBS_ERR_ARG);
    const int T = std::max(1, std::min(omp_get_max_threads(), kMaxThreads));
    int bad = 0;
#pragma omp parallel for num_threads(T) reduction(| : bad) schedule(static)
    for (int i = 0; i < n; ++i) bad |= partvec[i] < 0 || partvec[i] >= nparts;
    if (bad) return lfail(HBS_ERR_PARTITION);
    std::vector<int> locmap((size_t)n, -1), owned;
    for (int i = 0; i < n; ++i)
        if (partvec[i] == rank) {
            locmap[(size_t)i] = (int)owned.size();
            owned.push_back(i);
        }
    const int m_loc = (int)owned.size();
    if (m_loc == 0) return lfail(HBS_ERR_PARTITION);
    // owned rows in T runs of about equal entries (opre: entries before each)
    std::vector<long> opre((size_t)m_loc + 1, 0);
    for (int lr = 0; lr < m_loc; ++lr)
        opre[(size_t)lr + 1] = opre[(size_t)lr] + rowptr[owned[(size_t)lr] + 1] - rowptr[owned[(size_t)lr]];
    const std::vector<int> orun = runs(opre.data(), m_loc, T);
    // shr columns: non-owned columns referenced by owned rows
    long nnz_loc = 0, nnz_shr = 0;
    std::vector<char> used((size_t)n, 0);
#pragma omp parallel for num_threads(T) reduction(+ : nnz_loc, nnz_shr) schedule(static, 1)
    for (int t = 0; t < T; ++t)
        for (int lr = orun[(size_t)t]; lr < orun[(size_t)t + 1]; ++lr) {
            const int r = owned[(size_t)lr];
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const int c = col[j];
                if (locmap[(size_t)c] >= 0) {
                    ++nnz_loc;
                } else {
                    ++nnz_shr;
                    __atomic_store_n(&used[(size_t)c], (char)1, __ATOMIC_RELAXED);
                }
            }
        }
    if (nnz_loc + nnz_shr == 0) return lfail(HBS_ERR_PARTITION);
    if (nnz_loc > INT_MAX || nnz_shr > INT_MAX) return lfail(HBS_ERR_ARG);
    // cmap: column -> CSC column (loc columns first, then shr)
    std::vector<int> cmap(locmap), shrcols;
    for (int c = 0; c < n; ++c)
        if (used[(size_t)c]) {
            cmap[(size_t)c] = m_loc + (int)shrcols.size();
            shrcols.push_back(c);
        }
    const int n_shr = (int)shrcols.size();
    const int ncols = m_loc + n_shr;

    hbs_host_shard &s = *out;
    s.loc.m = m_loc;
    s.loc.n = m_loc;
    s.loc.nnz = (int)nnz_loc;
    s.shr.m = m_loc;
    s.shr.n = n_shr;
    s.shr.nnz = (int)nnz_shr;
    s.loc.I = hmalloc<int>((size_t)m_loc + 1);
    s.loc.J = hmalloc<int>((size_t)nnz_loc);
    s.loc.val = hmalloc<double>((size_t)nnz_loc);
    s.shr.I = hmalloc<int>((size_t)n_shr + 1);
    s.shr.J = hmalloc<int>((size_t)nnz_shr);
    s.shr.val = hmalloc<double>((size_t)nnz_shr);
    s.gind.nvals = (unsigned)ncols;
    s.gind.vals = hmalloc<int>((size_t)ncols);
    if (!s.loc.I || !s.loc.J || !s.loc.val || !s.shr.I || !s.shr.J || !s.shr.val || !s.gind.vals) {
        hbs_host_shard_free(out);
        return lfail(HBS_ERR_ALLOC);
    }
    std::copy(owned.begin(), owned.end(), s.gind.vals);
    std::copy(shrcols.begin(), shrcols.end(), s.gind.vals + m_loc);
    // CSC by a counting sort over the columns: thread t's entries of column k
    // go after those of threads 0..t-1, whose rows come first
    std::vector<int> cnt((size_t)T * ncols, 0);
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        int *ct = cnt.data() + (size_t)t * ncols;
        for (int lr = orun[(size_t)t]; lr < orun[(size_t)t + 1]; ++lr) {
            const int r = owned[(size_t)lr];
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) ct[cmap[(size_t)col[j]]]++;
        }
    }
    std::vector<int> total((size_t)ncols);
#pragma omp parallel for num_threads(T) schedule(static)
    for (int k = 0; k < ncols; ++k) {
        int run = 0;
        for (int t = 0; t < T; ++t) {
            int &c = cnt[(size_t)t * ncols + k];
            const int here = c;
            c = run;
            run += here;
        }
        total[(size_t)k] = run;
    }
    s.loc.I[0] = 0;
    for (int k = 0; k < m_loc; ++k) s.loc.I[k + 1] = s.loc.I[k] + total[(size_t)k];
    s.shr.I[0] = 0;
    for (int k = 0; k < n_shr; ++k) s.shr.I[k + 1] = s.shr.I[k] + total[(size_t)m_loc + k];
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        int *ct = cnt.data() + (size_t)t * ncols;
        for (int lr = orun[(size_t)t]; lr < orun[(size_t)t + 1]; ++lr) {
            const int r = owned[(size_t)lr];
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const int k = cmap[(size_t)col[j]];
                if (k < m_loc) {
                    const int d = s.loc.I[k] + ct[k]++;
                    s.loc.J[d] = lr;
                    s.loc.val[d] = val[j];
                } else {
                    const int d = s.shr.I[k - m_loc] + ct[k]++;
                    s.shr.J[d] = lr;
                    s.shr.val[d] = val[j];
                }
            }
        }
    }
    std::vector<int>().swap(cnt);
    // recv: owners ascending; per owner its shr columns ascending
    std::vector<int> ocount((size_t)nparts + 1, 0);
    for (int k = 0; k < n_shr; ++k) ocount[(size_t)partvec[shrcols[(size_t)k]] + 1]++;
    int rnum = 0;
    for (int p = 0; p < nparts; ++p)
        if (ocount[(size_t)p + 1]) ++rnum;
    for (int p = 0; p < nparts; ++p) ocount[(size_t)p + 1] += ocount[(size_t)p];
    s.recv.num = rnum;
    s.recv.I = hmalloc<int>((size_t)rnum + 1);
    s.recv.ranks = hmalloc<int>((size_t)rnum);
    s.recv.J = hmalloc<int>((size_t)n_shr);
    s.recv.val = static_cast<double *>(std::calloc((size_t)(n_shr ? n_shr : 1), sizeof(double)));
    if (!s.recv.I || !s.recv.ranks || !s.recv.J || !s.recv.val) {
        hbs_host_shard_free(out);
        return lfail(HBS_ERR_ALLOC);
    }
    s.recv.I[0] = 0;
    for (int p = 0, b = 0; p < nparts; ++p)
        if (ocount[(size_t)p + 1] > ocount[(size_t)p]) {
            s.recv.ranks[b] = p;
            s.recv.I[b + 1] = ocount[(size_t)p + 1];
            ++b;
        }
    {
        std::vector<int> fo(ocount.begin(), ocount.end() - 1);
        for (int k = 0; k < n_shr; ++k) s.recv.J[fo[(size_t)partvec[shrcols[(size_t)k]]]++] = k;
    }
    // send: the owned columns each peer's rows reference, per peer ascending.
    // Peers are taken in groups whose marks (one byte per peer and owned
    // column) fit in about 4n bytes; within a group each thread marks for a
    // run of the group's rows, and each peer's marks are read back in column
    // order.
    std::vector<int> pstart((size_t)nparts + 1, 0), prow((size_t)n);
    for (int r = 0; r < n; ++r) pstart[(size_t)partvec[r] + 1]++;
    for (int q = 0; q < nparts; ++q) pstart[(size_t)q + 1] += pstart[(size_t)q];
    {
        std::vector<int> fp(pstart.begin(), pstart.end() - 1);
        for (int r = 0; r < n; ++r) prow[(size_t)fp[(size_t)partvec[r]]++] = r;
    }
    const int G = (int)std::max<size_t>(1, std::min<size_t>((size_t)nparts, (size_t)4 * n / (size_t)m_loc));
    std::vector<char> mark((size_t)G * m_loc);
    std::vector<int> sendI((size_t)nparts + 1, 0), sendJ;
    for (int q0 = 0; q0 < nparts; q0 += G) {
        const int q1 = std::min(nparts, q0 + G), a = pstart[(size_t)q0], nr = pstart[(size_t)q1] - a;
        std::fill(mark.begin(), mark.begin() + (size_t)(q1 - q0) * m_loc, 0);
        std::vector<long> gpre((size_t)nr + 1, 0);
        for (int i = 0; i < nr; ++i) {
            const int r = prow[(size_t)a + i];
            gpre[(size_t)i + 1] = gpre[(size_t)i] + rowptr[r + 1] - rowptr[r];
        }
        const std::vector<int> grun = runs(gpre.data(), nr, T);
#pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int t = 0; t < T; ++t)
            for (int i = grun[(size_t)t]; i < grun[(size_t)t + 1]; ++i) {
                const int r = prow[(size_t)a + i], q = partvec[r];
                if (q == rank) continue;
                char *mq = mark.data() + (size_t)(q - q0) * m_loc;
                for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                    const int lc = locmap[(size_t)col[j]];
                    if (lc >= 0) __atomic_store_n(&mq[lc], (char)1, __ATOMIC_RELAXED);
                }
            }
#pragma omp parallel for num_threads(T) schedule(dynamic, 1)
        for (int q = q0; q < q1; ++q) {
            const char *mq = mark.data() + (size_t)(q - q0) * m_loc;
            int c = 0;
            for (int lc = 0; lc < m_loc; ++lc) c += mq[lc];
            sendI[(size_t)q + 1] = c;
        }
        for (int q = q0; q < q1; ++q) sendI[(size_t)q + 1] += sendI[(size_t)q];
        sendJ.resize((size_t)sendI[(size_t)q1]);
#pragma omp parallel for num_threads(T) schedule(dynamic, 1)
        for (int q = q0; q < q1; ++q) {
            const char *mq = mark.data() + (size_t)(q - q0) * m_loc;
            int d = sendI[(size_t)q];
            for (int lc = 0; lc < m_loc; ++lc)
                if (mq[lc]) sendJ[(size_t)d++] = lc;
        }
    }
    int snum = 0;
    for (int q = 0; q < nparts; ++q)
        if (sendI[(size_t)q + 1] > sendI[(size_t)q]) ++snum;
    s.send.num = snum;
    s.send.I = hmalloc<int>((size_t)snum + 1);
    s.send.ranks = hmalloc<int>((size_t)snum);
    s.send.J = hmalloc<int>(sendJ.size());
    s.send.val = static_cast<double *>(std::calloc(sendJ.size() ? sendJ.size() : 1, sizeof(double)));
    if (!s.send.I || !s.send.ranks || !s.send.J || !s.send.val) {
        hbs_host_shard_free(out);
        return lfail(HBS_ERR_ALLOC);
    }
    std::copy(sendJ.begin(), sendJ.end(), s.send.J);
    s.send.I[0] = 0;
    for (int q = 0, b = 0; q < nparts; ++q)
        if (sendI[(size_t)q + 1] > sendI[(size_t)q]) {
            s.send.ranks[b] = q;
            s.send.I[b + 1] = sendI[(size_t)q + 1];
            ++b;
        }
    return HBS_OK;
}
