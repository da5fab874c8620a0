// hbs_gen.cpp -- synthetic workloads of the benchmark configurations.
//
// Randomness: SplitMix64 (Steele, Lea, Flood 2014).  Every row (or block row)
// draws from its own stream, seeded with mix64(seed ^ mix64(row + 1)), so the
// generators run in parallel and give the same matrix for any thread count.
// u01 = (next() >> 11) * 2^-53.  Columns ascend within every row.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

#include "hbs.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) : s(mix64(seed ^ mix64(stream + 1))) {}
    uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ULL;
        return mix64(s);
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, n)
    long long below(long long n) { return (long long)(u01() * (double)n); }
    // standard normal, Box-Muller
    double normal()
    {
        const double u1 = 1.0 - u01();
        const double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

template <typename T>
T *hmalloc(size_t n)
{
    return static_cast<T *>(std::malloc(sizeof(T) * (n ? n : 1)));
}

// Allocates CSR arrays for row lengths len[0..m) (prefix sum into rowptr).
int alloc_csr(const std::vector<int> &len, int *nnz, int **rowptr, int **col, double **val)
{
    const size_t m = len.size();
    int *rp = hmalloc<int>(m + 1);
    if (!rp) return HBS_ERR_ALLOC;
    long long acc = 0;
    rp[0] = 0;
    for (size_t i = 0; i < m; ++i) {
        acc += len[i];
        if (acc > 0x7fffffffLL) {
            std::free(rp);
            return HBS_ERR_ARG;
        }
        rp[i + 1] = (int)acc;
    }
    int *c = hmalloc<int>((size_t)acc);
    double *v = hmalloc<double>((size_t)acc);
    if (!c || !v) {
        std::free(rp);
        std::free(c);
        std::free(v);
        return HBS_ERR_ALLOC;
    }
    *nnz = (int)acc;
    *rowptr = rp;
    *col = c;
    *val = v;
    return HBS_OK;
}

void sort_row(int *c, double *v, int k)
{
    std::vector<std::pair<int, double>> t((size_t)k);
    for (int i = 0; i < k; ++i) t[(size_t)i] = {c[i], v[i]};
    std::sort(t.begin(), t.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    for (int i = 0; i < k; ++i) {
        c[i] = t[(size_t)i].first;
        v[i] = t[(size_t)i].second;
    }
}

}  // namespace

extern "C" int hbs_gen_uniform(int n, double lo, double hi, uint64_t seed, double *out)
{
    if (n < 0 || (n > 0 && !out)) return HBS_ERR_ARG;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        Rng r(seed, (uint64_t)i);
        out[i] = lo + (hi - lo) * r.u01();
    }
    return HBS_OK;
}

// 2D 5-point convection-diffusion, lexicographic row-major grid (i = iy*nx+ix):
// south -1, west -1-conv (upwind), centre 4+conv, east -1, north -1.
extern "C" int hbs_gen_convdiff2d(int nx, int ny, double conv, int *m, int *nnz, int **rowptr, int **col,
                                  double **val)
{
    if (nx <= 0 || ny <= 0 || !m || !nnz || !rowptr || !col || !val) return HBS_ERR_ARG;
    const long long nn = (long long)nx * ny;
    if (nn > 0x7fffffffLL / 5) return HBS_ERR_ARG;
    std::vector<int> len((size_t)nn);
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            len[(size_t)iy * nx + ix] = 1 + (iy > 0) + (ix > 0) + (ix < nx - 1) + (iy < ny - 1);
    int rc = alloc_csr(len, nnz, rowptr, col, val);
    if (rc) return rc;
    int *rp = *rowptr, *c = *col;
    double *v = *val;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const int i = iy * nx + ix;
            int d = rp[i];
            if (iy > 0) { c[d] = i - nx; v[d++] = -1.0; }
            if (ix > 0) { c[d] = i - 1; v[d++] = -1.0 - conv; }
            c[d] = i; v[d++] = 4.0 + conv;
            if (ix < nx - 1) { c[d] = i + 1; v[d++] = -1.0; }
            if (iy < ny - 1) { c[d] = i + nx; v[d++] = -1.0; }
        }
    *m = (int)nn;
    return HBS_OK;
}

// 3D 7-point convection-diffusion, i = (iz*ny + iy)*nx + ix: -x neighbour
// -1-conv, other neighbours -1, centre 6+conv.
extern "C" int hbs_gen_convdiff3d(int nx, int ny, int nz, double conv, int *m, int *nnz, int **rowptr,
                                  int **col, double **val)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !m || !nnz || !rowptr || !col || !val) return HBS_ERR_ARG;
    const long long nn = (long long)nx * ny * nz;
    if (nn > 0x7fffffffLL / 7) return HBS_ERR_ARG;
    std::vector<int> len((size_t)nn);
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
                len[((size_t)iz * ny + iy) * nx + ix] =
                    1 + (iz > 0) + (iy > 0) + (ix > 0) + (ix < nx - 1) + (iy < ny - 1) + (iz < nz - 1);
    int rc = alloc_csr(len, nnz, rowptr, col, val);
    if (rc) return rc;
    int *rp = *rowptr, *c = *col;
    double *v = *val;
    const int sxy = nx * ny;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int i = (iz * ny + iy) * nx + ix;
                int d = rp[i];
                if (iz > 0) { c[d] = i - sxy; v[d++] = -1.0; }
                if (iy > 0) { c[d] = i - nx; v[d++] = -1.0; }
                if (ix > 0) { c[d] = i - 1; v[d++] = -1.0 - conv; }
                c[d] = i; v[d++] = 6.0 + conv;
                if (ix < nx - 1) { c[d] = i + 1; v[d++] = -1.0; }
                if (iy < ny - 1) { c[d] = i + nx; v[d++] = -1.0; }
                if (iz < nz - 1) { c[d] = i + sxy; v[d++] = -1.0; }
            }
    *m = (int)nn;
    return HBS_OK;
}

// Markov-like nonsymmetric matrix.  Row i (own stream): k = kmin +
// floor(u*(kmax-kmin+1)) distinct columns drawn uniformly from the band
// [i-band, i+band] (clipped, i excluded), then one column drawn uniformly
// outside the band (if any exists), then one U[0,1] value per off-diagonal
// in draw order; diagonal = dom * sum of the row's off-diagonal values.
extern "C" int hbs_gen_markov(int n, int kmin, int kmax, int band, double dom, uint64_t seed, int *nnz,
                              int **rowptr, int **col, double **val)
{
    if (n <= 0 || kmin < 0 || kmax < kmin || band < 0 || !nnz || !rowptr || !col || !val) return HBS_ERR_ARG;
    std::vector<int> len((size_t)n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        Rng r(seed, (uint64_t)i);
        int k = kmin + (int)r.below(kmax - kmin + 1);
        const int lo = std::max(0, i - band), hi = std::min(n - 1, i + band);
        k = std::min(k, hi - lo);
        const int far = (lo > 0 || hi < n - 1) ? 1 : 0;
        len[(size_t)i] = k + far + 1;
    }
    int rc = alloc_csr(len, nnz, rowptr, col, val);
    if (rc) return rc;
    int *rp = *rowptr, *c = *col;
    double *v = *val;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        Rng r(seed, (uint64_t)i);
        int k = kmin + (int)r.below(kmax - kmin + 1);
        const int lo = std::max(0, i - band), hi = std::min(n - 1, i + band);
        k = std::min(k, hi - lo);
        const int far = (lo > 0 || hi < n - 1) ? 1 : 0;
        int *cc = c + rp[i];
        double *vv = v + rp[i];
        int cnt = 0;
        while (cnt < k) {
            const int cand = lo + (int)r.below(hi - lo + 1);
            if (cand == i) continue;
            bool dup = false;
            for (int t = 0; t < cnt; ++t) dup |= cc[t] == cand;
            if (!dup) cc[cnt++] = cand;
        }
        if (far) {
            const int outside = n - (hi - lo + 1);
            int f = (int)r.below(outside);
            f = f < lo ? f : f + (hi - lo + 1);
            cc[cnt++] = f;
        }
        double sum = 0.0;
        for (int t = 0; t < cnt; ++t) {
            vv[t] = r.u01();
            sum += vv[t];
        }
        cc[cnt] = i;
        vv[cnt] = dom * sum;
        sort_row(cc, vv, cnt + 1);
    }
    return HBS_OK;
}

// Circuit-like matrix with power-law rows.  nheavy rows (chosen from stream
// n) get 1 + k off-diagonals with total length log-uniform in
// [heavy_lo, heavy_hi]; the other rows get k ~ Zipf(zipf_s) on 1..K, K the
// truncation whose mean brings the whole matrix closest to mean_nnz per row.
// Columns uniform, distinct, != i; values N(0,1) * 10^U(-3,3); diagonal =
// dom * sum |a_ij|.  equilibrate != 0 divides every row by its diagonal
// (the same solution for b scaled alike): without it the row scales span
// ~1e6 and unpreconditioned BiCGStab stagnates near 1e-5.
extern "C" int hbs_gen_circuit(int n, double mean_nnz, double zipf_s, int nheavy, int heavy_lo, int heavy_hi,
                               double dom, int equilibrate, uint64_t seed, int *nnz, int **rowptr, int **col,
                               double **val)
{
    if (n <= 1 || nheavy < 0 || nheavy > n || heavy_lo < 2 || heavy_hi < heavy_lo || heavy_hi > n ||
        !nnz || !rowptr || !col || !val)
        return HBS_ERR_ARG;
    // heavy rows and their lengths
    std::vector<int> heavy_len((size_t)n, 0);
    long long H = 0;
    {
        Rng r(seed, (uint64_t)n);
        int placed = 0;
        while (placed < nheavy) {
            const int row = (int)r.below(n);
            if (heavy_len[(size_t)row]) continue;
            const double lg = std::log((double)heavy_lo) +
                              r.u01() * (std::log((double)heavy_hi) - std::log((double)heavy_lo));
            int L = (int)std::lround(std::exp(lg));
            L = std::max(heavy_lo, std::min(heavy_hi, L));
            heavy_len[(size_t)row] = L;
            H += L;
            ++placed;
        }
    }
    // truncation of the Zipf law for the light rows
    const double mu = (mean_nnz * n - (double)H) / (double)(n - nheavy) - 1.0;
    auto zmean = [&](int K) {
        double a = 0.0, b = 0.0;
        for (int k = 1; k <= K; ++k) {
            const double p = std::pow((double)k, -zipf_s);
            a += k * p;
            b += p;
        }
        return a / b;
    };
    int K = 1;
    if (mu > 1.0) {
        int lo = 1, hi = 1;
        while (hi < 100000 && zmean(hi) < mu) hi *= 2;
        hi = std::min(hi, 100000);
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (zmean(mid) <= mu) lo = mid;
            else hi = mid - 1;
        }
        K = lo;
        if (K < 100000 && std::fabs(zmean(K + 1) - mu) < std::fabs(zmean(K) - mu)) ++K;
    }
    K = std::min(K, n - 1);
    std::vector<double> cdf((size_t)K);
    {
        double acc = 0.0;
        for (int k = 1; k <= K; ++k) {
            acc += std::pow((double)k, -zipf_s);
            cdf[(size_t)k - 1] = acc;
        }
        for (double &x : cdf) x /= acc;
    }
    std::vector<int> len((size_t)n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        if (heavy_len[(size_t)i]) {
            len[(size_t)i] = heavy_len[(size_t)i];
        } else {
            Rng r(seed, (uint64_t)i);
            const double u = r.u01();
            const int k = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
            len[(size_t)i] = std::min(k, K) + 1;
        }
    }
    int rc = alloc_csr(len, nnz, rowptr, col, val);
    if (rc) return rc;
    int *rp = *rowptr, *c = *col;
    double *v = *val;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int i = 0; i < n; ++i) {
        Rng r(seed, (uint64_t)i);
        if (!heavy_len[(size_t)i]) (void)r.u01();   // the length draw
        const int k = len[(size_t)i] - 1;
        int *cc = c + rp[i];
        double *vv = v + rp[i];
        if (k <= 64) {
            int cnt = 0;
            while (cnt < k) {
                const int cand = (int)r.below(n);
                if (cand == i) continue;
                bool dup = false;
                for (int t = 0; t < cnt; ++t) dup |= cc[t] == cand;
                if (!dup) cc[cnt++] = cand;
            }
        } else {
            std::vector<int> pick;
            pick.reserve((size_t)k);
            while ((int)pick.size() < k) {
                while ((int)pick.size() < k) {
                    const int cand = (int)r.below(n);
                    if (cand != i) pick.push_back(cand);
                }
                std::sort(pick.begin(), pick.end());
                pick.erase(std::unique(pick.begin(), pick.end()), pick.end());
            }
            std::copy(pick.begin(), pick.end(), cc);
        }
        double sum = 0.0;
        for (int t = 0; t < k; ++t) {
            const double mag = std::pow(10.0, -3.0 + 6.0 * r.u01());
            vv[t] = r.normal() * mag;
            sum += std::fabs(vv[t]);
        }
        double diag = dom * sum;
        if (diag == 0.0) diag = 1.0;
        cc[k] = i;
        vv[k] = diag;
        if (equilibrate) {
            for (int t = 0; t < k; ++t) vv[t] = vv[t] / diag;
            vv[k] = 1.0;
        }
        sort_row(cc, vv, k + 1);
    }
    return HBS_OK;
}

// FEM-like matrix of 3x3 blocks.  Block row I (own stream) couples to block
// columns: I itself, the nb_band = nblk_per_row - 2 nearest other block
// indices (window [I - nb_band/2, I + nb_band/2] shifted inside the range),
// and one block column drawn uniformly among the rest.  Block values N(0,1)
// in ascending block-column order, row-major inside a block; every scalar
// diagonal entry is then set to dom * sum |off-diagonal entries of its row|.
extern "C" int hbs_gen_fem(int nblockrows, int nblk_per_row, double dom, uint64_t seed, int *nnz,
                           int **rowptr, int **col, double **val)
{
    const int NB = nblockrows, K = nblk_per_row;
    if (NB <= 0 || K < 1 || K > NB || !nnz || !rowptr || !col || !val) return HBS_ERR_ARG;
    if ((long long)NB * 3 * K * 3 > 0x7fffffffLL) return HBS_ERR_ARG;
    const int n = 3 * NB;
    std::vector<int> len((size_t)n, 3 * K);
    int rc = alloc_csr(len, nnz, rowptr, col, val);
    if (rc) return rc;
    int *rp = *rowptr, *c = *col;
    double *v = *val;
    const int nb_band = std::max(0, K - 2);
    const int want_far = K >= 2 ? 1 : 0;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < NB; ++I) {
        Rng r(seed, (uint64_t)I);
        std::vector<int> bc;
        bc.reserve((size_t)K);
        bc.push_back(I);
        // band window of nb_band other blocks, shifted to stay in range
        int lo = I - nb_band / 2, hi = lo + nb_band;   // inclusive of I: nb_band + 1 slots
        if (lo < 0) { hi -= lo; lo = 0; }
        if (hi > NB - 1) { lo -= hi - (NB - 1); hi = NB - 1; }
        lo = std::max(lo, 0);
        for (int J = lo; J <= hi && (int)bc.size() < 1 + nb_band; ++J)
            if (J != I) bc.push_back(J);
        const int need = K - (int)bc.size();
        for (int f = 0; f < need && want_far; ++f) {
            for (;;) {
                const int J = (int)r.below(NB);
                if (std::find(bc.begin(), bc.end(), J) == bc.end()) {
                    bc.push_back(J);
                    break;
                }
            }
        }
        std::sort(bc.begin(), bc.end());
        for (int a = 0; a < 3; ++a) {
            const int row = 3 * I + a;
            int d = rp[row];
            for (int J : bc)
                for (int b = 0; b < 3; ++b) c[d++] = 3 * J + b;
        }
        for (size_t t = 0; t < bc.size(); ++t)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) v[rp[3 * I + a] + 3 * (int)t + b] = r.normal();
        for (int a = 0; a < 3; ++a) {
            const int row = 3 * I + a;
            double s = 0.0;
            int dpos = -1;
            for (int j = rp[row]; j < rp[row + 1]; ++j) {
                if (c[j] == row) dpos = j;
                else s += std::fabs(v[j]);
            }
            v[dpos] = dom * s;
        }
    }
    return HBS_OK;
}
