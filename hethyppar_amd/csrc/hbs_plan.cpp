// hbs_plan.cpp -- row-block, workgroup and vector-block planning.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <vector>

#include "hbs_internal.h"

namespace hbs {

namespace {
// Rows per block: fewer than 4 per thread of the SpMV workgroup (the kernel
// stages the block's nr + 1 row offsets with 4 loads per thread).
constexpr int max_rows_per_block(int threads) { return 4 * threads - 1; }

int pow2_floor(int v)
{
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}
}  // namespace

// Greedy row blocking: a block takes consecutive rows while their entries fit
// in block_nnz (the LDS product buffer).  A row longer than
// min(block_nnz, kLongRow) becomes one kind-1 block whose value is the
// ordered sum of its chunk sums, each chunk (kChunkLen entries) summed by its
// own workgroup of k_chunks, so no thread sums more than kLongRow products
// in turn.
// Blocks whose average row length is at most serial_max_avg use the
// order-exact serial row sum; others pow2_floor(threads / rows) lanes per
// row, threads = spmv_threads(block_nnz).
void plan_rows(int row0, int m, const int *rowptr, int block_nnz, int serial_max_avg, int sys,
               HostPlan &plan)
{
    const int end = row0 + m;
    const int threads = spmv_threads(block_nnz);
    const int max_rows = max_rows_per_block(threads);
    const int long_row = std::min(block_nnz, kLongRow);
    int r = row0;
    while (r < end) {
        const int len = rowptr[r + 1] - rowptr[r];
        plan.max_row = std::max(plan.max_row, len);
        if (len > long_row) {
            const int nch = (len + kChunkLen - 1) / kChunkLen;
            const int first = (int)plan.chunks.size();
            for (int c = 0; c < nch; ++c) {
                const int j0 = rowptr[r] + c * kChunkLen;
                plan.chunks.push_back(Chunk{j0, std::min(j0 + kChunkLen, rowptr[r + 1]), r, sys});
            }
            plan.blocks.push_back(RowBlock{r, r + 1, rowptr[r], rowptr[r + 1]});
            plan.aux.push_back(BlockAux{1, first, nch, 1});
            ++r;
            continue;
        }
        const int r0 = r;
        const int j0 = rowptr[r];
        while (r < end && rowptr[r + 1] - j0 <= block_nnz && rowptr[r + 1] - rowptr[r] <= long_row &&
               r - r0 < max_rows) {
            plan.max_row = std::max(plan.max_row, rowptr[r + 1] - rowptr[r]);
            ++r;
        }
        const int nr = r - r0;
        const int nz = rowptr[r] - j0;
        int lanes = 1;
        if (nr > 0 && nz > serial_max_avg * nr) {
            lanes = std::min(64, pow2_floor(threads / nr));
            if (lanes < 2) lanes = 1;
        }
        if (lanes == 1) plan.serial_rows += nr;
        plan.blocks.push_back(RowBlock{r0, r, j0, rowptr[r]});
        plan.aux.push_back(BlockAux{0, 0, 0, lanes});
    }
}

bool detect_b3(int m, int n, const int *rowptr, const int *col)
{
    if (m <= 0 || m % 3 != 0 || n % 3 != 0) return false;
    for (int r = 0; r < m; r += 3) {
        const int j0 = rowptr[r], len = rowptr[r + 1] - j0;
        if (len <= 0 || len % 3 != 0 || len > kLongRow) return false;
        for (int a = 1; a < 3; ++a)
            if (rowptr[r + a + 1] - rowptr[r + a] != len) return false;
        for (int k = 0; k < len; k += 3) {
            const int c0 = col[j0 + k];
            if (c0 % 3 != 0 || col[j0 + k + 1] != c0 + 1 || col[j0 + k + 2] != c0 + 2) return false;
        }
        for (int a = 1; a < 3; ++a)
            for (int t = 0; t < len; ++t)
                if (col[rowptr[r + a] + t] != col[j0 + t]) return false;
    }
    return true;
}

int plan_rows_b3(int row0, int m, const int *rowptr, int block_nnz, int sys, HostPlan &plan)
{
    const int end = row0 + m;
    const int max_rows = max_rows_per_block(kWideThreads);
    block_nnz = std::min(block_nnz, kB3BlockNnz);   // at most one block per thread
    int r = row0, nvar = 0;
    while (r < end) {
        const int r0 = r, j0 = rowptr[r], len = rowptr[r + 1] - rowptr[r];
        bool same = true;
        do {
            same = same && rowptr[r + 1] - rowptr[r] == len;
            plan.max_row = std::max(plan.max_row, rowptr[r + 1] - rowptr[r]);
            r += 3;
        } while (r < end && rowptr[r + 3] - j0 <= block_nnz && r + 3 - r0 <= max_rows);
        plan.serial_rows += r - r0;
        plan.blocks.push_back(RowBlock{r0, r, j0, rowptr[r]});
        plan.aux.push_back(BlockAux{3, 0, same ? len / 3 : 0, 1});
        nvar += !same;
    }
    (void)sys;
    return nvar;
}

void b3_layout(const HostPlan &plan, const int *rowptr, const int *col, std::vector<int> &bsrc,
               std::vector<int> &bcol)
{
    size_t nnz = 0;
    for (const RowBlock &rb : plan.blocks) nnz = std::max(nnz, (size_t)rb.j1);
    bsrc.resize(nnz);
    bcol.resize(nnz / 9);
    for (const RowBlock &rb : plan.blocks) {
        const int nbk = (rb.j1 - rb.j0) / 9;
        for (int r = rb.r0; r < rb.r1; r += 3) {
            const int kb = (rowptr[r + 1] - rowptr[r]) / 3;
            const int b0 = (rowptr[r] - rb.j0) / 9;   // the block row's first block in the row block
            for (int k = 0; k < kb; ++k) {
                bcol[(size_t)rb.j0 / 9 + b0 + k] = col[rowptr[r] + 3 * k];
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c)
                        bsrc[(size_t)rb.j0 + (size_t)(3 * a + c) * nbk + b0 + k] = rowptr[r + a] + 3 * k + c;
            }
        }
    }
}

int plan_row_order(const int *rowptr, HostPlan &plan, int m, std::vector<unsigned short> &rperm)
{
    rperm.assign((size_t)std::max(m, 1), 0);
    const int th = kWideThreads, waves = kWideThreads / 64;
    int nperm = 0;
    std::vector<int> ord;
    std::vector<long long> cost((size_t)waves);
    for (size_t b = 0; b < plan.blocks.size(); ++b) {
        const RowBlock &rb = plan.blocks[b];
        BlockAux &ax = plan.aux[b];
        const int nr = rb.r1 - rb.r0;
        for (int k = 0; k < nr; ++k) rperm[(size_t)rb.r0 + k] = (unsigned short)k;
        if (ax.kind == 1 || ax.lanes > 1 || nr < 2) continue;
        auto len = [&](int k) { return rowptr[rb.r0 + k + 1] - rowptr[rb.r0 + k]; };
        // slot s of the row sums belongs to wave (s % th) / 64; its rows are
        // summed in batches of 8 products, so a slot of 64 rows costs the
        // batches of its longest row
        auto estimate = [&](const std::vector<int> &o) {
            std::fill(cost.begin(), cost.end(), 0);
            for (int s0 = 0; s0 < nr; s0 += 64) {
                int mx = 0;
                for (int s = s0; s < std::min(nr, s0 + 64); ++s) mx = std::max(mx, len(o[(size_t)s]));
                cost[(size_t)((s0 % th) / 64)] += (mx + 7) / 8;
            }
            return *std::max_element(cost.begin(), cost.end());
        };
        ord.resize((size_t)nr);
        for (int k = 0; k < nr; ++k) ord[(size_t)k] = k;
        const long long c_row = estimate(ord);
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return len(x) > len(y); });
        const long long c_sorted = estimate(ord);
        if (c_sorted * 5 >= c_row * 4) continue;
        ax.first = 1;
        ++nperm;
        for (int k = 0; k < nr; ++k) rperm[(size_t)rb.r0 + k] = (unsigned short)ord[(size_t)k];
    }
    return nperm;
}

int plan_sorted(const int *col, int nnz, int max_line_pct, HostPlan &plan, std::vector<int> &src,
                std::vector<unsigned short> &pos)
{
    src.resize((size_t)nnz);
    pos.resize((size_t)nnz);
    for (int j = 0; j < nnz; ++j) src[(size_t)j] = j;
    int nsorted = 0;
    std::vector<int> idx;
    for (size_t b = 0; b < plan.blocks.size(); ++b) {
        const RowBlock &rb = plan.blocks[b];
        BlockAux &ax = plan.aux[b];
        const int nz = rb.j1 - rb.j0;
        for (int k = 0; k < nz; ++k) pos[(size_t)rb.j0 + k] = (unsigned short)(ax.kind == 1 ? 0 : k);
        if (ax.kind != 0 || nz < 2 || max_line_pct <= 0) continue;
        idx.resize((size_t)nz);
        for (int k = 0; k < nz; ++k) idx[(size_t)k] = k;
        // distinct 128-byte lines of x per wave instruction (64 consecutive
        // entries), in CSR order and in column order
        auto wave_lines = [&](bool sorted_order) {
            long long total = 0;
            std::vector<int> ls;
            for (int g = 0; g < nz; g += 64) {
                ls.clear();
                for (int k = g; k < std::min(nz, g + 64); ++k)
                    ls.push_back(col[rb.j0 + (sorted_order ? idx[(size_t)k] : k)] >> 4);
                std::sort(ls.begin(), ls.end());
                total += std::unique(ls.begin(), ls.end()) - ls.begin();
            }
            return total;
        };
        const long long csr_lines = wave_lines(false);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int c) { return col[rb.j0 + a] < col[rb.j0 + c]; });
        const long long sorted_lines = wave_lines(true);
        if (sorted_lines * 100 > (long long)max_line_pct * csr_lines) continue;
        ax.kind = 2;
        ++nsorted;
        for (int k = 0; k < nz; ++k) {
            src[(size_t)rb.j0 + k] = rb.j0 + idx[(size_t)k];
            pos[(size_t)rb.j0 + k] = (unsigned short)idx[(size_t)k];
        }
    }
    return nsorted;
}

double column_scatter(const int *col, int nnz, int window)
{
    if (nnz <= 0 || window <= 0) return 0.0;
    const long long stride = 16LL * window;   // one window in 16
    long long lines = 0, ents = 0;
    std::vector<int> ls;
    for (long long b = 0; b < nnz; b += stride) {
        const long long e = std::min<long long>(nnz, b + window);
        ls.assign(col + b, col + e);
        for (int &v : ls) v >>= 4;
        std::sort(ls.begin(), ls.end());
        lines += std::unique(ls.begin(), ls.end()) - ls.begin();
        ents += e - b;
    }
    return (double)lines / (double)ents;
}

// Row-block-major order: each block's entries sorted by (segment, CSR
// position).  Its runs of one segment are the chunks; segment-major order
// lists, per segment, its chunks in block order.  gch / delta name the chunk
// of every segment-major entry within its segment (delta < 64: a chunk holds
// at least one entry, and a granule 64).
void plan_xseg(const int *col, int n, const HostPlan &plan, XsegHost &xs)
{
    const int nseg = std::max(1, (n + kSegCols - 1) / kSegCols);
    const int nnz = plan.blocks.empty() ? 0 : plan.blocks.back().j1;
    xs.nseg = nseg;
    xs.ppos.assign((size_t)nnz, 0);
    std::vector<int> rbsrc((size_t)nnz);
    struct Run {
        int s, p0, cnt;
    };
    std::vector<Run> runs;
    std::vector<long long> seg_cnt((size_t)nseg, 0);
    std::vector<int> seg_chunks((size_t)nseg, 0);
    std::vector<unsigned long long> key;
    for (size_t b = 0; b < plan.blocks.size(); ++b) {
        const RowBlock &rb = plan.blocks[b];
        const bool long_row = plan.aux[b].kind == 1;
        const int nz = rb.j1 - rb.j0;
        key.resize((size_t)nz);
        for (int k = 0; k < nz; ++k)
            key[(size_t)k] = ((unsigned long long)(unsigned)(col[rb.j0 + k] / kSegCols) << 32) | (unsigned)k;
        std::sort(key.begin(), key.end());
        for (int k = 0; k < nz;) {
            const int s = (int)(key[(size_t)k] >> 32);
            int e = k;
            for (; e < nz && (int)(key[(size_t)e] >> 32) == s; ++e) {
                const int pos = (int)(key[(size_t)e] & 0xffffffffu);
                rbsrc[(size_t)rb.j0 + e] = rb.j0 + pos;
                xs.ppos[(size_t)rb.j0 + e] = (unsigned short)(long_row ? 0 : pos);
            }
            runs.push_back(Run{s, rb.j0 + k, e - k});
            seg_cnt[(size_t)s] += e - k;
            seg_chunks[(size_t)s]++;
            k = e;
        }
    }
    xs.sbase.assign((size_t)nseg + 1, 0);
    xs.cbase.assign((size_t)nseg + 1, 0);
    std::vector<int> &cbase = xs.cbase;
    for (int s = 0; s < nseg; ++s) {
        xs.sbase[(size_t)s + 1] = xs.sbase[(size_t)s] + (int)seg_cnt[(size_t)s];
        cbase[(size_t)s + 1] = cbase[(size_t)s] + seg_chunks[(size_t)s];
        xs.max_chunks = std::max(xs.max_chunks, seg_chunks[(size_t)s]);
    }
    const int nchunks = cbase[(size_t)nseg];
    xs.off.assign((size_t)std::max(nchunks, 1), 0);
    std::vector<int> cstart((size_t)nchunks + 1, nnz);
    std::vector<int> next_e(xs.sbase.begin(), xs.sbase.end() - 1), next_c(cbase.begin(), cbase.end() - 1);
    xs.src.assign((size_t)nnz, 0);
    xs.xcol.assign((size_t)nnz, 0);
    for (const Run &r : runs) {
        const int c = next_c[(size_t)r.s]++, g0 = next_e[(size_t)r.s];
        next_e[(size_t)r.s] += r.cnt;
        cstart[(size_t)c] = g0;
        xs.off[(size_t)c] = r.p0 - g0;
        for (int k = 0; k < r.cnt; ++k) {
            const int j = rbsrc[(size_t)r.p0 + k];
            xs.src[(size_t)g0 + k] = j;
            xs.xcol[(size_t)g0 + k] = (unsigned short)(col[j] - r.s * kSegCols);
        }
    }
    xs.gbase.assign((size_t)nseg + 1, 0);
    for (int s = 0; s < nseg; ++s) {
        const int ng = (int)((seg_cnt[(size_t)s] + 63) / 64);
        xs.gbase[(size_t)s + 1] = xs.gbase[(size_t)s] + ng;
        xs.max_granules = std::max(xs.max_granules, ng);
    }
    xs.gch.assign((size_t)std::max(xs.gbase[(size_t)nseg], 1), 0);
    xs.delta.assign((size_t)nnz, 0);
    for (int s = 0; s < nseg; ++s) {
        int c = cbase[(size_t)s];
        const int cnt = xs.sbase[(size_t)s + 1] - xs.sbase[(size_t)s];
        for (int f = 0; f < cnt; ++f) {
            const int g = xs.sbase[(size_t)s] + f;
            while (c + 1 < cbase[(size_t)s + 1] && cstart[(size_t)c + 1] <= g) ++c;
            int &first = xs.gch[(size_t)xs.gbase[(size_t)s] + f / 64];   // within the segment
            if (f % 64 == 0) first = c - cbase[(size_t)s];
            xs.delta[(size_t)g] = (unsigned char)(c - cbase[(size_t)s] - first);
        }
    }
}

// Contiguous, equal-count split of row blocks [rb0, rb1) over ngroups
// workgroups (row blocks carry about block_nnz entries each).
void plan_groups(int rb0, int rb1, int ngroups, int sys, HostPlan &plan)
{
    const int nb = rb1 - rb0;
    if (nb <= 0) return;
    ngroups = std::max(1, std::min(ngroups, nb));
    for (int p = 0; p < ngroups; ++p) {
        const int a = rb0 + (int)((long long)nb * p / ngroups);
        const int b = rb0 + (int)((long long)nb * (p + 1) / ngroups);
        plan.groups.push_back(Group{a, b, sys, 0});
    }
}

// A block's packs start at base = e0 rounded down to 4 elements (16-byte
// aligned for FP32 and FP64) and cover [base, base + elems_per_block); the
// kernels mask what lies outside [e0, e1).  elems_per_block is a multiple of
// 4, so only a system's first block can start below its e0.
void plan_vec(int e0, int n, int elems_per_block, int sys, std::vector<VecBlock> &blocks)
{
    const int end = e0 + n;
    for (int e = e0; e < end;) {
        const int base = e & ~3;
        const int e1 = std::min(base + elems_per_block, end);
        blocks.push_back(VecBlock{sys, e, e1, base});
        e = e1;
    }
}

int plan_vec_into(int e0, int n, int elems_per_block, int sys, VecBlock *out, int cap)
{
    const int end = e0 + n;
    int k = 0;
    for (int e = e0; e < end; ++k) {
        const int base = e & ~3;
        const int e1 = std::min(base + elems_per_block, end);
        if (k >= cap) return -1;
        if (out) out[k] = VecBlock{sys, e, e1, base};
        e = e1;
    }
    return k;
}

// Row runs of the multi-column SpMM on a 3-D structured grid (plan_grid_runs).
// The rows of a 7-point operator on an nx x ny x nz grid (x fastest) gather
// X rows r +- 1, r +- nx and r +- nx*ny.  In row order the +-z uses of an X
// row lie nx*ny rows apart -- 8 MiB of 64 FP64 columns at 128^2 -- so each
// one misses an XCD's 4 MiB L2 and X is fetched about once per use.  Swept
// in y-slabs of H lines, plane after plane, the three planes that use an X
// row are consecutive, and a slab's window of three planes of X stays in
// the L2; the +-y neighbours of the slab's edge lines are the only re-reads
// ((H + 2) / H of X).  The runs are the grid's x-lines (nx consecutive
// rows); each XCD takes an equal contiguous share of the runs, whole slabs
// when ny / 8 is a multiple of H.
namespace {
constexpr long long kL2PerXcd = 4LL << 20;
constexpr int kGridRowBytes = 64 * 8;     // one X row of a 64-column FP64 block
}  // namespace

int plan_grid_runs(int m, const int *rowptr, const int *col, std::vector<int> &runs, int &len)
{
    runs.clear();
    len = 0;
    if (m < 8 * 1024 || !rowptr || !col) return 0;
    // offsets |c - r| of a sample of rows
    std::map<long long, long long> cnt;
    long long sampled = 0, entries = 0;
    const int stride = std::max(1, m / 20000);
    for (int r = 0; r < m; r += stride) {
        ++sampled;
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            ++cnt[std::llabs((long long)col[j] - r)];
            ++entries;
        }
    }
    // the distinct offsets other than 0 must be 1, nx and nx * ny, each in at
    // least a fifth of the rows
    std::vector<long long> offs;
    long long covered = cnt.count(0) ? cnt[0] : 0;
    for (const auto &kv : cnt)
        if (kv.first > 0 && kv.second * 5 >= sampled) {
            offs.push_back(kv.first);
            covered += kv.second;
        }
    if (covered != entries || offs.size() != 3 || offs[0] != 1) return 0;
    const long long nx = offs[1], nxy = offs[2];
    if (nxy % nx != 0 || m % nxy != 0) return 0;
    const long long ny = nxy / nx, nz = m / nxy;
    // worth it only when a plane of X outgrows the L2
    if (nxy * kGridRowBytes <= kL2PerXcd || ny < 8 || nz < 3) return 0;
    if (nx % 2 != 0) return 0;   // runs hold whole steps of two rows
    // slab height: three planes of the slab within half the L2, a divisor of
    // each XCD's share of the y-lines when there is one
    long long h = g_grid_slab > 0 ? g_grid_slab : std::max(1LL, (kL2PerXcd / 2) / (3 * nx * kGridRowBytes));
    h = std::min(h, std::max(1LL, ny / 8));
    while (h > 1 && (ny / 8) % h != 0) --h;
    for (long long y0 = 0; y0 < ny; y0 += h) {
        const long long y1 = std::min(ny, y0 + h);
        for (long long z = 0; z < nz; ++z)
            for (long long y = y0; y < y1; ++y) runs.push_back((int)(z * nxy + y * nx));
    }
    len = (int)nx;
    return (int)runs.size();
}

int g_vec_target = 512;
int g_grid_slab = 0;
int g_spmm_groups = 0;
int g_spmm_v16 = 1;
int g_spmm_counters = kCtrPerXcd;
int g_xseg_fuse = 1;

// Elements per vector workgroup: 512 * u with u in {1, 2, 4, 8}, the
// smallest that keeps the workgroups (and the partials a consumer sums) at
// g_vec_target or fewer.  An FP64 thread then holds u 16-byte packs, an FP32
// thread max(1, u/2).
int vec_elems_per_block(int n)
{
    int u = 1;
    while (u < 8 && (long long)n > 512LL * u * g_vec_target) u *= 2;
    return 512 * u;
}

int vec_packs_per_thread(int elems_per_block, int elem_bytes)
{
    const int per_pack = 16 / elem_bytes;
    return std::max(1, (elems_per_block + kThreads * per_pack - 1) / (kThreads * per_pack));
}

}  // namespace hbs
