// hbs_part.cpp -- own hypergraph partitioner for the row partition vector.
//
// Replaces the closed PaToH call of CalcPartVec (hhp_matrix.c:225-300): the
// same model and inputs, our own algorithm.
//  * Model (as CalcPartVec sets it up): cells are the n columns of a square
//    CSR (cell j also owns row j), nets are the m rows, net i's pins are the
//    columns of row i (xpins = rowptr, pins = col), every net costs 1, cell
//    j weighs the entries of row j (the SpMV work of the row).  The metric is
//    connectivity - 1 (PATOH_CONPART): sum over rows of (distinct parts among
//    the row's columns - 1), the x values a row's owner must receive.
//  * Part weights follow target[k] (relative, as the weight file of
//    CalcPartVec gives them), within 1 + final_imbal.
//  * Algorithm: recursive bisection (the parts split into two ranges of
//    near-equal target weight), each bisection multilevel:
//      - coarsening by heavy-connectivity matching (a cell joins the free
//        neighbour sharing the most nets, each net counting 1 / (|net| - 1);
//        nets above kMatchNetCap pins are left out of the scores);
//      - initial bisection of the coarsest hypergraph: the cells split in
//        index order, and greedy growing from kInitTries random seeds; the
//        best after FM;
//      - uncoarsening with Fiduccia-Mattheyses refinement (cut-net gains,
//        balance-respecting moves, rollback to the best prefix).
//    Cut nets are split between the two sides for the recursion, so the
//    bisection cuts add up to connectivity - 1.
//  * Deterministic for a given seed (std::mt19937 seeded from it).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <queue>
#include <random>
#include <vector>

#include <omp.h>

#include "hbs_err.h"
#include "hbs_internal.h"

namespace hbs {
namespace {

constexpr int kCoarsestCells = 160;   // stop coarsening at about this many cells
constexpr int kMatchNetCap = 512;     // nets with more pins do not score matches
constexpr int kInitTries = 8;         // greedy-growing starts at the coarsest level
constexpr int kFmPasses = 4;          // FM passes per level (stops early without gain)

// Hypergraph: nets -> pins (xp, pin), cells -> nets (xn, net), cell weights,
// net costs.
struct Hyper {
    int nc = 0, nn = 0;
    std::vector<int> xp, pin, xn, net;
    std::vector<long long> cw;
    std::vector<int> cost;

    // Cells -> nets, each cell's nets ascending: a counting sort of the pins.
    // Large hypergraphs split the nets into runs of about equal pins, one per
    // thread; a cell's slots go to the runs in order, so the lists are the
    // same for any thread count.
    void build_incidence()
    {
        const int T = pin.size() < ((size_t)1 << 20) ? 1 : std::max(1, std::min(omp_get_max_threads(), 32));
        std::vector<int> run((size_t)T + 1, nn);
        run[0] = 0;
        for (int t = 1, e = 0; t < T; ++t) {
            const long long want = (long long)pin.size() * t / T;
            while (e < nn && xp[(size_t)e] < want) ++e;
            run[(size_t)t] = e;
        }
        std::vector<int> cnt((size_t)T * nc, 0);
#pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int t = 0; t < T; ++t) {
            int *ct = cnt.data() + (size_t)t * nc;
            for (int j = xp[(size_t)run[(size_t)t]]; j < xp[(size_t)run[(size_t)t + 1]]; ++j) ct[pin[(size_t)j]]++;
        }
        xn.assign((size_t)nc + 1, 0);
        for (int c = 0; c < nc; ++c) {
            int s = xn[(size_t)c];
            for (int t = 0; t < T; ++t) {
                int &k = cnt[(size_t)t * nc + c];
                const int here = k;
                k = s;
                s += here;
            }
            xn[(size_t)c + 1] = s;
        }
        net.resize(pin.size());
#pragma omp parallel for num_threads(T) schedule(static, 1)
        for (int t = 0; t < T; ++t) {
            int *ct = cnt.data() + (size_t)t * nc;
            for (int e = run[(size_t)t]; e < run[(size_t)t + 1]; ++e)
                for (int j = xp[(size_t)e]; j < xp[(size_t)e + 1]; ++j) net[(size_t)ct[pin[(size_t)j]]++] = e;
        }
    }
    long long total_weight() const { return std::accumulate(cw.begin(), cw.end(), 0LL); }
};

// Gain buckets of one bisection: per side, doubly linked lists of the free
// cells by gain (gains lie in [-D, D], D the largest cost-weighted cell
// degree), and the highest non-empty bucket.
struct Buckets {
    int D = 0;
    std::vector<int> head[2], nxt, prv;
    int top[2] = {-1, -1};
    void init(int nc, int d)
    {
        D = d;
        for (auto &h : head) h.assign((size_t)2 * d + 1, -1);
        nxt.assign((size_t)nc, -1);
        prv.assign((size_t)nc, -1);
        top[0] = top[1] = -1;
    }
    void insert(int c, int s, long long g)
    {
        const int i = (int)g + D;
        nxt[(size_t)c] = head[s][(size_t)i];
        prv[(size_t)c] = -1;
        if (head[s][(size_t)i] >= 0) prv[(size_t)head[s][(size_t)i]] = c;
        head[s][(size_t)i] = c;
        top[s] = std::max(top[s], i);
    }
    void remove(int c, int s, long long g)
    {
        const int i = (int)g + D;
        if (prv[(size_t)c] >= 0) nxt[(size_t)prv[(size_t)c]] = nxt[(size_t)c];
        else head[s][(size_t)i] = nxt[(size_t)c];
        if (nxt[(size_t)c] >= 0) prv[(size_t)nxt[(size_t)c]] = prv[(size_t)c];
    }
};

// Fiduccia-Mattheyses refinement of a bisection: side weights capped at
// maxw[0], maxw[1].  A pass moves free cells one at a time, best gain first
// among the moves the caps allow (a side over its cap only gives cells
// away), locks each, and keeps the best prefix.  Returns the cut after
// refinement.
long long fm_refine(const Hyper &h, std::vector<char> &side, const long long maxw[2], int passes)
{
    std::vector<int> cnt((size_t)h.nn * 2, 0);
    long long w[2] = {0, 0};
    for (int c = 0; c < h.nc; ++c) w[(int)side[(size_t)c]] += h.cw[(size_t)c];
    for (int e = 0; e < h.nn; ++e)
        for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) cnt[(size_t)e * 2 + side[(size_t)h.pin[(size_t)j]]]++;
    long long cut = 0;
    for (int e = 0; e < h.nn; ++e)
        if (cnt[(size_t)e * 2] && cnt[(size_t)e * 2 + 1]) cut += h.cost[(size_t)e];
    int D = 1;
    for (int c = 0; c < h.nc; ++c) {
        long long d = 0;
        for (int k = h.xn[(size_t)c]; k < h.xn[(size_t)c + 1]; ++k) d += h.cost[(size_t)h.net[(size_t)k]];
        D = (int)std::max<long long>(D, d);
    }
    std::vector<long long> gain((size_t)h.nc);
    std::vector<char> locked((size_t)h.nc);
    Buckets bk;
    auto over = [&](int s) { return w[s] > maxw[s]; };
    for (int pass = 0; pass < passes; ++pass) {
        const bool was_over = over(0) || over(1);
        std::fill(locked.begin(), locked.end(), 0);
        bk.init(h.nc, D);
        for (int c = 0; c < h.nc; ++c) {
            const int s = side[(size_t)c];
            long long g = 0;
            for (int k = h.xn[(size_t)c]; k < h.xn[(size_t)c + 1]; ++k) {
                const int e = h.net[(size_t)k];
                if (cnt[(size_t)e * 2 + s] == 1) g += h.cost[(size_t)e];
                if (cnt[(size_t)e * 2 + (1 - s)] == 0) g -= h.cost[(size_t)e];
            }
            gain[(size_t)c] = g;
            bk.insert(c, s, g);
        }
        std::vector<int> moves;
        long long run = 0, best = 0;
        size_t best_len = 0;
        const size_t stall = std::max<size_t>(64, (size_t)h.nc / 50);
        auto bump = [&](int u, long long d) {
            const int su = side[(size_t)u];
            bk.remove(u, su, gain[(size_t)u]);
            gain[(size_t)u] += d;
            bk.insert(u, su, gain[(size_t)u]);
        };
        for (;;) {
            // best feasible move: from an over-cap side if there is one, else
            // the higher gain of the two sides' best moves the caps allow
            int pick = -1;
            long long bg = 0;
            for (int s = 0; s < 2; ++s) {
                if ((over(0) || over(1)) && !over(s)) continue;
                while (bk.top[s] >= 0 && bk.head[s][(size_t)bk.top[s]] < 0) --bk.top[s];
                int seen = 0;
                for (int i = bk.top[s]; i >= 0 && seen < 64 && (pick < 0 || i - bk.D > bg); --i)
                    for (int c = bk.head[s][(size_t)i]; c >= 0 && seen < 64; c = bk.nxt[(size_t)c], ++seen)
                        if (over(s) || w[1 - s] + h.cw[(size_t)c] <= maxw[1 - s]) {
                            if (pick < 0 || i - bk.D > bg) {
                                pick = c;
                                bg = i - bk.D;
                            }
                            i = -1;   // best of this side found
                            break;
                        }
            }
            if (pick < 0) break;
            // move pick: s -> t, FM gain updates
            const int s = side[(size_t)pick], t = 1 - s;
            bk.remove(pick, s, gain[(size_t)pick]);
            locked[(size_t)pick] = 1;
            for (int k = h.xn[(size_t)pick]; k < h.xn[(size_t)pick + 1]; ++k) {
                const int e = h.net[(size_t)k];
                const int c = h.cost[(size_t)e];
                int &cf = cnt[(size_t)e * 2 + s], &ct = cnt[(size_t)e * 2 + t];
                if (ct == 0) {
                    for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
                        const int u = h.pin[(size_t)j];
                        if (!locked[(size_t)u]) bump(u, c);
                    }
                } else if (ct == 1) {
                    for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
                        const int u = h.pin[(size_t)j];
                        if (!locked[(size_t)u] && side[(size_t)u] == t) bump(u, -c);
                    }
                }
                --cf;
                ++ct;
                if (cf == 0) {
                    for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
                        const int u = h.pin[(size_t)j];
                        if (!locked[(size_t)u]) bump(u, -c);
                    }
                } else if (cf == 1) {
                    for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
                        const int u = h.pin[(size_t)j];
                        if (!locked[(size_t)u] && side[(size_t)u] == s) bump(u, c);
                    }
                }
            }
            side[(size_t)pick] = (char)t;
            w[s] -= h.cw[(size_t)pick];
            w[t] += h.cw[(size_t)pick];
            run += gain[(size_t)pick];
            moves.push_back(pick);
            // while a side is over its cap every move counts (balance first)
            if (run > best || (was_over && !over(0) && !over(1) && best_len < moves.size() &&
                               (best_len == 0 || run >= best))) {
                best = run;
                best_len = moves.size();
            } else if (moves.size() - best_len > stall) {
                break;
            }
        }
        // roll back past the best prefix
        for (size_t i = moves.size(); i > best_len; --i) {
            const int c = moves[i - 1];
            const int s = side[(size_t)c], t = 1 - s;
            for (int k = h.xn[(size_t)c]; k < h.xn[(size_t)c + 1]; ++k) {
                const int e = h.net[(size_t)k];
                cnt[(size_t)e * 2 + s]--;
                cnt[(size_t)e * 2 + t]++;
            }
            side[(size_t)c] = (char)t;
            w[s] -= h.cw[(size_t)c];
            w[t] += h.cw[(size_t)c];
        }
        cut = 0;
        for (int e = 0; e < h.nn; ++e)
            if (cnt[(size_t)e * 2] && cnt[(size_t)e * 2 + 1]) cut += h.cost[(size_t)e];
        if (best_len == 0 && !was_over) break;
    }
    return cut;
}

// One coarsening level: heavy-connectivity matching; cmap[c] = coarse cell.
bool coarsen(const Hyper &h, std::mt19937 &rng, long long maxcw, Hyper &co, std::vector<int> &cmap)
{
    std::vector<int> order((size_t)h.nc);
    std::iota(order.begin(), order.end(), 0);
    std::shuffle(order.begin(), order.end(), rng);
    std::vector<int> match((size_t)h.nc, -1);
    std::vector<double> score((size_t)h.nc, 0.0);
    std::vector<int> touched;
    for (int u : order) {
        if (match[(size_t)u] >= 0) continue;
        touched.clear();
        for (int k = h.xn[(size_t)u]; k < h.xn[(size_t)u + 1]; ++k) {
            const int e = h.net[(size_t)k];
            const int sz = h.xp[(size_t)e + 1] - h.xp[(size_t)e];
            if (sz < 2 || sz > kMatchNetCap) continue;
            const double s = (double)h.cost[(size_t)e] / (sz - 1);
            for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
                const int v = h.pin[(size_t)j];
                if (v == u || match[(size_t)v] >= 0 || h.cw[(size_t)u] + h.cw[(size_t)v] > maxcw) continue;
                if (score[(size_t)v] == 0.0) touched.push_back(v);
                score[(size_t)v] += s;
            }
        }
        int best = -1;
        double bs = 0.0;
        for (int v : touched) {
            if (score[(size_t)v] > bs || (score[(size_t)v] == bs && v < best)) {
                bs = score[(size_t)v];
                best = v;
            }
            score[(size_t)v] = 0.0;
        }
        match[(size_t)u] = best >= 0 ? best : u;
        if (best >= 0) match[(size_t)best] = u;
    }
    cmap.assign((size_t)h.nc, -1);
    int nc = 0;
    for (int c = 0; c < h.nc; ++c) {
        if (cmap[(size_t)c] >= 0) continue;
        cmap[(size_t)c] = nc;
        if (match[(size_t)c] != c) cmap[(size_t)match[(size_t)c]] = nc;
        ++nc;
    }
    if (nc > h.nc * 95 / 100) return false;   // too little progress
    co = Hyper();
    co.nc = nc;
    co.cw.assign((size_t)nc, 0);
    for (int c = 0; c < h.nc; ++c) co.cw[(size_t)cmap[(size_t)c]] += h.cw[(size_t)c];
    co.xp.push_back(0);
    std::vector<int> mark((size_t)nc, -1);
    for (int e = 0; e < h.nn; ++e) {
        const size_t start = co.pin.size();
        for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
            const int cc = cmap[(size_t)h.pin[(size_t)j]];
            if (mark[(size_t)cc] != e) {
                mark[(size_t)cc] = e;
                co.pin.push_back(cc);
            }
        }
        if (co.pin.size() - start < 2) {
            co.pin.resize(start);   // a single-pin net can never be cut
            continue;
        }
        co.cost.push_back(h.cost[(size_t)e]);
        co.xp.push_back((int)co.pin.size());
    }
    co.nn = (int)co.cost.size();
    co.build_incidence();
    return true;
}

// Greedy growing: side 1 grows from a random seed, the cell with the best
// FM gain joining next, until it holds target1 of the weight.
void grow(const Hyper &h, std::mt19937 &rng, long long target1, std::vector<char> &side)
{
    side.assign((size_t)h.nc, 0);
    std::vector<int> cnt1((size_t)h.nn, 0);
    std::vector<long long> g((size_t)h.nc, 0);
    std::vector<char> in((size_t)h.nc, 0);
    using Entry = std::pair<long long, int>;
    std::priority_queue<Entry> pq;
    long long w1 = 0;
    std::uniform_int_distribution<int> pickc(0, std::max(0, h.nc - 1));
    auto add = [&](int c) {
        in[(size_t)c] = 1;
        side[(size_t)c] = 1;
        w1 += h.cw[(size_t)c];
        for (int k = h.xn[(size_t)c]; k < h.xn[(size_t)c + 1]; ++k) {
            const int e = h.net[(size_t)k];
            if (cnt1[(size_t)e]++ == 0)
                for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
                    const int u = h.pin[(size_t)j];
                    if (!in[(size_t)u]) {
                        g[(size_t)u] += h.cost[(size_t)e];
                        pq.push({g[(size_t)u], u});
                    }
                }
        }
    };
    while (w1 < target1) {
        int next = -1;
        while (!pq.empty()) {
            const Entry t = pq.top();
            pq.pop();
            if (!in[(size_t)t.second] && t.first == g[(size_t)t.second]) {
                next = t.second;
                break;
            }
        }
        if (next < 0) {
            // disconnected: a random free cell
            for (int tries = 0; tries < 64 && next < 0; ++tries) {
                const int c = pickc(rng);
                if (!in[(size_t)c]) next = c;
            }
            for (int c = 0; c < h.nc && next < 0; ++c)
                if (!in[(size_t)c]) next = c;
            if (next < 0) break;
        }
        if (w1 + h.cw[(size_t)next] > target1 && w1 > 0 &&
            (w1 + h.cw[(size_t)next]) - target1 > target1 - w1)
            break;
        add(next);
    }
}

// Multilevel bisection of h with side-1 weight share frac1; caps at
// (1 + eps) times each side's target.
std::vector<char> bisect(const Hyper &h, double frac1, double eps, std::mt19937 &rng)
{
    const long long W = h.total_weight();
    const long long t1 = (long long)std::llround((double)W * frac1);
    const long long maxw[2] = {(long long)((double)(W - t1) * (1.0 + eps)) + 1,
                               (long long)((double)t1 * (1.0 + eps)) + 1};
    std::vector<Hyper> levels;
    std::vector<std::vector<int>> maps;
    levels.push_back(h);
    const long long maxcw = std::max<long long>(1, W / kCoarsestCells);
    while (levels.back().nc > kCoarsestCells) {
        Hyper co;
        std::vector<int> cmap;
        if (!coarsen(levels.back(), rng, maxcw, co, cmap)) break;
        // matching that merges cells without shrinking nets (random column
        // patterns) only repeats the same work level after level
        if (co.pin.size() * 10 > levels.back().pin.size() * 9) {
            levels.push_back(std::move(co));
            maps.push_back(std::move(cmap));
            break;
        }
        levels.push_back(std::move(co));
        maps.push_back(std::move(cmap));
    }
    // initial bisection: best of kInitTries greedy growings + FM (one when
    // coarsening stopped early on a large hypergraph)
    const Hyper &hc = levels.back();
    std::vector<char> best;
    long long best_cut = -1;
    bool best_ok = false;
    const int tries = hc.nc <= 4 * kCoarsestCells && hc.pin.size() <= 200000 ? kInitTries : 1;
    for (int t = -1; t < tries; ++t) {
        std::vector<char> side;
        if (t < 0) {
            // the cells in index order, split at the target weight: coarse
            // cells keep the order of their first fine cell, so this keeps
            // whatever locality the matrix's own ordering has
            side.assign((size_t)hc.nc, 1);
            long long w0 = 0;
            for (int c = 0; c < hc.nc && w0 + hc.cw[(size_t)c] <= W - t1; ++c) {
                side[(size_t)c] = 0;
                w0 += hc.cw[(size_t)c];
            }
        } else {
            grow(hc, rng, t1, side);
        }
        const long long cut = fm_refine(hc, side, maxw, kFmPasses);
        long long w1 = 0;
        for (int c = 0; c < hc.nc; ++c) w1 += side[(size_t)c] ? hc.cw[(size_t)c] : 0;
        // a balanced split beats any unbalanced one; then the lower cut
        const bool ok = w1 <= maxw[1] && W - w1 <= maxw[0];
        if (best_cut < 0 || (ok && !best_ok) || (ok == best_ok && cut < best_cut)) {
            best_cut = cut;
            best = side;
            best_ok = ok;
        }
    }
    // uncoarsen with refinement
    for (size_t l = levels.size() - 1; l > 0; --l) {
        const std::vector<int> &cmap = maps[l - 1];
        std::vector<char> fine((size_t)levels[l - 1].nc);
        for (int c = 0; c < levels[l - 1].nc; ++c) fine[(size_t)c] = best[(size_t)cmap[(size_t)c]];
        best.swap(fine);
        fm_refine(levels[l - 1], best, maxw, kFmPasses);
    }
    return best;
}

// The sub-hypergraph of the cells with side == s: cells renumbered in
// order, nets restricted to their pins there (nets with one pin dropped).
Hyper restrict_to(const Hyper &h, const std::vector<char> &side, char s, std::vector<int> &cells)
{
    std::vector<int> id((size_t)h.nc, -1);
    cells.clear();
    for (int c = 0; c < h.nc; ++c)
        if (side[(size_t)c] == s) {
            id[(size_t)c] = (int)cells.size();
            cells.push_back(c);
        }
    Hyper r;
    r.nc = (int)cells.size();
    r.cw.resize(cells.size());
    for (size_t i = 0; i < cells.size(); ++i) r.cw[i] = h.cw[(size_t)cells[i]];
    r.xp.push_back(0);
    for (int e = 0; e < h.nn; ++e) {
        const size_t start = r.pin.size();
        for (int j = h.xp[(size_t)e]; j < h.xp[(size_t)e + 1]; ++j) {
            const int c = id[(size_t)h.pin[(size_t)j]];
            if (c >= 0) r.pin.push_back(c);
        }
        if (r.pin.size() - start < 2) {
            r.pin.resize(start);
            continue;
        }
        r.cost.push_back(h.cost[(size_t)e]);
        r.xp.push_back((int)r.pin.size());
    }
    r.nn = (int)r.cost.size();
    r.build_incidence();
    return r;
}

// Parts [p0, p1) of the cells `cells` (global ids) of h; every bisection
// keeps each side within 1 + eps_level of its target.
void recurse(const Hyper &h, const std::vector<int> &cells, int p0, int p1, const std::vector<double> &target,
             double eps_level, std::mt19937 &rng, int *partvec)
{
    if (p1 - p0 == 1 || h.nc == 0) {
        for (int c : cells) partvec[c] = p0;
        return;
    }
    // split the part range where the target weights balance best
    double tot = 0.0;
    for (int p = p0; p < p1; ++p) tot += target[(size_t)p];
    int pm = p0 + 1;
    double left = target[(size_t)p0], best = std::fabs(2.0 * left - tot);
    for (int p = p0 + 1; p < p1 - 1; ++p) {
        left += target[(size_t)p];
        const double d = std::fabs(2.0 * left - tot);
        if (d < best) {
            best = d;
            pm = p + 1;
        }
    }
    double tl = 0.0;
    for (int p = p0; p < pm; ++p) tl += target[(size_t)p];
    const std::vector<char> side = bisect(h, tot > 0 ? 1.0 - tl / tot : 0.5, eps_level, rng);
    for (char s = 0; s < 2; ++s) {
        std::vector<int> sub_cells;
        const Hyper sub = restrict_to(h, side, s, sub_cells);
        for (int &c : sub_cells) c = cells[(size_t)c];
        if (s == 0) recurse(sub, sub_cells, p0, pm, target, eps_level, rng, partvec);
        else recurse(sub, sub_cells, pm, p1, target, eps_level, rng, partvec);
    }
}

}  // namespace

long long partition_cut(int n, const int *rowptr, const int *col, const int *partvec)
{
    long long cut = 0;
    std::vector<int> mark;
    for (int r = 0; r < n; ++r) {
        int lambda = 0;
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            const int p = partvec[col[j]];
            if (p >= (int)mark.size()) mark.resize((size_t)p + 1, -1);
            if (mark[(size_t)p] != r) {
                mark[(size_t)p] = r;
                ++lambda;
            }
        }
        if (lambda > 1) cut += lambda - 1;
    }
    return cut;
}

int partition_rows(int n, const int *rowptr, const int *col, int nparts, const double *target, double final_imbal,
                   int seed, int *partvec)
{
    // rows are nets over their columns; single-pin nets (a row whose only
    // column is its own diagonal) cannot be cut and are dropped
    Hyper h;
    h.nc = n;
    h.cw.resize((size_t)n);
    for (int r = 0; r < n; ++r) h.cw[(size_t)r] = std::max(1, rowptr[r + 1] - rowptr[r]);
    h.xp.push_back(0);
    for (int r = 0; r < n; ++r) {
        if (rowptr[r + 1] - rowptr[r] < 2) continue;
        h.pin.insert(h.pin.end(), col + rowptr[r], col + rowptr[r + 1]);
        h.cost.push_back(1);
        h.xp.push_back((int)h.pin.size());
    }
    h.nn = (int)h.cost.size();
    h.build_incidence();
    std::vector<double> t(target, target + nparts);
    std::vector<int> cells((size_t)n);
    std::iota(cells.begin(), cells.end(), 0);
    std::mt19937 rng((unsigned)seed);
    // the imbalance allowance spreads evenly over the levels of the recursion
    // (a part's weight compounds one factor per level)
    int depth = 0;
    while ((1 << depth) < nparts) ++depth;
    const double eps_level = std::pow(1.0 + final_imbal, 1.0 / std::max(1, depth)) - 1.0;
    recurse(h, cells, 0, nparts, t, eps_level, rng, partvec);
    return HBS_OK;
}

}  // namespace hbs

extern "C" int hbs_partition(int n, const int *rowptr, const int *col, int nparts, const double *target,
                             double final_imbal, int seed, int *partvec, long long *cut)
{
    using namespace hbs;
    if (n < 0 || nparts < 1 || !partvec || (n > 0 && (!rowptr || !col)) || !(final_imbal >= 0.0))
        return set_error(HBS_ERR_ARG, "bad partition arguments (n %d, nparts %d)", n, nparts);
    if (n > 0 && rowptr[0] != 0) return set_error(HBS_ERR_ARG, "rowptr[0] must be 0");
    for (int r = 0; r < n; ++r)
        if (rowptr[r + 1] < rowptr[r]) return set_error(HBS_ERR_ARG, "rowptr decreases at row %d", r);
    for (int j = 0; j < (n > 0 ? rowptr[n] : 0); ++j)
        if (col[j] < 0 || col[j] >= n) return set_error(HBS_ERR_ARG, "column %d out of range at entry %d", col[j], j);
    std::vector<double> t((size_t)nparts, 1.0);
    if (target)
        for (int p = 0; p < nparts; ++p) {
            if (!(target[p] > 0.0)) return set_error(HBS_ERR_ARG, "target weight %d must be positive", p);
            t[(size_t)p] = target[p];
        }
    if (n == 0) {
        if (cut) *cut = 0;
        return HBS_OK;
    }
    const int rc = partition_rows(n, rowptr, col, nparts, t.data(), final_imbal, seed, partvec);
    if (rc == HBS_OK && cut) *cut = partition_cut(n, rowptr, col, partvec);
    return rc;
}
