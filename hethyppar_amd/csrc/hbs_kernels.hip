// hbs_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the BiCGStab/SpMV engine.
//
// Everything here is bound by memory traffic (no dense contraction, so no
// MFMA): coalesced 16-byte streams, wave64 reductions, XCD-aware placement.
//
//  * k_spmv     CSR-stream SpMV.  A workgroup owns a contiguous run of row
//               blocks whose entries fit in LDS: passes stream col/val
//               coalesced and store one product per entry in LDS (in column
//               order where that shares x lines, at the CSR position); then
//               each row is summed.  The passes are software-pipelined: the
//               next pass's streams load while this pass's gathers, stores
//               and row sums run.  Short-row blocks sum serially in CSR order
//               (bit-identical to the reference CSR_spmxv_seq,
//               hhp_cpu.c:445-455); blocks of longer rows use 2..64 lanes per
//               row; rows longer than 256 entries (or than a block) are cut
//               into chunks summed by k_chunks and combined here in chunk
//               order.  Epilogues fuse the dot products BiCGStab needs after
//               each SpMV.
//  * k_vec_*    the three BiCGStab vector updates (reference k_vecop_P/S/XR,
//               hhp_dp_kernels.cu:51-74), each fused with the reductions of
//               bicgstab_cpu.c:203,219,229,242 that follow it.  Each thread
//               loads its 16-byte packs first, so the loads overlap the
//               prologue that derives the scalars.
//  * reductions atomic-free and deterministic: a producer writes one partial
//               per workgroup; every workgroup of the consumer kernel sums
//               them in the same fixed order, so all of them derive
//               bit-identical scalars (alpha, omega, beta, stop decisions) in
//               their prologue.  The first workgroup of a system stores the
//               updated scalars in the other state slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <map>
#include <mutex>

#include "hbs_device.h"
#include "hbs_internal.h"

namespace hbs {
namespace {

template <int EPI>
struct EpiDots {
    static constexpr int n = (EPI == EPI_DOT2 || EPI == EPI_RESID2 || EPI == EPI_RESIDB) ? 2 : (EPI == EPI_AXPBY ? 0 : 1);
};

// Row epilogue: write y[r] and feed the fused dots (wr = w[r], loaded by the
// caller for the epilogues that read w).
template <typename T, int EPI>
__device__ __forceinline__ void row_out(const SpmvArgs<T> &a, int r, T acc, T wr, Acc<T> &d0, Acc<T> &d1)
{
    if constexpr (EPI == EPI_AXPBY) {
        T out = a.alpha * acc;
        if (a.beta != T(0)) out = out + a.beta * a.y[r];
        a.y[r] = out;
    } else if constexpr (EPI == EPI_DOT1) {
        a.y[r] = acc;
        d0.add(wr, acc);
    } else if constexpr (EPI == EPI_DOT2) {
        a.y[r] = acc;
        d0.add(acc, wr);
        d1.add(acc, acc);
    } else {
        const T out = wr - acc;
        a.y[r] = out;
        d0.add(out, out);
        if constexpr (EPI == EPI_RESID2) d1.add(out, out);
        if constexpr (EPI == EPI_RESIDB) d1.add(wr, wr);
    }
}

template <int EPI>
__device__ __forceinline__ constexpr bool reads_w()
{
    return EPI != EPI_AXPBY;
}

// Diagonal operands of row r of an x-segmented matrix whose diagonal is
// kept apart (DG 1: unit diagonal, x[r]; DG 2: diag[r] and x[r]), loaded
// with the row's epilogue operand; diag_add adds the row's diagonal term to
// the sum of its off-diagonal products.
template <typename T, int DG>
struct DiagOp {
    T x, d;
};
template <typename T, int DG>
__device__ __forceinline__ DiagOp<T, DG> diag_load(const SpmvArgs<T> &a, int r)
{
    DiagOp<T, DG> o{T(0), T(0)};
    if constexpr (DG >= 1) o.x = a.x[r];
    if constexpr (DG == 2) o.d = a.diag[r];
    return o;
}
template <typename T, int DG>
__device__ __forceinline__ T diag_add(T acc, const DiagOp<T, DG> &o)
{
    if constexpr (DG == 1) return acc + o.x;
    else if constexpr (DG == 2) return acc + o.d * o.x;
    else return acc;
}

// One pass over a row block: the entries base + t + TH * i (i < NPT) of the
// block, as loaded column indices, values and (kind 2) CSR positions -- or,
// for x-segmented matrices (PRE), the products and their CSR positions.  A
// wave instruction covers 64 consecutive entries.
template <typename T, int NPT>
struct PassLoad {
    int c[NPT];
    T v[NPT];
    unsigned short pos[NPT];
    int rp[4];   // PRE, first pass of a block: its row pointers (kRowsPerThread)
    int rq[4];   // PRE, blocks summed in rperm order: the rows of this thread's slots
};

// PRE: the block's row pointers, with its first pass's loads (staged in LDS
// when that pass is consumed).  Later passes load the block's first row
// pointer in every lane instead (one request per wave), so no branch guards
// them.
template <typename T, int NPT, int TH>
__device__ __forceinline__ void issue_rowptr(const SpmvArgs<T> &a, const RowBlock &rb, int base, PassLoad<T, NPT> &L)
{
#pragma unroll
    for (int q = 0; q < 4; ++q)
        L.rp[q] = a.rowptr[rb.r0 + (base == 0 ? min((int)threadIdx.x + q * TH, rb.r1 - rb.r0) : 0)];
    if (a.rperm) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            L.rq[q] = a.rperm[rb.r0 + (base == 0 ? min((int)threadIdx.x + q * TH, max(rb.r1 - rb.r0 - 1, 0)) : 0)];
    }
}

// Issue the stream loads of one pass.  Every lane loads: entries at or past
// the block's end read its last entry again (its first, padding included,
// for a block without entries) and are ignored later.  Loads that always
// issue let the compiler wait for exactly the ones it needs; loads behind a
// branch make it wait for every load in flight, the next pass's included.
template <typename T, int NPT, int TH, bool PRE>
__device__ __forceinline__ void issue_pass(const SpmvArgs<T> &a, const RowBlock &rb, const BlockAux &ax, int base,
                                           PassLoad<T, NPT> &L)
{
    const int hi = max(rb.j1 - rb.j0 - base - 1, 0);
    if constexpr (PRE) {
        const T *__restrict__ pr = a.prod + rb.j0 + base;
        const unsigned short *__restrict__ pp = a.ppos + rb.j0 + base;
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            const int k = min((int)threadIdx.x + i * TH, hi);
            L.v[i] = pr[k];
            L.pos[i] = pp[k];
        }
        issue_rowptr<T, NPT, TH>(a, rb, base, L);
        return;
    }
    const bool sorted = ax.kind == 2;
    const int *__restrict__ colp = (sorted ? a.scol : a.col) + rb.j0 + base;
    const T *__restrict__ valp = (sorted ? a.sval : a.val) + rb.j0 + base;
    if (sorted && !a.pos_shift) {
        const unsigned short *__restrict__ pp = a.sperm + rb.j0 + base;
#pragma unroll
        for (int i = 0; i < NPT; ++i) L.pos[i] = pp[min((int)threadIdx.x + i * TH, hi)];
    }
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int k = min((int)threadIdx.x + i * TH, hi);
        if (a.stream_nt) {
            L.c[i] = __builtin_nontemporal_load(colp + k);
            L.v[i] = __builtin_nontemporal_load(valp + k);
        } else {
            L.c[i] = colp[k];
            L.v[i] = valp[k];
        }
    }
}

// Rows per row block: at most kRowsPerThread per thread, so the serial row
// sums take one round of loads.  The block's row pointers, relative to its
// first entry, are staged in LDS as 16-bit offsets (block_nnz <= 16384).
constexpr int kRowsPerThread = 4;
static_assert(kRowsPerThread == 4, "PassLoad::rp holds 4 row pointers per thread");
template <int TH>
constexpr int rowptr_lds_bytes()
{
    return ((kRowsPerThread * TH + 1) * 2 + 15) / 16 * 16;
}

// Epilogue operand of this thread's first row of a block, loaded long before
// the row sums need it.
template <typename T, int EPI>
__device__ __forceinline__ T issue_w0(const SpmvArgs<T> &a, const RowBlock &rb, const BlockAux &ax)
{
    const int first = ax.lanes <= 1 ? (int)threadIdx.x : (int)threadIdx.x / ax.lanes;
    if (reads_w<EPI>()) return a.w[rb.r0 + min(first, rb.r1 - rb.r0 - 1)];
    return T(0);
}

// Adds lds[j0], lds[j0 + step], ... below e to acc in that order, loading B
// of them at a time so a long row costs one LDS round trip per batch rather
// than per product.  Past e the batch reads +0.0, and acc + 0.0 == acc
// exactly: acc starts at +0.0, so it is never -0.0.
template <typename T, int B>
__device__ __forceinline__ T sum_lds(const T *lds, int j0, int e, int step)
{
    T acc = T(0);
    for (int j = j0; j < e; j += B * step) {
        T v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) v[u] = j + u * step < e ? lds[j + u * step] : T(0);
#pragma unroll
        for (int u = 0; u < B; ++u) acc += v[u];
    }
    return acc;
}

// A row of a kind-3 block, its products block-major in LDS: row a of a block
// row of kb blocks reads the triples at off + 9k, k = 0 .. kb - 1, added in
// CSR order (k, then the triple's columns) from +0.0, so the sum equals the
// serial CSR sum; 4 triples are loaded per round.
template <typename T>
__device__ __forceinline__ T sum_b3(const T *p, int kb)
{
    constexpr int U = 4;
    T acc = T(0);
    for (int k0 = 0; k0 < kb; k0 += U) {
        T v[U][3];
        // past the row's last triple: read that triple again, add +0.0
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T r = p[9 * min(k0 + u, kb - 1) + c];
                v[u][c] = k0 + u < kb ? r : T(0);
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc += v[u][c];
    }
    return acc;
}

// Row t of kind-3 row block rb (rb.r0 is a multiple of 3, so row t is row
// a = t % 3 of its block row): its blocks per row kb and the LDS offset of
// its first triple.  Block rows of one length K (ax.nchunks): block row t / 3
// starts at 9K(t / 3).  Otherwise (VAR) from the row's own pointers rp =
// (rowptr[r], rowptr[r + 1]): kb = (rp.y - rp.x) / 3, and the block row's
// products start where its entries do, at rp.x - 3a kb - rb.j0.
template <bool VAR>
__device__ __forceinline__ void b3_row(const RowBlock &rb, const BlockAux &ax, int t, int2 rp, int &kb, int &off)
{
    const int ra = t % 3;
    if constexpr (VAR) {
        kb = (rp.y - rp.x) / 3;
        off = rp.x - rb.j0 + 3 * ra * (1 - kb);
    } else {
        kb = ax.nchunks;
        off = (t / 3) * 9 * kb + 3 * ra;
    }
}

// Row pointers of this thread's first row of a kind-3 block (VAR kernels),
// loaded a row block ahead like the epilogue operand.
template <bool VAR>
__device__ __forceinline__ int2 issue_rp(const int *rowptr, const RowBlock &rb)
{
    if constexpr (!VAR) return make_int2(0, 0);
    const int r = rb.r0 + min((int)threadIdx.x, rb.r1 - rb.r0 - 1);
    return make_int2(rowptr[r], rowptr[r + 1]);
}

// Epilogue operands (w[r], diagonal operands) of this thread's row slots
// of a serial block, rows in rperm order where the block takes it (PRE).
// The x-segmented row sums load them before the next pass's streams: the
// memory counter retires loads in order, so operands loaded after those
// streams would make the row sums wait for the whole next pass.
template <typename T, int DG>
struct EpiOps {
    T wr[kRowsPerThread];
    DiagOp<T, DG> dg[kRowsPerThread];
};
template <typename T, int EPI, int TH, int DG>
__device__ __forceinline__ void epi_load(const SpmvArgs<T> &a, const RowBlock &rb, const BlockAux &ax,
                                         const int (&rq)[4], EpiOps<T, DG> &o)
{
    // rperm order only in kind-0 blocks (a kind-1 block's `first` is its
    // first chunk); every row index is clamped into the block
    const int hi = max(rb.r1 - rb.r0 - 1, 0);
    const bool perm = ax.kind == 0 && ax.first;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
        const int t = perm ? rq[q] : (int)threadIdx.x + q * TH;
        const int r = rb.r0 + min(max(t, 0), hi);
        o.wr[q] = reads_w<EPI>() ? a.w[r] : T(0);
        o.dg[q] = diag_load<T, DG>(a, r);
    }
}

// Sums the rows of a block whose products (lds) and row offsets (lrp) are in
// LDS and feeds the epilogue.  Serial blocks add each row's products in CSR
// order from 0 (the reference's order): thread t sums rows t, t + TH, ...,
// with their epilogue operands loaded together (the first one, w0, was
// prefetched).  Other blocks use `lanes` lanes per row and an xor butterfly.
template <typename T, int EPI, int TH, bool PRE, int DG>
__device__ __forceinline__ void sum_rows(const SpmvArgs<T> &a, const RowBlock &rb, const BlockAux &ax, T w0,
                                         const T *lds, const unsigned short *lrp, const int (&rq)[4],
                                         const EpiOps<T, DG> &eo, Acc<T> &d0, Acc<T> &d1)
{
    constexpr int B = PRE ? 8 : 2;
    const int nr = rb.r1 - rb.r0;
    const int lanes = ax.lanes;
    if (PRE && lanes <= 1 && ax.first) {
        // rows in the block's rperm order (longest first): slot t + q * TH
        // sums row rq[q], in CSR order as ever; its epilogue operand and
        // output sit at that row
        T acc[kRowsPerThread];
#pragma unroll
        for (int q = 0; q < kRowsPerThread; ++q) {
            const int t = (int)threadIdx.x + q * TH;
            acc[q] = t < nr ? sum_lds<T, B>(lds, lrp[rq[q]], lrp[rq[q] + 1], 1) : T(0);
        }
#pragma unroll
        for (int q = 0; q < kRowsPerThread; ++q) {
            const int t = (int)threadIdx.x + q * TH;
            if (t < nr) row_out<T, EPI>(a, rb.r0 + rq[q], diag_add<T, DG>(acc[q], eo.dg[q]), eo.wr[q], d0, d1);
        }
        return;
    }
    if (lanes <= 1) {
        T wr[kRowsPerThread];
        DiagOp<T, DG> dg[kRowsPerThread];
#pragma unroll
        for (int q = 0; q < kRowsPerThread; ++q) {
            const int t = (int)threadIdx.x + q * TH;
            if constexpr (PRE) {
                wr[q] = eo.wr[q];
                dg[q] = eo.dg[q];
            } else {
                wr[q] = q == 0 ? w0 : (reads_w<EPI>() ? a.w[rb.r0 + min(t, nr - 1)] : T(0));
                dg[q] = diag_load<T, DG>(a, rb.r0 + min(t, nr - 1));
            }
        }
        // all sums first, then the stores: a store between two rows' LDS
        // reads makes the next row wait for every memory operation in flight
        // (the next pass's loads among them)
        T acc[kRowsPerThread];
#pragma unroll
        for (int q = 0; q < kRowsPerThread; ++q) {
            const int t = (int)threadIdx.x + q * TH;
            acc[q] = t < nr ? sum_lds<T, B>(lds, lrp[t], lrp[t + 1], 1) : T(0);
        }
#pragma unroll
        for (int q = 0; q < kRowsPerThread; ++q) {
            const int t = (int)threadIdx.x + q * TH;
            if (t < nr) row_out<T, EPI>(a, rb.r0 + t, diag_add<T, DG>(acc[q], dg[q]), wr[q], d0, d1);
        }
    } else {
        const int gi = threadIdx.x / lanes, l = threadIdx.x & (lanes - 1);
        const int ng = TH / lanes;
        for (int t0 = 0; t0 < nr; t0 += ng) {
            const int t = t0 + gi;
            T acc = T(0);
            T wr = w0;
            DiagOp<T, DG> dg{T(0), T(0)};
            if (t < nr) {
                if (t0 > 0 && reads_w<EPI>() && l == 0) wr = a.w[rb.r0 + t];
                if (l == 0) dg = diag_load<T, DG>(a, rb.r0 + t);
                const int e = lrp[t + 1];
                acc = sum_lds<T, B>(lds, lrp[t] + l, e, lanes);
            }
            for (int o = lanes >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (t < nr && l == 0) row_out<T, EPI>(a, rb.r0 + t, diag_add<T, DG>(acc, dg), wr, d0, d1);
        }
    }
}

// Position of a workgroup in its run of row blocks: block b (rb, ax), entry
// offset base of the pass in flight.
struct Cursor {
    int b, base;
    RowBlock rb;
    BlockAux ax;
};

// A finished system's workgroups return before any load; its first one
// hands the state on.
template <typename T>
__device__ __forceinline__ bool spmv_skip(const SpmvArgs<T> &a, const Group &g, int p)
{
    if (a.st_in && a.check_done && a.st_in[g.sys].done) {
        if (a.st_out && p == a.sys[g.sys].g0 && threadIdx.x == 0) a.st_out[g.sys] = a.st_in[g.sys];
        return true;
    }
    return false;
}

// ||s|| test: the system's first workgroup applies it to the state it hands
// on; the product is formed either way (when the test fires, the consumer
// ignores it), so no other workgroup waits on the reduction.  Then the
// pending mark.
template <typename T, int NW>
__device__ __forceinline__ void spmv_state(const SpmvArgs<T> &a, const Group &g, int p, double *red, SolverState &sh)
{
    if (a.st_in && a.fin_in == FIN_S && p == a.sys[g.sys].g0) {
        const SysRange sr = a.sys[g.sys];
        if (threadIdx.x == 0) sh = a.st_in[g.sys];
        double v[1];
        reduce_parts<1, NW>(a.part_in, a.nparts_in, sr.v0, sr.v1, v, red);
        if (threadIdx.x == 0) {
            apply_fin(sh, FIN_S, v[0], 0.0);
            a.st_out[g.sys] = sh;
        }
    }
    if (a.st_mark && p == a.sys[g.sys].g0 && threadIdx.x == 0) a.st_mark[g.sys].pending = 1;
}

// The workgroup's partials of the fused dots, one per dot.
template <typename T, int EPI, int NW>
__device__ __forceinline__ void spmv_parts(const SpmvArgs<T> &a, int p, const Acc<T> &d0, const Acc<T> &d1,
                                           double *red)
{
    constexpr int ND = EpiDots<EPI>::n;
    if constexpr (ND > 0) {
        double v[ND];
        v[0] = d0.get();
        if constexpr (ND > 1) v[1] = d1.get();
        block_sum<ND, NW>(v, red);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < ND; ++q) a.part_out[(size_t)q * a.ngroups + p] = v[q];
        }
    }
}

// Workgroup barrier ordering LDS accesses only.  __syncthreads' release
// fence makes a wave wait for its global stores, and on CDNA stores and
// loads share one counter, so it waits for every load in flight too.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// TH threads, NPT entries per thread per pass: a row block of up to
// block_nnz entries takes ceil(block_nnz / (NPT * TH)) passes, each storing
// its products in LDS, before the row sums.  The passes of a workgroup's
// blocks form one stream, software-pipelined over two register sets: the
// stream loads of pass i + 1 are issued right after the x gathers of pass i,
// so they are in flight while pass i's products are stored and, after a
// block's last pass, while its rows are summed.  A finished system returns
// before any load; otherwise the first pass's loads are issued before the
// rest of the state prologue, so they overlap it.  PRE (x-segmented
// matrices): the passes load the products k_xseg formed, with their
// positions, instead of gathering x.
template <typename T, int EPI, int NPT, int TH, bool PRE, int DG = 0>
__global__ __launch_bounds__(TH) __attribute__((amdgpu_waves_per_eu(4))) void k_spmv(SpmvArgs<T> a)
{
    constexpr int NW = TH / 64;
    constexpr int PASS = NPT * TH;
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned short *lrp = reinterpret_cast<unsigned short *>(smem);
    T *lds = reinterpret_cast<T *>(smem + rowptr_lds_bytes<TH>());
    __shared__ double red[kMaxDots * NW];
    __shared__ SolverState sh;

    const int p = xcd_remap(blockIdx.x, a.ngroups);
    const Group g = a.groups[p];
    if (spmv_skip(a, g, p)) return;
    Cursor cur{g.rb0, 0, a.blocks[g.rb0], a.aux[g.rb0]};
    PassLoad<T, NPT> LA, LB;
    if (cur.ax.kind != 1) issue_pass<T, NPT, TH, PRE>(a, cur.rb, cur.ax, 0, LA);
    T w0 = issue_w0<T, EPI>(a, cur.rb, cur.ax);
    spmv_state<T, NW>(a, g, p, red, sh);

    Acc<T> d0, d1;
    // PRE: the block's rperm rows (loaded with its first pass) and its
    // epilogue operands
    int rqv[kRowsPerThread] = {0, 0, 0, 0};
    EpiOps<T, DG> eo;
    // One pass of the stream: the current pass's loads are in `in`; the next
    // pass's go to `out`.  Returns false after the workgroup's last block.
    auto step = [&](const PassLoad<T, NPT> &in, PassLoad<T, NPT> &out) -> bool {
        const int nz = cur.rb.j1 - cur.rb.j0;
        const bool gather = cur.ax.kind != 1;
        const bool last = !gather || cur.base + PASS >= nz;
        T xv[NPT];
        // kind-2 blocks with packed positions: column in the low pos_shift bits
        const bool packed = cur.ax.kind == 2 && a.pos_shift;
        const unsigned cmask = packed ? (1u << a.pos_shift) - 1u : ~0u;
        if (gather && !PRE) {
            // every lane gathers (past the block's end, at a column of the
            // block or the padding's column 0); 32-bit byte offset from the
            // uniform base (n * sizeof(T) < 2^32)
#pragma unroll
            for (int i = 0; i < NPT; ++i)
                xv[i] = *reinterpret_cast<const T *>(reinterpret_cast<const char *>(a.x) +
                                                     ((unsigned)in.c[i] & cmask) * (unsigned)sizeof(T));
        }
        // first pass of a block: its row pointers, staged in LDS below (PRE:
        // loaded with the pass, one step earlier)
        int rpv[kRowsPerThread];
        const bool stage = gather && cur.base == 0;
        if constexpr (PRE) {
            // every pass, before its scatter and the next pass's loads; the
            // rperm rows come with a block's first pass (a long-row block
            // loads no pass: its `in` is another block's, or unset)
            if (a.rperm) {
#pragma unroll
                for (int q = 0; q < kRowsPerThread; ++q) rqv[q] = stage ? in.rq[q] : rqv[q];
            }
            epi_load<T, EPI, TH, DG>(a, cur.rb, cur.ax, rqv, eo);
        }
        if (stage) {
#pragma unroll
            for (int q = 0; q < kRowsPerThread; ++q)
                rpv[q] = PRE ? in.rp[q]
                             : a.rowptr[cur.rb.r0 + min((int)threadIdx.x + q * TH, cur.rb.r1 - cur.rb.r0)];
        }
        Cursor nxt = cur;
        if (last) {
            nxt.b = cur.b + 1;
            nxt.base = 0;
            if (nxt.b < g.rb1) {
                nxt.rb = a.blocks[nxt.b];
                nxt.ax = a.aux[nxt.b];
            }
        } else {
            nxt.base = cur.base + PASS;
        }
        const bool more = nxt.b < g.rb1;
        if constexpr (!PRE) {
            if (more && nxt.ax.kind != 1) issue_pass<T, NPT, TH, PRE>(a, nxt.rb, nxt.ax, nxt.base, out);
        }
        if (gather) {
            const int lim = nz - cur.base;
            const bool sorted = cur.ax.kind == 2;
#pragma unroll
            for (int i = 0; i < NPT; ++i) {
                const int k = (int)threadIdx.x + i * TH;
                if constexpr (PRE) {
                    if (k < lim) lds[in.pos[i]] = in.v[i];
                } else {
                    const int at = !sorted ? cur.base + k
                                           : packed ? (int)((unsigned)in.c[i] >> a.pos_shift) : (int)in.pos[i];
                    if (k < lim) lds[at] = in.v[i] * xv[i];
                }
            }
        }
        if (stage) {
#pragma unroll
            for (int q = 0; q < kRowsPerThread; ++q) {
                const int t = (int)threadIdx.x + q * TH;
                if (t <= cur.rb.r1 - cur.rb.r0) lrp[t] = (unsigned short)(rpv[q] - cur.rb.j0);
            }
        }
        // PRE: one register set -- the next pass loads into it once this
        // pass's products and row pointers are in LDS, still ahead of the
        // row sums
        if constexpr (PRE) {
            if (more && nxt.ax.kind != 1) issue_pass<T, NPT, TH, PRE>(a, nxt.rb, nxt.ax, nxt.base, out);
        }
        if (last) {
            if (!gather) {
                // long row: ordered sum of its chunk sums.  The first wave
                // loads up to 64 of them at once and every lane adds them
                // in chunk order from its neighbours' registers, so the
                // loads are one round trip per 64 chunks, not one per chunk.
                if (threadIdx.x < 64) {
                    const T wr = reads_w<EPI>() && threadIdx.x == 0 ? a.w[cur.rb.r0] : T(0);
                    const DiagOp<T, DG> dg = diag_load<T, DG>(a, cur.rb.r0);
                    T total = T(0);
                    for (int c0 = 0; c0 < cur.ax.nchunks; c0 += 64) {
                        const int c = c0 + (int)threadIdx.x;
                        const T v = c < cur.ax.nchunks ? (T)a.chunk_sums[cur.ax.first + c] : T(0);
                        const int lim = min(64, cur.ax.nchunks - c0);
                        for (int l = 0; l < lim; ++l) total += __shfl(v, l, 64);
                    }
                    if (threadIdx.x == 0) row_out<T, EPI>(a, cur.rb.r0, diag_add<T, DG>(total, dg), wr, d0, d1);
                }
            } else {
                // LDS-only barriers: the next pass's loads stay in flight
                // through the row sums
                lds_barrier();
                sum_rows<T, EPI, TH, PRE, DG>(a, cur.rb, cur.ax, w0, lds, lrp, rqv, eo, d0, d1);
                lds_barrier();   // LDS is reused by the next row block
            }
            if (more) w0 = issue_w0<T, EPI>(a, nxt.rb, nxt.ax);
        }
        cur = nxt;
        return more;
    };
    if constexpr (PRE) {
        while (step(LA, LA)) {
        }
    } else {
        for (;;) {
            if (!step(LA, LB)) break;
            if (!step(LB, LA)) break;
        }
    }
    spmv_parts<T, EPI, NW>(a, p, d0, d1, red);
}

// A row block's descriptors loaded by vector loads, made uniform for the
// scalar address arithmetic that uses them.
__device__ __forceinline__ RowBlock uniform_rb(RowBlock r)
{
    r.r0 = __builtin_amdgcn_readfirstlane(r.r0);
    r.r1 = __builtin_amdgcn_readfirstlane(r.r1);
    r.j0 = __builtin_amdgcn_readfirstlane(r.j0);
    r.j1 = __builtin_amdgcn_readfirstlane(r.j1);
    return r;
}
__device__ __forceinline__ BlockAux uniform_ax(BlockAux x)
{
    x.kind = __builtin_amdgcn_readfirstlane(x.kind);
    x.lanes = __builtin_amdgcn_readfirstlane(x.lanes);
    x.nchunks = __builtin_amdgcn_readfirstlane(x.nchunks);
    x.first = __builtin_amdgcn_readfirstlane(x.first);
    return x;
}

// One 3x3 block of a kind-3 row block per thread: its first column and its
// 9 values, v[3a + c] = row a, column c of the block.
template <typename T>
struct B3Load {
    int bc;
    T v[9];
};

// The loads of row block rb's 3x3 blocks: thread t takes its block t, whose
// values sit in the row block's 9 planes of nbk values (plane 3a + c holds
// entry (a, c) of every block), so each load instruction reads 64
// consecutive values.  Threads past the row block's nbk blocks read the
// next ones (the arrays are padded by kWideThreads entries past the last
// block) and are ignored later: no load is clamped or behind a branch.
template <typename T>
__device__ __forceinline__ void issue_b3(const SpmvArgs<T> &a, const RowBlock &rb, B3Load<T> &L)
{
    const int nbk = (rb.j1 - rb.j0) / 9;
    const int *__restrict__ cb = a.bcol + rb.j0 / 9;
    L.bc = cb[threadIdx.x];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const T *__restrict__ plane = a.bval + rb.j0 + q * nbk;
        L.v[q] = plane[threadIdx.x];
    }
}

// Sums the rows of a kind-3 block (serial, CSR order) and feeds the
// epilogue: row t in straight-line code (no loop join ahead of its epilogue
// operand, prefetched as w0, and its row pointers, rp); rows past TH (block
// rows of fewer than 3 blocks) one at a time after it.
template <typename T, int EPI, int TH, bool VAR>
__device__ __forceinline__ void sum_rows3(const SpmvArgs<T> &a, const RowBlock &rb, const BlockAux &ax, T w0, int2 rp,
                                          const T *lds, Acc<T> &d0, Acc<T> &d1)
{
    const int nr = rb.r1 - rb.r0;
    const int t = (int)threadIdx.x;
    int kb, off;
    b3_row<VAR>(rb, ax, min(t, nr - 1), rp, kb, off);
    const T acc = sum_b3<T>(lds + off, kb);
    if (t < nr) row_out<T, EPI>(a, rb.r0 + t, acc, w0, d0, d1);
#pragma unroll 1
    for (int q = 1; q < kRowsPerThread && q * TH < nr; ++q) {
        const int tq = t + q * TH, tc = min(tq, nr - 1);
        const T wr = reads_w<EPI>() ? a.w[rb.r0 + tc] : T(0);
        const int2 rq = VAR ? make_int2(a.rowptr[rb.r0 + tc], a.rowptr[rb.r0 + tc + 1]) : make_int2(0, 0);
        b3_row<VAR>(rb, ax, tc, rq, kb, off);
        const T aq = sum_b3<T>(lds + off, kb);
        if (tq < nr) row_out<T, EPI>(a, rb.r0 + tq, aq, wr, d0, d1);
    }
}

// SpMV of a matrix of 3x3 blocks (kind-3 row blocks of at most TH blocks,
// one pass each), bit-identical to k_spmv's serial rows: thread t forms the
// 9 products of block t (its three x values gathered once) and stores them
// block-major in LDS, where the rows are summed in CSR order (sum_b3).  The
// stream reads 8 + 4/9 bytes per entry, and per 9 entries a thread issues
// 13 loads (1 column, 9 values, 3 x values) from uniform bases.  The loads
// of the workgroup's next row block are issued right after this one's x
// gathers, unconditionally (the last row block loads itself again), so no
// branch join makes the products wait for them: they stay in flight while
// the products are stored and the rows summed.  The descriptors of the row
// block after that are loaded then too, a step before their use.
template <typename T, int EPI, bool VAR>
__global__ __launch_bounds__(kWideThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_spmv3(SpmvArgs<T> a)
{
    constexpr int TH = kWideThreads, NW = TH / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    // 9 * TH products: every thread stores its block's, past nbk too
    T *lds = reinterpret_cast<T *>(smem);
    __shared__ double red[kMaxDots * NW];
    __shared__ SolverState sh;

    const int p = xcd_remap(blockIdx.x, a.ngroups);
    const Group g = a.groups[p];
    if (spmv_skip(a, g, p)) return;
    Cursor cur{g.rb0, 0, a.blocks[g.rb0], a.aux[g.rb0]};
    B3Load<T> LA, LB;
    issue_b3<T>(a, cur.rb, LA);
    // descriptors of row block min(cur.b + 1, g.rb1 - 1)
    RowBlock prb = a.blocks[min(g.rb0 + 1, g.rb1 - 1)];
    BlockAux pax = a.aux[min(g.rb0 + 1, g.rb1 - 1)];
    T w0 = issue_w0<T, EPI>(a, cur.rb, cur.ax);
    int2 rp = issue_rp<VAR>(a.rowptr, cur.rb);
    spmv_state<T, NW>(a, g, p, red, sh);

    Acc<T> d0, d1;
    auto step = [&](const B3Load<T> &in, B3Load<T> &out) -> bool {
        // the block's three x values: 32-bit byte offset from the uniform
        // base (n * sizeof(T) < 2^32), the columns as immediate offsets
        const T *__restrict__ xp =
            reinterpret_cast<const T *>(reinterpret_cast<const char *>(a.x) + (unsigned)in.bc * (unsigned)sizeof(T));
        const T x0 = xp[0], x1 = xp[1], x2 = xp[2];
        const bool more = cur.b + 1 < g.rb1;
        const int nb = more ? cur.b + 1 : cur.b;
        const RowBlock nrb = uniform_rb(prb);
        const BlockAux nax = uniform_ax(pax);
        issue_b3<T>(a, nrb, out);
        prb = a.blocks[min(nb + 1, g.rb1 - 1)];
        pax = a.aux[min(nb + 1, g.rb1 - 1)];
        T *q = lds + 9 * (int)threadIdx.x;
        q[0] = in.v[0] * x0;
        q[1] = in.v[1] * x1;
        q[2] = in.v[2] * x2;
        q[3] = in.v[3] * x0;
        q[4] = in.v[4] * x1;
        q[5] = in.v[5] * x2;
        q[6] = in.v[6] * x0;
        q[7] = in.v[7] * x1;
        q[8] = in.v[8] * x2;
        lds_barrier();
        sum_rows3<T, EPI, TH, VAR>(a, cur.rb, cur.ax, w0, rp, lds, d0, d1);
        lds_barrier();   // LDS is reused by the next row block
        w0 = issue_w0<T, EPI>(a, nrb, nax);
        rp = issue_rp<VAR>(a.rowptr, nrb);
        cur = Cursor{nb, 0, nrb, nax};
        return more;
    };
    for (;;) {
        if (!step(LA, LB)) break;
        if (!step(LB, LA)) break;
    }
    spmv_parts<T, EPI, NW>(a, p, d0, d1, red);
}

// One workgroup per chunk of a long row: ordered tree sum of its products
// (formed here, or by k_xseg for x-segmented matrices).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_chunks(ChunkArgs<T> a)
{
    __shared__ T red[kWaves];
    static_assert(kChunkLen <= 4 * kThreads, "a chunk is at most four strides");
    const Chunk c = a.chunks[blockIdx.x];
    if (a.st && a.st[c.sys].done) return;
    // thread t adds the terms t, t + 256, ... of the chunk in turn; the
    // four strides' terms load at once (indices clamped into the chunk), so
    // the sum waits for one round trip, not four
    const int j = c.j0 + (int)threadIdx.x;
    T t[4];
    if (a.prod) {
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = a.prod[min(j + k * kThreads, c.j1 - 1)];
    } else {
        int cc[4];
        T vv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int jk = min(j + k * kThreads, c.j1 - 1);
            cc[k] = a.col[jk];
            vv[k] = a.val[jk];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = vv[k] * a.x[cc[k]];
    }
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (j + k * kThreads < c.j1) acc += t[k];
    acc = block_sum1<T>(acc, red);
    if (threadIdx.x == 0) a.chunk_sums[blockIdx.x] = (double)acc;
}

// x-segmented SpMV, product pass (layout in XsegArgs).  For columns spread
// over all of x, the gathers of the CSR-stream SpMV each miss L2 and are
// bound by the misses in flight.  Here workgroup s copies its x segment
// (kSegCols values, coalesced) and its chunk tables into LDS once, then
// streams its entries: column within the segment, value, and the chunk that
// places the product in the row-block-major product array, which
// k_spmv<PRE> / k_chunks then sum in the usual order.  The passes are
// software-pipelined over two register sets (the next pass's streams load
// while this pass's products are formed and stored).  The workgroups of
// neighbouring segments share an XCD, so the chunks one XCD writes into a
// row block are neighbours.
template <typename T, int NPT>
struct XLoad {
    int c[NPT], d[NPT];
    T v[NPT];
};

template <typename T, int NPT>
__device__ __forceinline__ void xseg_issue(const XsegArgs<T> &a, int g0, int lim, int base, XLoad<T, NPT> &L)
{
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int f = base + i * kWideThreads + (int)threadIdx.x;
        const bool in = f < lim;
        L.c[i] = in ? (int)a.xcol[g0 + f] : 0;
        L.d[i] = in ? (int)a.delta[g0 + f] : 0;
        L.v[i] = in ? a.xval[g0 + f] : T(0);
    }
}

// Solver prologue of a fused product pass (XsegArgs::fuse): the state
// after the pending reduction (FIN_XR for XFUSE_P, FIN_ALPHA for XFUSE_S),
// derived by every workgroup from the same partials in the same order;
// workgroup 0 hands it on.  Returns whether the system is done.
template <typename T, int FUSE>
__device__ __forceinline__ bool xseg_state(const XsegArgs<T> &a, SolverState &sh, double *red)
{
    constexpr int NW = kWideThreads / 64;
    if (threadIdx.x == 0) sh = *a.st_in;
    __syncthreads();
    const bool act = FUSE == XFUSE_P ? (!sh.done && sh.pending) : !sh.done;
    if (act) {
        double v[2] = {0.0, 0.0};
        if constexpr (FUSE == XFUSE_P) {
            reduce_parts<2, NW>(a.part_in, a.nparts_in, a.pin0, a.pin1, v, red);
        } else {
            double v1[1];
            reduce_parts<1, NW>(a.part_in, a.nparts_in, a.pin0, a.pin1, v1, red);
            v[0] = v1[0];
        }
        if (threadIdx.x == 0) apply_fin(sh, FUSE == XFUSE_P ? FIN_XR : FIN_ALPHA, v[0], v[1]);
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.st_out = sh;
    return sh.done != 0;
}

// Loads of one round of a fused segment update: U elements per thread,
// i0 + u * kWideThreads (clamped into the segment) of R, V (and P).
template <typename T>
constexpr int stage_unroll()
{
    return sizeof(T) == 4 ? 8 : 4;
}
template <typename T, int FUSE>
struct StageLoad {
    static constexpr int U = stage_unroll<T>();
    T r[U], v[U], p[U];
};
template <typename T, int FUSE>
__device__ __forceinline__ void stage_issue(const XsegArgs<T> &a, int c0, int nc, int i0, StageLoad<T, FUSE> &L)
{
#pragma unroll
    for (int u = 0; u < StageLoad<T, FUSE>::U; ++u) {
        const int idx = c0 + min(i0 + u * kWideThreads, nc - 1);
        L.r[u] = a.R[idx];
        L.v[u] = a.V[idx];
        L.p[u] = FUSE == XFUSE_P ? a.x[idx] : T(0);
    }
}

// The segment's x staged in LDS.  Fused: each element of the vector is
// updated here (the same arithmetic and roundings as k_vec_P / k_vec_S),
// written out by the segment's first share and, for XFUSE_S, its square
// added to this workgroup's ||S||^2 partial (written by every workgroup, 0
// by the other shares).  The first round's loads (L) were issued before the
// state prologue; each later round loads after the previous one is stored.
template <typename T, int FUSE>
__device__ __forceinline__ void xseg_stage(const XsegArgs<T> &a, const SolverState &sh, T *xs, int c0, int nc,
                                           bool first_share, int w, double *red, StageLoad<T, FUSE> &L)
{
    if constexpr (FUSE == XFUSE_NONE) {
        for (int i = (int)threadIdx.x; i < nc; i += kWideThreads) xs[i] = a.x[c0 + i];
    } else {
        constexpr int NW = kWideThreads / 64, U = StageLoad<T, FUSE>::U;
        const T om = (T)sh.omega, be = (T)sh.beta, al = (T)sh.alpha;
        Acc<T> d;
        for (int i0 = (int)threadIdx.x;;) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kWideThreads;
                if (i < nc) {
                    T o;
                    if constexpr (FUSE == XFUSE_P) o = ((L.p[u] - L.v[u] * om) * be) + L.r[u];
                    else o = L.r[u] - L.v[u] * al;
                    xs[i] = o;
                    if (first_share) {
                        a.xv_out[c0 + i] = o;
                        if constexpr (FUSE == XFUSE_S) d.add(o, o);
                    }
                }
            }
            i0 += U * kWideThreads;
            if (i0 >= nc) break;
            stage_issue<T, FUSE>(a, c0, nc, i0, L);
        }
        if constexpr (FUSE == XFUSE_S) {
            double v1[1] = {d.get()};
            block_sum<1, NW>(v1, red);
            if (threadIdx.x == 0) a.part_out[w] = v1[0];
        }
    }
}

template <typename T, int NPT, int FUSE>
__global__ __launch_bounds__(kWideThreads) void k_xseg(XsegArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *xs = reinterpret_cast<T *>(smem);
    int *offl = reinterpret_cast<int *>(smem + (size_t)a.seg_cols * sizeof(T));
    int *gchl = offl + a.max_chunks;
    __shared__ double red[kMaxDots * (kWideThreads / 64)];
    __shared__ SolverState sh;
    // workgroup w takes share w % split of segment w / split (the shares of
    // one segment and neighbouring segments on one XCD); entries f in
    // [f0, f1) of the segment, numbered from its first entry
    const int w = xcd_remap(blockIdx.x, a.nseg * a.split);
    const int s = w / a.split, part = w - s * a.split;
    if constexpr (FUSE == XFUSE_NONE) {
        if (a.st && a.st->done) return;
    }
    const int g0 = a.sbase[s], cnt = a.sbase[s + 1] - g0;
    const int f0 = (int)((long long)cnt * part / a.split), f1 = (int)((long long)cnt * (part + 1) / a.split);
    // fused: every share updates the segment (its first writes it) and
    // writes its partial, with or without entries of its own
    if (FUSE == XFUSE_NONE && f1 <= f0) return;
    XLoad<T, NPT> LA, LB;
    const int c0 = s * a.seg_cols, nc = min(a.seg_cols, a.n - c0);
    StageLoad<T, FUSE> L;
    if constexpr (FUSE != XFUSE_NONE) {
        // the update's first loads overlap the state prologue; the entries'
        // first pass loads behind it (fewer registers live across it)
        stage_issue<T, FUSE>(a, c0, nc, (int)threadIdx.x, L);
        if (xseg_state<T, FUSE>(a, sh, red)) return;
    }
    xseg_issue<T, NPT>(a, g0, f1, f0, LA);
    xseg_stage<T, FUSE>(a, sh, xs, c0, nc, part == 0, w, red, L);
    const int cb = a.cbase[s], ncs = a.cbase[s + 1] - cb;
    for (int i = (int)threadIdx.x; i < ncs; i += kWideThreads) offl[i] = a.off[cb + i];
    const int gb = a.gbase[s], ngs = a.gbase[s + 1] - gb;
    for (int i = (int)threadIdx.x; i < ngs; i += kWideThreads) gchl[i] = a.gch[gb + i];
    __syncthreads();
    if (FUSE != XFUSE_NONE && f1 <= f0) return;
    T *__restrict__ prod = a.prod + g0;
    constexpr int PASS = NPT * kWideThreads;
    auto step = [&](const XLoad<T, NPT> &in, XLoad<T, NPT> &out, int base) -> bool {
        const bool more = base + PASS < f1;
        if (more) xseg_issue<T, NPT>(a, g0, f1, base + PASS, out);
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            const int f = base + i * kWideThreads + (int)threadIdx.x;
            if (f < f1) prod[f + offl[gchl[f >> 6] + in.d[i]]] = in.v[i] * xs[in.c[i]];
        }
        return more;
    };
    for (int base = f0;; base += 2 * PASS) {
        if (!step(LA, LB, base)) break;
        if (!step(LB, LA, base + PASS)) break;
    }
}

// ---- streaming vector kernels --------------------------------------------

template <typename T>
struct alignas(16) Pk {
    static constexpr int N = 16 / sizeof(T);
    T v[N];
};

// Load / store the 16-byte pack at element idx, masked against [e0, e1)
// (elements outside load as 0 and are not stored).
template <typename T>
__device__ __forceinline__ Pk<T> ldm(const T *p, int idx, const VecBlock &vb)
{
    Pk<T> r;
    if (idx >= vb.e0 && idx + Pk<T>::N <= vb.e1) {
        r = *reinterpret_cast<const Pk<T> *>(p + idx);
    } else {
#pragma unroll
        for (int k = 0; k < Pk<T>::N; ++k) r.v[k] = (idx + k >= vb.e0 && idx + k < vb.e1) ? p[idx + k] : T(0);
    }
    return r;
}
template <typename T>
__device__ __forceinline__ void stm(T *p, int idx, const VecBlock &vb, const Pk<T> &x)
{
    if (idx >= vb.e0 && idx + Pk<T>::N <= vb.e1) {
        *reinterpret_cast<Pk<T> *>(p + idx) = x;
    } else {
#pragma unroll
        for (int k = 0; k < Pk<T>::N; ++k)
            if (idx + k >= vb.e0 && idx + k < vb.e1) p[idx + k] = x.v[k];
    }
}

// Element index of pack u of this thread (base is 16-byte aligned).
template <typename T>
__device__ __forceinline__ int pk_idx(const VecBlock &vb, int u)
{
    return vb.base + (u * kThreads + (int)threadIdx.x) * Pk<T>::N;
}

template <typename T>
__device__ __forceinline__ bool first_of_sys(const VecArgs<T> &a, const VecBlock &vb)
{
    return (int)blockIdx.x == a.sys[vb.sys].v0;
}

// Sum the producer's partials of this system (all threads) and apply fin to
// the shared state (thread 0).
template <int ND, typename T>
__device__ __forceinline__ void apply_in(const VecArgs<T> &a, const VecBlock &vb, SolverState &sh, int fin,
                                         double *red)
{
    const SysRange sr = a.sys[vb.sys];
    double v[ND];
    if (a.in_spmv) reduce_parts<ND>(a.part_in, a.nparts_in, sr.g0, sr.g1, v, red);
    else reduce_parts<ND>(a.part_in, a.nparts_in, sr.v0, sr.v1, v, red);
    if (threadIdx.x == 0) apply_fin(sh, fin, v[0], ND > 1 ? v[ND - 1] : 0.0);
    __syncthreads();
}

template <int ND, typename T>
__device__ __forceinline__ void write_parts(const VecArgs<T> &a, Acc<T> &d0, Acc<T> &d1, double *red)
{
    double v[ND];
    v[0] = d0.get();
    if constexpr (ND > 1) v[1] = d1.get();
    block_sum<ND>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < ND; ++q) a.part_out[(size_t)q * a.nvb + blockIdx.x] = v[q];
    }
}

// Start of an iteration: apply the pending (r, r0), ||r||^2 of the previous
// one (stop tests, beta), then P = ((P - V*omega)*beta) + R
// (bicgstab_cpu.c:203-213; k_vecop_P).
template <typename T, int U>
__global__ __launch_bounds__(kThreads) void k_vec_P(VecArgs<T> a)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ SolverState sh;
    const VecBlock vb = a.blocks[blockIdx.x];
    if (threadIdx.x == 0) sh = a.st_in[vb.sys];
    Pk<T> p[U], v[U], r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = pk_idx<T>(vb, u);
        p[u] = ldm(a.P, i, vb);
        v[u] = ldm(a.V, i, vb);
        r[u] = ldm(a.R, i, vb);
    }
    __syncthreads();
    if (!sh.done && sh.pending) apply_in<2, T>(a, vb, sh, FIN_XR, red);
    if (first_of_sys(a, vb) && threadIdx.x == 0) a.st_out[vb.sys] = sh;
    if (sh.done) return;
    const T om = (T)sh.omega, be = (T)sh.beta;
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int k = 0; k < Pk<T>::N; ++k) p[u].v[k] = ((p[u].v[k] - v[u].v[k] * om) * be) + r[u].v[k];
        stm(a.P, pk_idx<T>(vb, u), vb, p[u]);
    }
}

// alpha = rho / (r0, v) from the SpMV partials; S = R - V*alpha, ||S||^2
// (bicgstab_cpu.c:219-223; k_vecop_S).
template <typename T, int U>
__global__ __launch_bounds__(kThreads) void k_vec_S(VecArgs<T> a)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ SolverState sh;
    const VecBlock vb = a.blocks[blockIdx.x];
    if (threadIdx.x == 0) sh = a.st_in[vb.sys];
    Pk<T> v[U], r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = pk_idx<T>(vb, u);
        v[u] = ldm(a.V, i, vb);
        r[u] = ldm(a.R, i, vb);
    }
    __syncthreads();
    if (!sh.done) apply_in<1, T>(a, vb, sh, FIN_ALPHA, red);
    if (first_of_sys(a, vb) && threadIdx.x == 0) a.st_out[vb.sys] = sh;
    if (sh.done) return;
    const T al = (T)sh.alpha;
    Acc<T> d0, d1;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        Pk<T> s;
#pragma unroll
        for (int k = 0; k < Pk<T>::N; ++k) {
            s.v[k] = r[u].v[k] - v[u].v[k] * al;
            d0.add(s.v[k], s.v[k]);
        }
        stm(a.S, pk_idx<T>(vb, u), vb, s);
    }
    write_parts<1, T>(a, d0, d1, red);
}

// omega = (t, s) / (t, t) from the SpMV partials; X = (X + P*alpha) + S*omega;
// R = S - T*omega; partials (R, R0), ||R||^2 for the next iteration
// (bicgstab_cpu.c:229-242, :203; k_vecop_XR).  After the ||s|| test fired:
// X = X + P*alpha only (bicgstab_gpu.c:219-236).
template <typename T, int U>
__global__ __launch_bounds__(kThreads) void k_vec_XR(VecArgs<T> a)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ SolverState sh;
    const VecBlock vb = a.blocks[blockIdx.x];
    if (threadIdx.x == 0) sh = a.st_in[vb.sys];
    Pk<T> x[U], p[U], s[U], t[U], r0[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = pk_idx<T>(vb, u);
        x[u] = ldm(a.X, i, vb);
        p[u] = ldm(a.P, i, vb);
        s[u] = ldm(a.S, i, vb);
        t[u] = ldm(a.T_, i, vb);
        r0[u] = ldm(a.R0, i, vb);
    }
    __syncthreads();
    if (!sh.done && !sh.s_conv) apply_in<2, T>(a, vb, sh, FIN_OMEGA, red);
    if (threadIdx.x == 0 && !sh.done) sh.pending = 1;
    __syncthreads();
    if (first_of_sys(a, vb) && threadIdx.x == 0) a.st_out[vb.sys] = sh;
    if (sh.done) return;
    const T al = (T)sh.alpha, om = (T)sh.omega;
    if (sh.s_conv) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int k = 0; k < Pk<T>::N; ++k) x[u].v[k] = x[u].v[k] + p[u].v[k] * al;
            stm(a.X, pk_idx<T>(vb, u), vb, x[u]);
        }
        return;
    }
    Acc<T> d0, d1;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        Pk<T> r;
#pragma unroll
        for (int k = 0; k < Pk<T>::N; ++k) {
            x[u].v[k] = (x[u].v[k] + p[u].v[k] * al) + s[u].v[k] * om;
            r.v[k] = s[u].v[k] - t[u].v[k] * om;
            d0.add(r.v[k], r0[u].v[k]);
            d1.add(r.v[k], r.v[k]);
        }
        const int i = pk_idx<T>(vb, u);
        stm(a.X, i, vb, x[u]);
        stm(a.R, i, vb, r);
    }
    write_parts<2, T>(a, d0, d1, red);
}

// Apply a pending FIN_XR in place, one workgroup per system (end of a chunk
// of iterations, so the host sees the state of the last iteration), and
// store the state in the host-mapped mirror the host polls (if any).
__global__ __launch_bounds__(kThreads) void k_apply_xr(SolverState *st, const SysRange *sys, const double *part,
                                                      int nvb, SolverState *mirror, int stamp)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ SolverState sh;
    const int q = blockIdx.x;
    if (threadIdx.x == 0) sh = st[q];
    __syncthreads();
    if (!sh.done && sh.pending) {
        const SysRange sr = sys[q];
        double v[2];
        reduce_parts<2>(part, nvb, sr.v0, sr.v1, v, red);
        if (threadIdx.x == 0) {
            apply_fin(sh, FIN_XR, v[0], v[1]);
            st[q] = sh;
        }
    }
    if (mirror && threadIdx.x == 0) {
        mirror[q] = sh;
        if (stamp) {
            // the stamp lands after the state: a host that reads the stamp
            // then the state sees this chunk's state
            __threadfence_system();
            *reinterpret_cast<volatile int *>(&mirror[q].pad) = stamp;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_finalize(SolverState *st, const SysRange *sys, const double *part,
                                                      int nparts, int in_spmv, int fin)
{
    __shared__ double red[kMaxDots * kWaves];
    __shared__ SolverState sh;
    const int q = blockIdx.x;
    if (threadIdx.x == 0) sh = st[q];
    const SysRange sr = sys[q];
    const int p0 = in_spmv ? sr.g0 : sr.v0, p1 = in_spmv ? sr.g1 : sr.v1;
    double v[2] = {0.0, 0.0};
    if (fin == FIN_RESIDB) {
        reduce_parts<2>(part, nparts, p0, p1, v, red);
    } else {
        double v1[1];
        reduce_parts<1>(part, nparts, p0, p1, v1, red);
        v[0] = v1[0];
    }
    if (threadIdx.x == 0) {
        apply_fin(sh, fin, v[0], v[1]);
        st[q] = sh;
    }
}

// R0 = R ; P = V = 0 (bicgstab_cpu.c:180-188); partials (R, R0) = ||R||^2
// for the first iteration's rho and the threshold tol * ||r0||, pending in
// st_out.
__global__ __launch_bounds__(kThreads) void k_init_vectors(VecArgs<double> a)
{
    __shared__ double red[kMaxDots * kWaves];
    const VecBlock vb = a.blocks[blockIdx.x];
    Acc<double> d0, d1;
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) {
        const double r = a.R[i];
        const_cast<double *>(a.R0)[i] = r;
        a.P[i] = 0.0;
        a.V[i] = 0.0;
        d0.add(r, r);
    }
    d1 = d0;
    if (first_of_sys(a, vb) && threadIdx.x == 0) a.st_out[vb.sys].pending = 1;
    write_parts<2, double>(a, d0, d1, red);
}

// Mixed sweep start (FP32 inner solve of A d = r / ||r||).
__global__ __launch_bounds__(kThreads) void k_ir_start(VecArgs<float> a, const double *r64,
                                                      const SolverState *outer)
{
    __shared__ double red[kMaxDots * kWaves];
    const VecBlock vb = a.blocks[blockIdx.x];
    const double nr = sqrt(outer[vb.sys].resid2);
    const double inv = nr > 0.0 ? 1.0 / nr : 0.0;
    Acc<float> d0, d1;
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) {
        const float r = (float)(r64[i] * inv);
        a.R[i] = r;
        const_cast<float *>(a.R0)[i] = r;
        a.X[i] = 0.f;
        a.P[i] = 0.f;
        a.V[i] = 0.f;
        d0.add(r, r);
    }
    d1 = d0;
    if (first_of_sys(a, vb) && threadIdx.x == 0) {
        a.st_out[vb.sys].scale = nr;
        a.st_out[vb.sys].pending = 1;
    }
    write_parts<2, float>(a, d0, d1, red);
}

__global__ __launch_bounds__(kThreads) void k_ir_update(VecArgs<float> a, const double *x_in, double *x_out)
{
    const VecBlock vb = a.blocks[blockIdx.x];
    const double sc = a.st_in[vb.sys].scale;
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) x_out[i] = x_in[i] + (double)a.X[i] * sc;
}

// States into the host-mapped mirror the host spins on: the copies first,
// then, after a system-scope fence, the stamp in each copy.
__global__ __launch_bounds__(kThreads) void k_mirror(const SolverState *a, const SolverState *b, int n,
                                                    SolverState *mirror, int stamp)
{
    const int tot = b ? 2 * n : n;
    for (int q = threadIdx.x; q < tot; q += blockDim.x) mirror[q] = q < n ? a[q] : b[q - n];
    if (!stamp) return;
    __syncthreads();
    __threadfence_system();
    for (int q = threadIdx.x; q < tot; q += blockDim.x) *reinterpret_cast<volatile int *>(&mirror[q].pad) = stamp;
}

__global__ __launch_bounds__(kThreads) void k_dot(VecArgs<double> a, const double *x, const double *y)
{
    __shared__ double red[kMaxDots * kWaves];
    const VecBlock vb = a.blocks[blockIdx.x];
    Acc<double> d0, d1;
    for (int i = vb.e0 + threadIdx.x; i < vb.e1; i += kThreads) d0.add(x[i], y[i]);
    write_parts<1, double>(a, d0, d1, red);
}

// Start of a solve: X = x (grid-stride; none when x is null) and, in
// workgroup 0, the states.
__global__ __launch_bounds__(kThreads) void k_solve_start(const double *x, double *X, int n, SolverState *st,
                                                         int nstates, double tol, int maxit, int fixed)
{
    if (x)
        for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) X[i] = x[i];
    if (blockIdx.x != 0) return;
    for (int q = threadIdx.x; q < nstates; q += blockDim.x) {
        SolverState z{};
        z.rho = 1.0;
        z.alpha = 1.0;
        z.omega = 1.0;
        z.tol = tol;
        z.maxit = maxit;
        z.fixed = fixed;
        z.iter = -1;
        st[q] = z;
    }
}

__global__ void k_state_init(SolverState *st, int nstates, double tol, int maxit, int fixed)
{
    for (int q = threadIdx.x; q < nstates; q += blockDim.x) {
        SolverState z{};
        z.rho = 1.0;
        z.alpha = 1.0;
        z.omega = 1.0;
        z.tol = tol;
        z.maxit = maxit;
        z.fixed = fixed;
        z.iter = -1;
        st[q] = z;
    }
}

template <typename T, int NPT, int TH, bool PRE = false, int DG = 0>
hipError_t spmv_epi(const SpmvArgs<T> &a, int epi, size_t shmem, hipStream_t s)
{
    switch (epi) {
    case EPI_AXPBY: k_spmv<T, EPI_AXPBY, NPT, TH, PRE, DG><<<a.ngroups, TH, shmem, s>>>(a); break;
    case EPI_DOT1: k_spmv<T, EPI_DOT1, NPT, TH, PRE, DG><<<a.ngroups, TH, shmem, s>>>(a); break;
    case EPI_DOT2: k_spmv<T, EPI_DOT2, NPT, TH, PRE, DG><<<a.ngroups, TH, shmem, s>>>(a); break;
    case EPI_RESID: k_spmv<T, EPI_RESID, NPT, TH, PRE, DG><<<a.ngroups, TH, shmem, s>>>(a); break;
    case EPI_RESID2: k_spmv<T, EPI_RESID2, NPT, TH, PRE, DG><<<a.ngroups, TH, shmem, s>>>(a); break;
    case EPI_RESIDB: k_spmv<T, EPI_RESIDB, NPT, TH, PRE, DG><<<a.ngroups, TH, shmem, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, bool VAR>
hipError_t spmv3_epi(const SpmvArgs<T> &a, int epi, hipStream_t s)
{
    const size_t shmem = spmv3_lds_bytes(sizeof(T));
    switch (epi) {
    case EPI_AXPBY: k_spmv3<T, EPI_AXPBY, VAR><<<a.ngroups, kWideThreads, shmem, s>>>(a); break;
    case EPI_DOT1: k_spmv3<T, EPI_DOT1, VAR><<<a.ngroups, kWideThreads, shmem, s>>>(a); break;
    case EPI_DOT2: k_spmv3<T, EPI_DOT2, VAR><<<a.ngroups, kWideThreads, shmem, s>>>(a); break;
    case EPI_RESID: k_spmv3<T, EPI_RESID, VAR><<<a.ngroups, kWideThreads, shmem, s>>>(a); break;
    case EPI_RESID2: k_spmv3<T, EPI_RESID2, VAR><<<a.ngroups, kWideThreads, shmem, s>>>(a); break;
    case EPI_RESIDB: k_spmv3<T, EPI_RESIDB, VAR><<<a.ngroups, kWideThreads, shmem, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Raise a kernel's dynamic LDS limit to at least `bytes` (above the 64 KiB
// default).  Limits only grow, so a plan made later with smaller blocks does
// not lower the limit an earlier plan's launches need.
hipError_t raise_lds_limit(const void *f, size_t bytes)
{
    static std::mutex mu;
    static std::map<const void *, size_t> limit;
    std::lock_guard<std::mutex> lock(mu);
    size_t &have = limit[f];
    if (bytes <= have) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) have = bytes;
    return e;
}

// Allow the dynamic LDS of the wide blocks.
template <typename T, int NPT, int TH, bool PRE = false, int DG = 0>
hipError_t spmv_allow_lds(size_t shmem)
{
    hipError_t e = hipSuccess;
    const void *k[] = {(const void *)k_spmv<T, EPI_AXPBY, NPT, TH, PRE, DG>,
                       (const void *)k_spmv<T, EPI_DOT1, NPT, TH, PRE, DG>,
                       (const void *)k_spmv<T, EPI_DOT2, NPT, TH, PRE, DG>,
                       (const void *)k_spmv<T, EPI_RESID, NPT, TH, PRE, DG>,
                       (const void *)k_spmv<T, EPI_RESID2, NPT, TH, PRE, DG>,
                       (const void *)k_spmv<T, EPI_RESIDB, NPT, TH, PRE, DG>};
    for (const void *f : k)
        if (e == hipSuccess) e = raise_lds_limit(f, shmem);
    return e;
}

template <typename T, bool VAR>
hipError_t spmv3_allow_lds()
{
    hipError_t e = hipSuccess;
    const void *k[] = {(const void *)k_spmv3<T, EPI_AXPBY, VAR>, (const void *)k_spmv3<T, EPI_DOT1, VAR>,
                       (const void *)k_spmv3<T, EPI_DOT2, VAR>,  (const void *)k_spmv3<T, EPI_RESID, VAR>,
                       (const void *)k_spmv3<T, EPI_RESID2, VAR>, (const void *)k_spmv3<T, EPI_RESIDB, VAR>};
    for (const void *f : k)
        if (e == hipSuccess) e = raise_lds_limit(f, spmv3_lds_bytes(sizeof(T)));
    return e;
}

// Row blocks up to 2048 entries: 256 threads, one pass; larger ones: 512
// threads, passes of npt * 512 entries (npt 4: under 72 VGPRs, so three such
// workgroups share a CU when their LDS fits; npt 8: under 128 VGPRs, two).
template <typename T>
hipError_t spmv_dispatch(const SpmvArgs<T> &a, int epi, int block_nnz, int npt, hipStream_t s)
{
    if (a.ngroups <= 0) return hipSuccess;
    if (block_nnz <= 0 || block_nnz > kMaxBlockNnz) return hipErrorInvalidValue;
    const size_t prod = (size_t)block_nnz * sizeof(T);
    if (a.dmode != 0 && !a.prod) return hipErrorInvalidValue;   // a diagonal apart only with x segments
    if (a.prod) {
        // x-segmented: 512-thread workgroups, one pass per row block
        if (spmv_threads(block_nnz) != kWideThreads || npt != kPreNpt) return hipErrorInvalidValue;
        const size_t shmem = prod + rowptr_lds_bytes<kWideThreads>();
        if constexpr (sizeof(T) == 4) {
            // FP32 copies of matrices with their diagonal apart
            if (a.dmode == 1) return spmv_epi<T, kPreNpt, kWideThreads, true, 1>(a, epi, shmem, s);
            if (a.dmode == 2) return spmv_epi<T, kPreNpt, kWideThreads, true, 2>(a, epi, shmem, s);
        }
        if (a.dmode != 0) return hipErrorInvalidValue;
        return spmv_epi<T, kPreNpt, kWideThreads, true>(a, epi, shmem, s);
    }
    if (a.bval) {
        // 3x3-block matrices: row blocks of at most kWideThreads blocks
        if (block_nnz != kB3BlockNnz || npt != kB3Npt) return hipErrorInvalidValue;
        return a.b3var ? spmv3_epi<T, true>(a, epi, s) : spmv3_epi<T, false>(a, epi, s);
    }
    if (block_nnz <= 4 * kThreads) return spmv_epi<T, 4, kThreads>(a, epi, prod + rowptr_lds_bytes<kThreads>(), s);
    if (block_nnz <= 8 * kThreads) return spmv_epi<T, 8, kThreads>(a, epi, prod + rowptr_lds_bytes<kThreads>(), s);
    const size_t shmem = prod + rowptr_lds_bytes<kWideThreads>();
    if (npt == 4) return spmv_epi<T, 4, kWideThreads>(a, epi, shmem, s);
    return spmv_epi<T, 8, kWideThreads>(a, epi, shmem, s);
}

template <typename T>
hipError_t chunk_dispatch(const ChunkArgs<T> &a, hipStream_t s)
{
    if (a.nchunks <= 0) return hipSuccess;
    k_chunks<T><<<a.nchunks, kThreads, 0, s>>>(a);
    return hipGetLastError();
}

// a.upb = 16-byte packs per thread = ceil(elements per block / (256 * pack)).
#define HBS_VEC_KERNEL(NAME)                                                      \
    template <typename T>                                                          \
    hipError_t NAME##_dispatch(const VecArgs<T> &a, hipStream_t s)                 \
    {                                                                              \
        if (a.nvb <= 0) return hipSuccess;                                         \
        switch (a.upb) {                                                           \
        case 1: NAME<T, 1><<<a.nvb, kThreads, 0, s>>>(a); break;                   \
        case 2: NAME<T, 2><<<a.nvb, kThreads, 0, s>>>(a); break;                   \
        case 4: NAME<T, 4><<<a.nvb, kThreads, 0, s>>>(a); break;                   \
        case 8: NAME<T, 8><<<a.nvb, kThreads, 0, s>>>(a); break;                   \
        default: return hipErrorInvalidValue;                                      \
        }                                                                          \
        return hipGetLastError();                                                  \
    }
HBS_VEC_KERNEL(k_vec_P)
HBS_VEC_KERNEL(k_vec_S)
HBS_VEC_KERNEL(k_vec_XR)
#undef HBS_VEC_KERNEL

}  // namespace

hipError_t prepare_spmv(int block_nnz, int npt)
{
    if (spmv_threads(block_nnz) == kThreads) return hipSuccess;
    const size_t rows = rowptr_lds_bytes<kWideThreads>();
    const size_t d = (size_t)block_nnz * sizeof(double) + rows, f = (size_t)block_nnz * sizeof(float) + rows;
    hipError_t e;
    if (npt == kPreNpt) {
        e = spmv_allow_lds<double, kPreNpt, kWideThreads, true>(d);
        if (e == hipSuccess) e = spmv_allow_lds<float, kPreNpt, kWideThreads, true>(f);
        if (e == hipSuccess) e = spmv_allow_lds<float, kPreNpt, kWideThreads, true, 1>(f);
        if (e == hipSuccess) e = spmv_allow_lds<float, kPreNpt, kWideThreads, true, 2>(f);
    } else if (npt == kB3Npt) {
        e = spmv3_allow_lds<double, false>();
        if (e == hipSuccess) e = spmv3_allow_lds<float, false>();
        if (e == hipSuccess) e = spmv3_allow_lds<double, true>();
        if (e == hipSuccess) e = spmv3_allow_lds<float, true>();
    } else if (npt == 4) {
        e = spmv_allow_lds<double, 4, kWideThreads>(d);
        if (e == hipSuccess) e = spmv_allow_lds<float, 4, kWideThreads>(f);
    } else {
        e = spmv_allow_lds<double, 8, kWideThreads>(d);
        if (e == hipSuccess) e = spmv_allow_lds<float, 8, kWideThreads>(f);
    }
    return e;
}
int spmv_resident_groups(int block_nnz, int npt, int device, bool pre)
{
    const int th = spmv_threads(block_nnz);
    const bool b3 = npt == kB3Npt;
    const size_t shmem = b3 ? spmv3_lds_bytes(sizeof(double))
                            : (size_t)block_nnz * sizeof(double) +
                                  (th == kWideThreads ? rowptr_lds_bytes<kWideThreads>() : rowptr_lds_bytes<kThreads>());
    const void *f = th == kThreads ? (block_nnz <= 4 * kThreads ? (const void *)k_spmv<double, EPI_DOT2, 4, kThreads, false>
                                                                 : (const void *)k_spmv<double, EPI_DOT2, 8, kThreads, false>)
                    : pre          ? (const void *)k_spmv<double, EPI_DOT2, kPreNpt, kWideThreads, true>
                    : b3           ? (const void *)k_spmv3<double, EPI_DOT2, false>
                                   : (npt == 4 ? (const void *)k_spmv<double, EPI_DOT2, 4, kWideThreads, false>
                                               : (const void *)k_spmv<double, EPI_DOT2, 8, kWideThreads, false>);
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, th, shmem) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return per_cu * cus;
}
hipError_t launch_spmv(const SpmvArgs<double> &a, int epi, int block_nnz, int npt, hipStream_t s)
{
    return spmv_dispatch<double>(a, epi, block_nnz, npt, s);
}
hipError_t launch_spmv(const SpmvArgs<float> &a, int epi, int block_nnz, int npt, hipStream_t s)
{
    return spmv_dispatch<float>(a, epi, block_nnz, npt, s);
}
hipError_t launch_chunks(const ChunkArgs<double> &a, int, hipStream_t s) { return chunk_dispatch<double>(a, s); }
hipError_t launch_chunks(const ChunkArgs<float> &a, int, hipStream_t s) { return chunk_dispatch<float>(a, s); }
size_t xseg_lds_bytes(int elem_bytes, int max_chunks, int max_granules)
{
    return (size_t)kSegCols * (size_t)elem_bytes + 4 * ((size_t)max_chunks + (size_t)max_granules);
}
hipError_t prepare_xseg(int max_chunks, int max_granules)
{
    hipError_t e = hipSuccess;
    const void *k[] = {(const void *)k_xseg<double, xseg_npt<double>(), XFUSE_NONE>,
                       (const void *)k_xseg<double, xseg_npt<double>(), XFUSE_P>,
                       (const void *)k_xseg<double, xseg_npt<double>(), XFUSE_S>,
                       (const void *)k_xseg<float, xseg_npt<float>(), XFUSE_NONE>,
                       (const void *)k_xseg<float, xseg_npt<float>(), XFUSE_P>,
                       (const void *)k_xseg<float, xseg_npt<float>(), XFUSE_S>};
    for (int i = 0; i < 6 && e == hipSuccess; ++i)
        e = raise_lds_limit(k[i], xseg_lds_bytes(i < 3 ? 8 : 4, max_chunks, max_granules));
    return e;
}
template <typename T>
hipError_t xseg_dispatch(const XsegArgs<T> &a, hipStream_t s)
{
    if (a.nseg <= 0) return hipSuccess;
    if (a.seg_cols != kSegCols || a.split < 1 || (long long)a.nseg * a.split > INT_MAX) return hipErrorInvalidValue;
    const size_t lds = xseg_lds_bytes((int)sizeof(T), a.max_chunks, a.max_granules);
    const int grid = a.nseg * a.split;
    switch (a.fuse) {
    case XFUSE_NONE: k_xseg<T, xseg_npt<T>(), XFUSE_NONE><<<grid, kWideThreads, lds, s>>>(a); break;
    case XFUSE_P:
        if (!a.st_in || !a.st_out || !a.R || !a.V || !a.xv_out) return hipErrorInvalidValue;
        k_xseg<T, xseg_npt<T>(), XFUSE_P><<<grid, kWideThreads, lds, s>>>(a);
        break;
    case XFUSE_S:
        if (!a.st_in || !a.st_out || !a.R || !a.V || !a.xv_out || !a.part_out) return hipErrorInvalidValue;
        k_xseg<T, xseg_npt<T>(), XFUSE_S><<<grid, kWideThreads, lds, s>>>(a);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_xseg(const XsegArgs<double> &a, hipStream_t s) { return xseg_dispatch<double>(a, s); }
hipError_t launch_xseg(const XsegArgs<float> &a, hipStream_t s) { return xseg_dispatch<float>(a, s); }

hipError_t launch_vec_P(const VecArgs<double> &a, hipStream_t s) { return k_vec_P_dispatch<double>(a, s); }
hipError_t launch_vec_S(const VecArgs<double> &a, hipStream_t s) { return k_vec_S_dispatch<double>(a, s); }
hipError_t launch_vec_XR(const VecArgs<double> &a, hipStream_t s) { return k_vec_XR_dispatch<double>(a, s); }
hipError_t launch_vec_P(const VecArgs<float> &a, hipStream_t s) { return k_vec_P_dispatch<float>(a, s); }
hipError_t launch_vec_S(const VecArgs<float> &a, hipStream_t s) { return k_vec_S_dispatch<float>(a, s); }
hipError_t launch_vec_XR(const VecArgs<float> &a, hipStream_t s) { return k_vec_XR_dispatch<float>(a, s); }

hipError_t launch_apply_xr(SolverState *st, const SysRange *sys, int nsys, const double *part, int nvb,
                           SolverState *mirror, int stamp, hipStream_t s)
{
    if (nsys <= 0) return hipSuccess;
    k_apply_xr<<<nsys, kThreads, 0, s>>>(st, sys, part, nvb, mirror, stamp);
    return hipGetLastError();
}

hipError_t launch_finalize(SolverState *st, const SysRange *sys, int nsys, const double *part, int nparts,
                           int in_spmv, int fin, hipStream_t s)
{
    if (nsys <= 0) return hipSuccess;
    k_finalize<<<nsys, kThreads, 0, s>>>(st, sys, part, nparts, in_spmv, fin);
    return hipGetLastError();
}

hipError_t launch_init_vectors(const VecArgs<double> &a, hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_init_vectors<<<a.nvb, kThreads, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_ir_start(const VecArgs<float> &a, const double *r64, const SolverState *outer, hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_ir_start<<<a.nvb, kThreads, 0, s>>>(a, r64, outer);
    return hipGetLastError();
}

hipError_t launch_ir_update(const VecArgs<float> &a, const double *x_in, double *x_out, hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_ir_update<<<a.nvb, kThreads, 0, s>>>(a, x_in, x_out);
    return hipGetLastError();
}

hipError_t launch_mirror(const SolverState *a, const SolverState *b, int n, SolverState *mirror, int stamp,
                         hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_mirror<<<1, kThreads, 0, s>>>(a, b, n, mirror, stamp);
    return hipGetLastError();
}

hipError_t launch_dot(const VecArgs<double> &a, const double *x, const double *y, hipStream_t s)
{
    if (a.nvb <= 0) return hipSuccess;
    k_dot<<<a.nvb, kThreads, 0, s>>>(a, x, y);
    return hipGetLastError();
}

hipError_t launch_solve_start(const double *x, double *X, int n, SolverState *st, int nstates, double tol,
                              int maxit, int fixed, hipStream_t s)
{
    const int grid = x ? std::max(1, std::min(2048, (n + kThreads * 4 - 1) / (kThreads * 4))) : 1;
    k_solve_start<<<grid, kThreads, 0, s>>>(x, X, n, st, nstates, tol, maxit, fixed);
    return hipGetLastError();
}

hipError_t launch_state_init(SolverState *st, int nstates, double tol, int maxit, int fixed, hipStream_t s)
{
    k_state_init<<<1, kThreads, 0, s>>>(st, nstates, tol, maxit, fixed);
    return hipGetLastError();
}

}  // namespace hbs
