// hbs_internal.h -- shared definitions of the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "hbs.h"

namespace hbs {

constexpr int kThreads = 256;          // workgroup size of every kernel (4 wave64)
constexpr int kWaves = kThreads / 64;
constexpr int kWideThreads = 512;      // SpMV workgroup size for row blocks above 2048 entries
constexpr int kMaxBlockNnz = 16384;    // upper bound of nnz per SpMV row block (128 KiB of FP64 LDS)
constexpr int kLongRow = 256;          // rows longer than this are summed in chunks (k_chunks)
constexpr int kChunkLen = 1024;        // entries per chunk of such a row
constexpr int kXcds = 8;                // MI355X: 8 XCDs, workgroups dispatched round robin
constexpr int kRowPad = 8;             // entries of padding past nnz in col / val / val32
constexpr int kSegCols = 8192;         // x-segmented SpMV: columns per segment (64 KiB of FP64 x in LDS)
constexpr int kSegBlockNnz = 8192;     // x-segmented SpMV: entries per row block

// Threads of the SpMV workgroups that sum row blocks of block_nnz entries.
constexpr int spmv_threads(int block_nnz) { return block_nnz > 2048 ? kWideThreads : kThreads; }
constexpr int kMaxDots = 2;            // reductions fused into one pass
constexpr int kMaxGroups = 4096;       // SpMV workgroups by default = partials a consumer sums

// One SpMV row block: rows [r0, r1) with entries [j0, j1).
struct alignas(16) RowBlock {
    int r0, r1, j0, j1;
};
// kind 0: the block's entries fit in LDS; lanes = lanes per row in the
// reduction (1 = the order-exact serial sum).  kind 1: one long row r0 whose
// value is the sum, in order, of the chunk sums [first, first + nchunks).
// kind 2: as kind 0, but the products are formed in column order (the
// column-sorted copy scol/sval of the block's entries) and stored at their
// CSR positions (sperm, or the high bits of scol), so a wave's x gathers fall
// on fewer cache lines.
struct alignas(16) BlockAux {
    int kind, first, nchunks, lanes;
};
// A piece [j0, j1) of a long row, summed by one workgroup of k_chunks.
struct alignas(16) Chunk {
    int j0, j1, row, sys;
};
// SpMV workgroup p handles row blocks [rb0, rb1), all of system sys.
struct alignas(16) Group {
    int rb0, rb1, sys, pad;
};
// Vector workgroup: elements [e0, e1) of system sys; its 16-byte packs start
// at base (e0 rounded down to a multiple of 4 elements).
struct alignas(16) VecBlock {
    int sys, e0, e1, base;
};
// Per system: SpMV workgroups [g0, g1) and vector workgroups [v0, v1).
struct alignas(16) SysRange {
    int g0, g1, v0, v1;
};

// Device-resident BiCGStab scalars of one system.  A solve keeps two copies
// (slots A and B): every kernel that updates scalars reads one slot and its
// first workgroup of the system writes the other, so no workgroup ever reads
// a value another workgroup of the same kernel is writing.
struct alignas(16) SolverState {
    double rho, alpha, omega, beta;
    double rho_next;     // (r, r0) of the newest r
    double rv, ts, tt;   // (r0, v), (t, s), (t, t)
    double ss, rr;       // ||s||^2, ||r||^2
    double thr;          // tol * ||r0||
    double rr0;          // ||r0||^2
    double tol;
    double scale;        // mixed: ||r|| of the outer residual the sweep solves for
    double resid2;       // ||b - Ax||^2 of the last residual pass
    double out;          // generic reduction result
    int iter, maxit, done, stop;
    int s_conv, fixed, pending, pad;   // pending: the (r,r0), ||r||^2 partials await FIN_XR
};

enum Epi : int {
    EPI_AXPBY = 0,   // y = alpha*Ax + beta*y
    EPI_DOT1 = 1,    // y = Ax ; d0 = (w, y)
    EPI_DOT2 = 2,    // y = Ax ; d0 = (y, w), d1 = (y, y)
    EPI_RESID = 3,   // y = w - Ax ; d0 = (y, y)
    EPI_RESID2 = 4,  // y = w - Ax ; d0 = d1 = (y, y)   (start of a solve)
    EPI_RESIDB = 5,  // y = w - Ax ; d0 = (y, y), d1 = (w, w)   (end of a solve)
};

enum Fin : int {
    FIN_NONE = 0,
    FIN_ALPHA = 1,   // rv = d0, alpha = rho / rv
    FIN_OMEGA = 2,   // ts = d0, tt = d1, omega = ts / tt
    FIN_XR = 3,      // rho_next = d0, rr = d1: iteration count, stop tests, next beta
    FIN_S = 4,       // ss = d0: ||s|| stop test
    FIN_RESID = 5,   // resid2 = d0
    FIN_STORE = 6,   // out = d0
    FIN_RESIDB = 7,  // resid2 = d0, out = d1
};

struct SpmvPlan {
    int nblk = 0;
    int ngroups = 0;
    int nchunks = 0;
    int nsys = 1;
    int block_nnz = 0;
    int npt = 0;                  // SpMV entries per thread per pass
    int serial_rows = 0;
    int max_row = 0;
    RowBlock *blocks = nullptr;   // device [nblk]
    BlockAux *aux = nullptr;      // device [nblk]
    Chunk *chunks = nullptr;      // device [nchunks]
    Group *groups = nullptr;      // device [ngroups]
};

// Launch parameters of the SpMV kernels.
template <typename T>
struct SpmvArgs {
    const RowBlock *blocks;
    const BlockAux *aux;
    const Group *groups;
    int ngroups;
    const SysRange *sys;
    const int *rowptr;
    const int *col;
    const T *val;
    const T *x;
    T *y;
    const T *w;
    T alpha, beta;
    const int *scol;            // kind-2 blocks: entries in column order,
    const T *sval;              //   with their positions in the block:
    const unsigned short *sperm;    //   sperm[j], or (pos_shift > 0)
    int pos_shift;              //   scol[j] >> pos_shift over column bits
    const double *chunk_sums;   // sums of the long-row chunks (k_chunks, same x)
    double *part_out;           // [ngroups * kMaxDots] partials of the fused dots
    const double *part_in;      // partials of the previous vector kernel (FIN_S)
    int nparts_in;
    const SolverState *st_in;   // null: no state
    SolverState *st_out;
    int fin_in;                 // FIN_NONE or FIN_S
    int check_done;             // return when the system is done (or ||s|| fired)
    int stream_nt;              // load col/val non-temporally
    SolverState *st_mark;       // start of a solve: set pending in this slot
    const T *prod;              // x-segmented matrices: the products of k_xseg, row-block major
    const unsigned short *ppos; //   and each one's position in its row block (CSR order)
    const unsigned short *rperm;    // x-segmented matrices: serial row-sum order (plan_row_order)
    const int *bcol;            // kind-3 (3x3-block) blocks: each block's first column
    const T *bval;              //   and the values block-major (b3_layout)
    int b3var;                  //   some kind-3 blocks hold block rows of different lengths
    int dmode;                  // x-segmented row sums add a diagonal term: 0 none, 1 x[r], 2 diag[r] * x[r]
    const T *diag;
};

// x-segmented SpMV, product pass (k_xseg): workgroup s stages the x segment
// [s * seg_cols, (s + 1) * seg_cols) in LDS and forms the products of the
// entries whose column lies in it.  Those entries are stored segment-major
// (xcol: column within the segment; xval) in chunks, one per (segment, row
// block); product g goes to prod[g + off[c]], c the chunk of entry g, found
// as cbase[s] + gch[granule of g] + delta[g] (granules of 64 entries of a
// segment).  The segment's off and gch entries are staged in LDS too: at
// most max_chunks and max_granules of them.
template <typename T>
struct XsegArgs {
    int nseg, seg_cols, n;
    int split;                   // workgroups per segment, each a contiguous share of its entries
    int max_chunks, max_granules;
    const int *sbase;            // [nseg + 1] first entry of each segment
    const int *cbase;            // [nseg + 1] first chunk of each segment
    const int *gbase;            // [nseg + 1] first granule of each segment
    const int *gch;              // [granules] chunk of each granule's first entry, within its segment
    const int *off;              // [chunks] row-block-major start minus segment-major start
    const unsigned short *xcol;
    const unsigned char *delta;
    const T *xval;
    const T *x;
    T *prod;
    const SolverState *st;       // null: always run; else skip when st->done
    // Solver iterations: the vector the products read, updated by the pass
    // itself (one system).  XFUSE_P: x = P = ((P - V omega) beta) + R after
    // FIN_XR of the previous iteration's (r, r0), ||r||^2 partials;
    // XFUSE_S: x = S = R - V alpha after FIN_ALPHA of the (r0, v) partials,
    // with a ||S||^2 partial per workgroup.  Each segment's first share
    // writes its elements of the vector (xv_out, may be x) and its partial
    // (the others 0); workgroup 0 writes the state it derived to st_out.
    int fuse;
    T *xv_out;
    const T *R, *V;
    const SolverState *st_in;
    SolverState *st_out;
    const double *part_in;
    int nparts_in, pin0, pin1;   // partials [pin0, pin1), nparts_in apart per dot
    double *part_out;            // XFUSE_S: [nseg * split]
};
enum XsegFuse : int { XFUSE_NONE = 0, XFUSE_P = 1, XFUSE_S = 2 };

template <typename T>
struct ChunkArgs {
    const Chunk *chunks;
    int nchunks;
    const int *col;
    const T *val;
    const T *x;
    double *chunk_sums;
    const SolverState *st;      // null: always run
    const T *prod;              // x-segmented matrices: sum these products instead
};

// Launch parameters of the vector kernels of one BiCGStab iteration.
template <typename T>
struct VecArgs {
    const VecBlock *blocks;
    const SysRange *sys;
    int nvb;
    int upb;                    // 16-byte packs per thread (1, 2, 4, 8)
    T *X, *R, *P, *V, *S, *T_;
    const T *R0;
    const SolverState *st_in;
    SolverState *st_out;
    const double *part_in;      // partials the prologue applies
    int nparts_in;
    int in_spmv;                // 1: part_in from an SpMV (ranges g0..g1), 0: from a vector kernel
    double *part_out;           // [nvb * kMaxDots]
};

// ---- k right-hand sides (hbs_block.hip) ----
// A group of kc <= 64 columns of row-major n x ld blocks (pointers already
// offset to the group's first column).  A wave holds 64 / kp rows x kp
// columns (kp = kc rounded up to a power of two): lane l works on column
// l % kp of row offset l / kp, so one row's columns are contiguous in HBM.
constexpr int kBlockCols = 64;
// Row traversal of the multi-column SpMM kernels: runs of `len` consecutive
// rows (len even), run k starting at row run[k], the runs covering every
// row once; each XCD takes an equal contiguous share of the runs.  run ==
// nullptr: the rows in order, each XCD an equal contiguous share of them.
// ctr: step counters, zero between launches, for ctr_tiles column tiles:
// per tile and XCD a set of kCtrSet counters, kCtrStride ints apart -- the
// kCtrPerXcd step counters, then the arrival count.  The waves of a tile
// with counters take their steps from them, the others in a fixed stride.
struct RowRuns {
    const int *run = nullptr;
    int len = 0, nruns = 0;
    int *ctr = nullptr;
    int ctr_tiles = 0;
    int ctr_per_xcd = 0;   // step counters in use per XCD (at most kCtrPerXcd)
};
constexpr int kSpmmCtrTiles = 16;   // column tiles with step counters per matrix
constexpr int kCtrPerXcd = 16;      // step counters per XCD (steps dealt round robin over them)
constexpr int kCtrSet = kCtrPerXcd + 1;
constexpr int kCtrStride = 64;      // ints between two counters (256 bytes)

struct BlockArgs {
    int n, ld, kc, kp;
    int ngroups;                  // workgroups (a multiple of 8) = partials per column
    const int *rowptr, *col;
    const double *val;
    const double *x;              // SpMM input
    double *y;                    // SpMM output
    const double *w;              // SpMM epilogue operand
    double *X, *R, *P, *V, *S, *T;
    const double *R0;
    const SolverState *st;        // [kc], read-only in the streaming kernels
    int skip_done;                // 1: columns whose state is done are skipped
    double *part;                 // [kMaxDots][kBlockCols][ngroups]
    RowRuns rr;                   // SpMM row traversal
    int rb;                       // SpMM rows per lane step (1, 2, 4)
    int v16;                      // 64-column SpMM with 16-byte X gathers (k_bspmm2; rb unused)
};
enum BlockEpi : int {
    BEPI_DOT1 = 0,    // y = A x ; d0 = (w, y)
    BEPI_DOT2 = 1,    // y = A x ; d0 = (y, w), d1 = (y, y)   (skips columns with s_conv)
    BEPI_RESID = 2,   // y = w - A x ; d0 = d1 = (y, y)
};
enum BlockVec : int {
    BVEC_P = 0,       // P = ((P - V omega) beta) + R
    BVEC_S = 1,       // S = R - V alpha ; d0 = (S, S)
    BVEC_XR = 2,      // X = (X + P alpha) + S omega ; R = S - T omega ; d0 = (R, R0), d1 = (R, R)
    BVEC_INIT = 3,    // R0 = R ; P = V = 0
    BVEC_DOT = 4,     // d0 = (x, x) of a.x
};
hipError_t launch_block_spmm(const BlockArgs &a, int epi, hipStream_t s);
hipError_t launch_block_vec(const BlockArgs &a, int kind, hipStream_t s);
// Reduce the partials of every column into st[c] with fin (one workgroup per column).
hipError_t launch_block_fin(SolverState *st, int kc, const double *part, int ngroups, int fin, hipStream_t s);

// ---- row partitioner (hbs_part.cpp) ----
// partvec[n] for a square CSR: nparts parts with weights proportional to
// target[nparts], within 1 + final_imbal; see hbs_partition.
int partition_rows(int n, const int *rowptr, const int *col, int nparts, const double *target, double final_imbal,
                   int seed, int *partvec);
// Connectivity - 1 of partvec over the rows of the CSR.
long long partition_cut(int n, const int *rowptr, const int *col, const int *partvec);

// ---- host-side planning (hbs_plan.cpp) ----
struct HostPlan {
    std::vector<RowBlock> blocks;
    std::vector<BlockAux> aux;
    std::vector<Chunk> chunks;
    std::vector<Group> groups;
    int serial_rows = 0;
    int max_row = 0;
};
// Order of the serial row sums of each block of 512-thread workgroups:
// rperm[r0 + k] = the block's k-th longest row (local index, ties in row
// order), so the rows one wave sums together have similar lengths and no
// lane waits for the longest row of a mixed set.  A block takes that order
// (aux.first = 1) when its estimated cost -- per wave, the sum over its row
// slots of the batches of 8 products the longest row of the slot needs;
// the block's cost is its slowest wave's -- falls below 80% of the row
// order's.  Returns the number of such blocks; rperm covers every row
// (identity elsewhere).
int plan_row_order(const int *rowptr, HostPlan &plan, int m, std::vector<unsigned short> &rperm);
// 3x3-block matrices: every block row (rows 3R..3R+2) has three rows of the
// same entries' columns, in aligned triples 3J, 3J+1, 3J+2, at most kLongRow
// per row.  Their row blocks (kind 3, BlockAux{3, 0, K, 1}: whole block rows
// of K blocks each, at most kWideThreads blocks and block_nnz entries; K = 0
// when the block rows' lengths differ, which the kernel then reads from the
// row pointers) are summed by k_spmv3, one block per thread.  plan_rows_b3
// returns the number of row blocks with K = 0.  A row block's nbk blocks are
// numbered in CSR order (block row, then block); its values are stored in 9
// planes of nbk, plane 3a + c holding entry (a, c) of every block, and each
// block's first column once (bcol, at the row block's first entry / 9 + its
// number).  b3_layout gives the CSR entry of each stored value (bsrc, nnz
// entries) and bcol (nnz / 9 entries); the device copies carry
// kWideThreads entries of padding (zero values, column 0) past the end.
bool detect_b3(int m, int n, const int *rowptr, const int *col);
int plan_rows_b3(int row0, int m, const int *rowptr, int block_nnz, int sys, HostPlan &plan);
void b3_layout(const HostPlan &plan, const int *rowptr, const int *col, std::vector<int> &bsrc,
               std::vector<int> &bcol);
// Row blocks of rows [row0, row0 + m) of system sys, appended to the plan.
void plan_rows(int row0, int m, const int *rowptr, int block_nnz, int serial_max_avg, int sys,
               HostPlan &plan);
// Marks kind-0 blocks whose gathers, taken in column order, touch at most
// max_line_pct percent as many 128-byte lines of an FP64 x per wave
// instruction (64 consecutive entries) as in CSR order (kind 2), and fills
// src[j] (the CSR entry at sorted position j) and pos[j] (its position in
// the block); identity outside kind-2 blocks.  Returns the number of kind-2
// blocks.
int plan_sorted(const int *col, int nnz, int max_line_pct, HostPlan &plan, std::vector<int> &src,
                std::vector<unsigned short> &pos);
// x-segmented layout of the row blocks of a plan: every block's entries
// ordered by column segment (kSegCols columns), CSR order inside a segment
// (row-block major, ppos = CSR position in the block), and the same entries
// segment-major in chunks of one (segment, row block); see XsegArgs.
struct XsegHost {
    int nseg = 0;
    int max_chunks = 0, max_granules = 0;   // per segment
    std::vector<int> sbase, cbase, gbase, gch, off;
    std::vector<int> src;                  // segment-major entry -> CSR entry
    std::vector<unsigned short> xcol;      // segment-major: column within the segment
    std::vector<unsigned char> delta;      // segment-major: chunk minus its granule's first chunk
    std::vector<unsigned short> ppos;      // row-block major: CSR position in the block
};
void plan_xseg(const int *col, int n, const HostPlan &plan, XsegHost &xs);
// Distinct 128-byte lines of an FP64 x touched per entry, over windows of
// window consecutive entries (sampled): near 1 for scattered columns.
double column_scatter(const int *col, int nnz, int window);
// Split the row blocks [rb0, rb1) of one system over ngroups workgroups.
void plan_groups(int rb0, int rb1, int ngroups, int sys, HostPlan &plan);
// Vector blocks of elements [e0, e0 + n) of one system.
void plan_vec(int e0, int n, int elems_per_block, int sys, std::vector<VecBlock> &blocks);
// The same blocks into out[cap] (out may be null to count); -1 if more than cap.
int plan_vec_into(int e0, int n, int elems_per_block, int sys, VecBlock *out, int cap);
// Runs of the multi-column SpMM for a square matrix that is a 7-point
// operator on a 3-D grid too large for an XCD's L2 (y-slab order): the run
// starts and len = nx; returns the number of runs, 0 (runs empty) when the
// matrix is not such a grid.
int plan_grid_runs(int m, const int *rowptr, const int *col, std::vector<int> &runs, int &len);
int vec_elems_per_block(int n);
extern int g_vec_target;      // vector workgroups a plan aims at (hbs_set_vec_blocks)
extern int g_grid_slab;       // y-lines per slab of plan_grid_runs (0: from the L2 size)
extern int g_spmm_groups;     // workgroups of the multi-column SpMM (0: 2048)
extern int g_spmm_v16;        // 64-column SpMM: 16-byte gathers where aligned (1) or 8/4-byte (0)
extern int g_spmm_counters;   // step counters per XCD of the dealt 64-column sweeps
extern int g_xseg_fuse;       // solvers created afterwards: P and S updates in x-segmented product passes
int vec_packs_per_thread(int elems_per_block, int elem_bytes);

// ---- launchers (hbs_kernels.hip) ----
// Raise the dynamic LDS limit of the SpMV kernels for row blocks of block_nnz
// entries, npt entries per thread per pass (once per plan, outside any
// stream capture).
hipError_t prepare_spmv(int block_nnz, int npt);
// FP64 SpMV workgroups of row blocks of block_nnz entries (npt per thread per
// pass) resident at once on the device (0 when the query fails).
int spmv_resident_groups(int block_nnz, int npt, int device, bool pre = false);
hipError_t launch_spmv(const SpmvArgs<double> &a, int epi, int block_nnz, int npt, hipStream_t s);
hipError_t launch_spmv(const SpmvArgs<float> &a, int epi, int block_nnz, int npt, hipStream_t s);
hipError_t launch_chunks(const ChunkArgs<double> &a, int block_nnz, hipStream_t s);
hipError_t launch_chunks(const ChunkArgs<float> &a, int block_nnz, hipStream_t s);
hipError_t launch_xseg(const XsegArgs<double> &a, hipStream_t s);
hipError_t launch_xseg(const XsegArgs<float> &a, hipStream_t s);
// Dynamic LDS of k_xseg for values of elem_bytes and the given table sizes.
size_t xseg_lds_bytes(int elem_bytes, int max_chunks, int max_granules);
// Allow k_xseg that much LDS (once per plan, outside any stream capture).
hipError_t prepare_xseg(int max_chunks, int max_granules);
// Entries per thread per pass of the x-segmented row sums (k_spmv<PRE>): one
// pass covers a row block of kSegBlockNnz entries.
constexpr int kPreNpt = kSegBlockNnz / kWideThreads;
// 3x3-block SpMV (k_spmv3): one block of 9 entries per thread, one pass per
// row block of at most kB3BlockNnz entries, its products in LDS.
constexpr int kB3Npt = 9;
constexpr int kB3BlockNnz = kB3Npt * kWideThreads;
constexpr size_t spmv3_lds_bytes(size_t elem_bytes) { return (size_t)kB3BlockNnz * elem_bytes; }
// k_xseg entries per thread per pass: as many as the registers allow with
// two workgroups per CU, so each workgroup keeps more bytes in flight
template <typename T>
constexpr int xseg_npt() { return sizeof(T) == 8 ? 12 : 16; }
constexpr int kXsegMaxLds = 96 * 1024; // k_xseg LDS cap (x segment + chunk tables); above it no x-segmenting
hipError_t launch_vec_P(const VecArgs<double> &a, hipStream_t s);
hipError_t launch_vec_S(const VecArgs<double> &a, hipStream_t s);
hipError_t launch_vec_XR(const VecArgs<double> &a, hipStream_t s);
hipError_t launch_vec_P(const VecArgs<float> &a, hipStream_t s);
hipError_t launch_vec_S(const VecArgs<float> &a, hipStream_t s);
hipError_t launch_vec_XR(const VecArgs<float> &a, hipStream_t s);
// Apply the pending FIN_XR of every system in place (one workgroup per
// system) and copy the states to mirror (device view of host memory; may be
// null); a nonzero stamp is then written into each mirrored state's pad,
// after a system-scope fence.
hipError_t launch_apply_xr(SolverState *st, const SysRange *sys, int nsys, const double *part, int nvb,
                           SolverState *mirror, int stamp, hipStream_t s);
// Reduce partials into st[sys] with fin (FIN_RESID / FIN_STORE), one workgroup per system.
hipError_t launch_finalize(SolverState *st, const SysRange *sys, int nsys, const double *part,
                           int nparts, int in_spmv, int fin, hipStream_t s);
// R0 = R ; P = V = 0 (FP64, after the residual SpMV wrote R)
hipError_t launch_init_vectors(const VecArgs<double> &a, hipStream_t s);
// Mixed sweep start: R = R0 = (float)(r64 / ||r64||), X = P = V = 0, partials
// d0 = d1 = ||R||^2 in part_out; st_out (the inner slot A) gets pending and
// scale = ||r64|| from outer.resid2.
hipError_t launch_ir_start(const VecArgs<float> &a, const double *r64, const SolverState *outer,
                           hipStream_t s);
// Mixed sweep end: x64 += scale * (double)x32 (scale from st_in).
// x_out = x_in + X32 * scale of the sweep (x_out may be x_in).
hipError_t launch_ir_update(const VecArgs<float> &a, const double *x_in, double *x_out, hipStream_t s);
// Copies n states of a (then n of b, if b is not null) into the host-mapped
// mirror, then (stamp != 0) stamps every copy after a system-scope fence.
hipError_t launch_mirror(const SolverState *a, const SolverState *b, int n, SolverState *mirror, int stamp,
                         hipStream_t s);
// d0 = (x, y) partials into part_out.
hipError_t launch_dot(const VecArgs<double> &a, const double *x, const double *y, hipStream_t s);
hipError_t launch_state_init(SolverState *st, int nstates, double tol, int maxit, int fixed, hipStream_t s);
// X = x and the initial states of nstates slots, in one launch.
hipError_t launch_solve_start(const double *x, double *X, int n, SolverState *st, int nstates, double tol,
                              int maxit, int fixed, hipStream_t s);

// BLAS-1 helpers on raw device pointers.
hipError_t launch_axpy_f64(int n, double a, const double *x, double *y, hipStream_t s);
hipError_t launch_scal_f64(int n, double a, double *x, hipStream_t s);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_copy_f64(size_t n, const double *src, double *dst, hipStream_t s);
hipError_t launch_to_f32(int n, const double *src, float *dst, hipStream_t s);

// Halo kernels: out[k] = in[idx[k]] and out[idx[k]] = in[k].
hipError_t launch_gather_f64(int n, const double *in, const int *idx, double *out, hipStream_t s);
// out[i] = (float)in[idx[i]]
hipError_t launch_gather_to_f32(int n, const double *in, const int *idx, float *out, hipStream_t s);
hipError_t launch_scatter_f64(int n, const double *in, const int *idx, double *out, hipStream_t s);

// Adjoint gradient on the pattern: g[j] = -lambda[row(j)] * x[col[j]].
hipError_t launch_grad_pattern(int m, const int *rowptr, const int *col, const double *lambda,
                               const double *x, double *g, hipStream_t s);

// SpMM (row-major blocks, ld = k).
hipError_t launch_spmm_f64(int m, const int *rowptr, const int *col, const double *val, int k,
                           const double *X, double *Y, int tile, const RowRuns &rr, hipStream_t s);
hipError_t launch_spmm_f32(int m, const int *rowptr, const int *col, const float *val, int k,
                           const float *X, float *Y, int tile, const RowRuns &rr, hipStream_t s);

// Reference-shaped fused kernels (hhp_dp_kernels.h).
hipError_t launch_dp_vecop_P(double *P, const double *V, const double *R, const double *omega,
                             const double *beta, int n, hipStream_t s);
hipError_t launch_dp_vecop_S(double *S, const double *R, const double *V,
                             const double *neg_alpha, int n, hipStream_t s);
hipError_t launch_dp_vecop_XR(double *X, double *R, const double *P, const double *S,
                              const double *T, const double *alpha, const double *omega,
                              const double *neg_omega, int n, hipStream_t s);
// which: 0 beta, 1 alpha, 2 omega, 3 add
hipError_t launch_dp_scalar(int which, double *a, double *b, const double *c, const double *d,
                            const double *e, hipStream_t s);

}  // namespace hbs
