// hbs_plan.cpp -- row-block and vector-block planning for the HIP kernels.
#include <algorithm>

#include "hbs_internal.h"

namespace hbs {

namespace {
constexpr int kMaxRowsPerBlock = 1024;   // 4 rows per thread in the serial sum

int pow2_floor(int v)
{
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}
}  // namespace

// Greedy row blocking: a block takes consecutive rows while their entries fit
// in block_nnz (the LDS product buffer).  A row longer than block_nnz becomes
// a run of chunk blocks.  Blocks whose average row length is at most
// serial_max_avg use the order-exact serial row sum; others get
// pow2_floor(256 / rows) lanes per row (2..64).
void plan_rows(int row0, int m, const int *rowptr, int block_nnz, int serial_max_avg, int sys,
               std::vector<RowBlock> &blocks, std::vector<BlockAux> &aux, int *serial_rows,
               int *max_row)
{
    const int end = row0 + m;
    int r = row0;
    while (r < end) {
        const int len = rowptr[r + 1] - rowptr[r];
        if (max_row && len > *max_row) *max_row = len;
        if (len > block_nnz) {
            const int nch = (len + block_nnz - 1) / block_nnz;
            const int first = (int)blocks.size();
            for (int c = 0; c < nch; ++c) {
                const int j0 = rowptr[r] + c * block_nnz;
                const int j1 = std::min(j0 + block_nnz, rowptr[r + 1]);
                blocks.push_back(RowBlock{r, r + 1, j0, j1});
                aux.push_back(BlockAux{c, nch, first, 1 | (sys << 8)});
            }
            ++r;
            continue;
        }
        const int r0 = r;
        const int j0 = rowptr[r];
        while (r < end && rowptr[r + 1] - j0 <= block_nnz && r - r0 < kMaxRowsPerBlock) {
            ++r;
            if (r < end && max_row) {
                const int l = rowptr[r + 1] - rowptr[r];
                if (l > *max_row) *max_row = l;
            }
        }
        const int nr = r - r0;
        const int nz = rowptr[r] - j0;
        int lanes = 1;
        if (nr > 0 && nz > serial_max_avg * nr) {
            lanes = std::min(64, pow2_floor(kThreads / nr));
            if (lanes < 2) lanes = 1;
        }
        if (lanes == 1 && serial_rows) *serial_rows += nr;
        blocks.push_back(RowBlock{r0, r, j0, rowptr[r]});
        aux.push_back(BlockAux{-1, 0, 0, lanes | (sys << 8)});
    }
}

void plan_vec(int e0, int n, int elems_per_block, int sys, std::vector<VecBlock> &blocks)
{
    for (int e = e0; e < e0 + n; e += elems_per_block)
        blocks.push_back(VecBlock{sys, e, std::min(e + elems_per_block, e0 + n), 0});
}

}  // namespace hbs
