// hbs_patpart -- row partition of a Matrix Market matrix for the multi-rank
// drivers, with the arguments of the `patpart` program the reference's
// tools/python/matrix_partition.py runs (cmd_patpart) and the output of
// CalcPartVec (hhp_matrix.c:190-300): the partition vector, one part id per
// line (the `.part` files the drivers read with MPI_ivector_read_bcast),
// the entries per part and the cut.  The partition comes from
// hbs_partition, this library's own partitioner, in place of PaToH.
//
//   hbs_patpart MATRIX.mtx NPARTS WEIGHTS.txt IMBAL_PERCENT SEED OUT.part [LOG]
//
// WEIGHTS.txt holds NPARTS relative part weights (whitespace separated, the
// target weight file of CalcPartVec).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hbs.h"

namespace {

int die(const char *what)
{
    fprintf(stderr, "%s: %s\n", what, hbs_last_error());
    return EXIT_FAILURE;
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 7 || argc > 8) {
        fprintf(stderr, "usage: %s MATRIX.mtx NPARTS WEIGHTS.txt IMBAL_PERCENT SEED OUT.part [LOG]\n", argv[0]);
        return EXIT_FAILURE;
    }
    const char *mtx = argv[1], *wfile = argv[3], *out = argv[6], *logf = argc == 8 ? argv[7] : nullptr;
    const int nparts = atoi(argv[2]);
    const double imbal = atof(argv[4]) / 100.0;
    const int seed = atoi(argv[5]);
    if (nparts < 1) {
        fprintf(stderr, "NPARTS must be positive\n");
        return EXIT_FAILURE;
    }
    std::vector<double> target;
    {
        FILE *f = fopen(wfile, "r");
        if (!f) {
            fprintf(stderr, "file can't be opened %s file \n", wfile);
            return EXIT_FAILURE;
        }
        double w;
        while ((int)target.size() < nparts && fscanf(f, "%lf", &w) == 1) target.push_back(w);
        fclose(f);
        if ((int)target.size() != nparts) {
            fprintf(stderr, "%s holds %zu weights, %d parts asked\n", wfile, target.size(), nparts);
            return EXIT_FAILURE;
        }
    }
    const double t0 = now();
    int m = 0, n = 0, nnz = 0, *rowptr = nullptr, *col = nullptr;
    double *val = nullptr;
    if (hbs_mm_read(mtx, &m, &n, &nnz, &rowptr, &col, &val) != 0) return die("matrix read");
    if (m != n) {
        fprintf(stderr, "the row partition needs a square matrix (%d x %d)\n", m, n);
        return EXIT_FAILURE;
    }
    const double t1 = now();
    std::vector<int> partvec((size_t)n);
    long long cut = 0;
    if (hbs_partition(n, rowptr, col, nparts, target.data(), imbal, seed, partvec.data(), &cut) != 0)
        return die("partition");
    const double t2 = now();
    FILE *f = fopen(out, "w");
    if (!f) {
        fprintf(stderr, "file can't be opened %s file \n", out);
        return EXIT_FAILURE;
    }
    for (int r = 0; r < n; ++r) fprintf(f, "%d\n", partvec[(size_t)r]);
    fclose(f);
    // entries per part, as nnzDistribution prints them (hhp_matrix.c:190-222)
    std::vector<long long> cnt((size_t)nparts, 0);
    for (int r = 0; r < n; ++r) cnt[(size_t)partvec[(size_t)r]] += rowptr[r + 1] - rowptr[r];
    FILE *lg = logf ? fopen(logf, "w") : nullptr;
    for (int p = 0; p < nparts; ++p) {
        printf("[%d] %lld ", p, cnt[(size_t)p]);
        if (lg) fprintf(lg, " [%d] : %lld ", p, cnt[(size_t)p]);
    }
    printf("\ncut : %lld\n", cut);
    printf("file_read : %lf\n", t1 - t0);
    printf("partition : %lf\n", t2 - t1);
    if (lg) {
        fprintf(lg, "\n\ncut  : %lld \n", cut);
        fclose(lg);
    }
    hbs_host_free(rowptr);
    hbs_host_free(col);
    hbs_host_free(val);
    return EXIT_SUCCESS;
}
