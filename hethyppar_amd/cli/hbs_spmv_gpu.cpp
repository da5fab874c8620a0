// hbs_spmv_gpu -- command-line SpMV benchmark on one MI355X with the options
// and stdout of the reference's spmv_gpu (src/entry/spmv_gpu.c: options
// :35-55, run and output :57-152): one warm-up SpMV, -n timed repetitions of
// Y = A*X, the maximum error of Y against [1, 2, ..., m] (the reference's
// inputs are built so that A*X is that vector), Y written with "%lf".
//
//   hbs_spmv_gpu -m A.mtx -x X.txt -o Y.txt -n REPS
#include <argp.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hbs.h"

namespace {

struct Args {
    const char *input_matrix = nullptr;
    const char *input_x = nullptr;
    const char *output_y = nullptr;
    int n_iters = 1;
};

argp_option options[] = {
    {"input-matrix", 'm', "FILE", 0, "Input matrix market file path", 0},
    {"input-x", 'x', "FILE", 0, "Input X vector file path", 0},
    {"output-y", 'o', "FILE", 0, "Output Y vector file path", 0},
    {"n-iters", 'n', "POSITIVE-INTEGER", 0, "Number of SpMV repetitions", 0},
    {nullptr, 0, nullptr, 0, nullptr, 0}};

error_t parse_opt(int key, char *arg, argp_state *state)
{
    Args *a = static_cast<Args *>(state->input);
    switch (key) {
    case 'm': a->input_matrix = arg; break;
    case 'x': a->input_x = arg; break;
    case 'o': a->output_y = arg; break;
    case 'n': a->n_iters = (int)strtol(arg, nullptr, 10); break;
    case ARGP_KEY_ARG: return 0;
    default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

char doc[] = "SpMV-only benchmark on one MI355X (libhbs) -- repeatedly computes Y = A*X and times it";
argp parser = {options, parse_opt, "", doc, nullptr, nullptr, nullptr};

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int die(const char *what)
{
    fprintf(stderr, "%s: %s\n", what, hbs_last_error());
    return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char **argv)
{
    Args args;
    argp_parse(&parser, argc, argv, 0, nullptr, &args);
    if (!args.input_matrix || !args.input_x || !args.output_y) {
        fprintf(stderr, "Error: -m, -x, and -o must all be specified.\n");
        return EXIT_FAILURE;
    }
    for (const char *f : {args.input_matrix, args.input_x}) {
        if (access(f, F_OK) == -1) {
            fprintf(stderr, "Input file '%s' does not exist\n", f);
            return EXIT_FAILURE;
        }
    }
    const int niters = args.n_iters < 1 ? 1 : args.n_iters;
    const double t_begin = now();
    hbs_ctx ctx = nullptr;
    if (hbs_ctx_create(0, nullptr, &ctx) != HBS_OK) return die("hbs_ctx_create");
    const double t_read0 = now();
    int m = 0, n = 0, nnz = 0;
    int *rp = nullptr, *col = nullptr;
    double *val = nullptr;
    if (hbs_mm_read(args.input_matrix, &m, &n, &nnz, &rp, &col, &val) != HBS_OK) return die("hbs_mm_read");
    hbs_csr A = nullptr;
    if (hbs_csr_create(ctx, m, n, nnz, rp, col, val, 0, &A) != HBS_OK) return die("hbs_csr_create");
    hbs_host_free(rp);
    hbs_host_free(col);
    hbs_host_free(val);
    std::vector<double> hx((size_t)n), hy((size_t)m);
    if (hbs_vector_read(args.input_x, n, hx.data()) != HBS_OK) return die("hbs_vector_read");
    double *dx = nullptr, *dy = nullptr;
    if (hbs_dev_alloc(ctx, sizeof(double) * (size_t)n, (void **)&dx) != HBS_OK ||
        hbs_dev_alloc(ctx, sizeof(double) * (size_t)m, (void **)&dy) != HBS_OK ||
        hbs_h2d(ctx, dx, hx.data(), sizeof(double) * (size_t)n) != HBS_OK ||
        hbs_dev_zero(ctx, dy, sizeof(double) * (size_t)m) != HBS_OK)
        return die("device vectors");
    const double t_read1 = now();
    printf("Matrix name : %s\n", args.input_matrix);
    if (hbs_csr_spmv_f64(A, 1.0, dx, 0.0, dy) != HBS_OK || hbs_ctx_sync(ctx) != HBS_OK) return die("warm-up");
    const double t0 = now();
    for (int i = 0; i < niters; ++i)
        if (hbs_csr_spmv_f64(A, 1.0, dx, 0.0, dy) != HBS_OK) return die("spmv");
    if (hbs_ctx_sync(ctx) != HBS_OK) return die("sync");
    const double t1 = now();
    if (hbs_d2h(ctx, hy.data(), dy, sizeof(double) * (size_t)m) != HBS_OK) return die("hbs_d2h");
    double max_abs_err = 0.0;
    for (int i = 0; i < m; ++i) max_abs_err = std::fmax(max_abs_err, std::fabs(hy[(size_t)i] - (double)(i + 1)));
    const double total = t1 - t0;
    printf("n_iters : %d \n"
           "spmv : %lf \n"
           "spmv_per_iter : %e \n"
           "file_read : %lf \n"
           "max_abs_err_vs_1ton : %e \n",
           niters, total, total / niters, t_read1 - t_read0, max_abs_err);
    printf("everything_total : %lf\n", now() - t_begin);
    printf("\n----------------------------------------------------------------------\n");
    if (hbs_vector_write(args.output_y, m, hy.data()) != HBS_OK) return die("hbs_vector_write");
    hbs_csr_destroy(A);
    hbs_dev_free(ctx, dx);
    hbs_dev_free(ctx, dy);
    hbs_ctx_destroy(ctx);
    return EXIT_SUCCESS;
}
