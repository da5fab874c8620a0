// hbs_bicgstab_gpu -- command-line BiCGStab on one MI355X with the options
// and stdout of the reference's bicgstab_gpu (src/entry/bicgstab_gpu.c:
// argp options :24-84, output lines :570-610), so its benchmark scripts can
// run this binary in its place.
//
//   hbs_bicgstab_gpu -m A.mtx -x X0.txt -y B.txt -o X.txt -n ITERS
//
// As in the reference, -n iterations are run with no stop test
// (bicgstab_gpu.c:431-560).  Extensions: -t TOL (> 0: stop at
// ||r|| < TOL*||r0||, at most -n iterations), --mixed (FP32 sweeps inside
// FP64 refinement), --check-every K.  The "spmv" line is the wall time of
// the solve, like the reference's timed loop.  HHP_PROFILE=1 adds the
// PROFILE_* lines: the fused kernels compute every dot inside an SpMV or
// vector pass, so PROFILE_dot is 0 and PROFILE_spmv is the iterations'
// SpMV launches timed with HIP events.
#include <argp.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hbs.h"

namespace {

struct Args {
    const char *input_matrix = nullptr;
    const char *input_x = nullptr;
    const char *input_y = nullptr;
    const char *output_x = nullptr;
    int n_iters = 0;
    double tol = 0.0;
    int mixed = 0;
    int check_every = 4;
};

enum { OPT_MIXED = 1000, OPT_CHECK = 1001 };

argp_option options[] = {
    {"input-matrix", 'm', "FILE", 0, "Input matrix market file path", 0},
    {"input-x", 'x', "FILE", 0, "Input X vector file path", 0},
    {"input-y", 'y', "FILE", 0, "Input Y target vector file path", 0},
    {"output-x", 'o', "FILE", 0, "Output X vector file path", 0},
    {"n-iters", 'n', "POSITIVE-INTEGER", 0, "Number of iterations", 0},
    {"tol", 't', "TOL", 0, "Stop at ||r|| < TOL*||r0|| (0: run exactly n-iters)", 0},
    {"mixed", OPT_MIXED, nullptr, 0, "FP32 BiCGStab sweeps inside FP64 refinement", 0},
    {"check-every", OPT_CHECK, "K", 0, "Iterations per convergence poll", 0},
    {nullptr, 0, nullptr, 0, nullptr, 0}};

error_t parse_opt(int key, char *arg, argp_state *state)
{
    Args *a = static_cast<Args *>(state->input);
    switch (key) {
    case 'm': a->input_matrix = arg; break;
    case 'x': a->input_x = arg; break;
    case 'y': a->input_y = arg; break;
    case 'o': a->output_x = arg; break;
    case 'n': a->n_iters = (int)strtol(arg, nullptr, 10); break;
    case 't': a->tol = strtod(arg, nullptr); break;
    case OPT_MIXED: a->mixed = 1; break;
    case OPT_CHECK: a->check_every = (int)strtol(arg, nullptr, 10); break;
    case ARGP_KEY_ARG: return 0;
    default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

char doc[] = "BiCGStab on one MI355X (libhbs) -- solves A x = y from the initial x";
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
    if (!args.input_matrix || !args.input_x || !args.output_x) {
        fprintf(stderr, "Error: All input and output files must be specified.\n");
        return EXIT_FAILURE;
    }
    for (const char *f : {args.input_matrix, args.input_x, args.input_y}) {
        if (f && access(f, F_OK) == -1) {
            fprintf(stderr, "Input file '%s' does not exist\n", f);
            return EXIT_FAILURE;
        }
    }
    if (!args.input_y) {
        fprintf(stderr, "Error: -y (right-hand side) must be specified.\n");
        return EXIT_FAILURE;
    }
    const bool profile = getenv("HHP_PROFILE") != nullptr;
    const double t_begin = now();

    hbs_ctx ctx = nullptr;
    if (hbs_ctx_create(0, nullptr, &ctx) != HBS_OK) return die("hbs_ctx_create");
    const double t_read0 = now();
    int m = 0, n = 0, nnz = 0;
    int *rp = nullptr, *col = nullptr;
    double *val = nullptr;
    if (hbs_mm_read(args.input_matrix, &m, &n, &nnz, &rp, &col, &val) != HBS_OK) return die("hbs_mm_read");
    hbs_csr A = nullptr;
    if (hbs_csr_create(ctx, m, n, nnz, rp, col, val, args.mixed ? HBS_CSR_F32 : 0, &A) != HBS_OK)
        return die("hbs_csr_create");
    hbs_host_free(rp);
    hbs_host_free(col);
    hbs_host_free(val);
    printf("Matrix name : %s\n", args.input_matrix);
    std::vector<double> hx((size_t)n), hb((size_t)m);
    if (hbs_vector_read(args.input_x, n, hx.data()) != HBS_OK) return die("hbs_vector_read");
    if (hbs_vector_read(args.input_y, m, hb.data()) != HBS_OK) return die("hbs_vector_read");
    double *dx = nullptr, *db = nullptr;
    if (hbs_dev_alloc(ctx, sizeof(double) * (size_t)n, (void **)&dx) != HBS_OK ||
        hbs_dev_alloc(ctx, sizeof(double) * (size_t)m, (void **)&db) != HBS_OK ||
        hbs_h2d(ctx, dx, hx.data(), sizeof(double) * (size_t)n) != HBS_OK ||
        hbs_h2d(ctx, db, hb.data(), sizeof(double) * (size_t)m) != HBS_OK)
        return die("device vectors");
    const double t_read1 = now();

    hbs_solver S = nullptr;
    if (hbs_solver_create(A, &S) != HBS_OK) return die("hbs_solver_create");
    hbs_solve_opts o;
    hbs_solve_opts_default(&o);
    o.maxit = args.n_iters;
    o.tol = args.tol;
    o.fixed_iters = args.tol > 0.0 ? 0 : 1;
    o.check_every = args.check_every;
    o.max_sweeps = 64;
    o.inner_maxit = args.n_iters;
    hbs_solve_info info{};
    const double t_spmv0 = now();
    const int rc = args.mixed ? hbs_bicgstab_mixed(S, db, dx, &o, &info) : hbs_bicgstab_f64(S, db, dx, &o, &info);
    const double t_spmv1 = now();
    if (rc != HBS_OK && rc != HBS_ERR_NOCONV) return die("BiCGStab");
    const double t_end = now();

    printf("n_iters : %d \n"
           "spmv : %lf \n"
           "file_read : %lf \n"
           "relative_residual : %E\n",
           info.iters, t_spmv1 - t_spmv0, t_read1 - t_read0, info.relres);
    printf("everything_total : %lf\n", t_end - t_begin);
    if (profile) {
        double ms = 0.0;
        double *dy = nullptr;
        if (hbs_dev_alloc(ctx, sizeof(double) * (size_t)m, (void **)&dy) != HBS_OK ||
            hbs_bench_spmv_f64(A, dx, dy, 10, &ms) != HBS_OK)
            return die("profile");
        const double loop = t_spmv1 - t_spmv0;
        const double spmv = std::fmin(loop, 2.0 * info.iters * ms * 1e-3);
        const double vec = loop - spmv;
        const double sum = spmv + vec;
        printf("PROFILE_enabled : 1\n");
        printf("PROFILE_loop_total : %lf\n", loop);
        printf("PROFILE_dot : %lf\n", 0.0);
        printf("PROFILE_spmv : %lf\n", spmv);
        printf("PROFILE_vec : %lf\n", vec);
        printf("PROFILE_sum : %lf\n", sum);
        printf("PROFILE_dot_pct : %.2f\n", 0.0);
        printf("PROFILE_spmv_pct : %.2f\n", sum > 0 ? 100.0 * spmv / sum : 0.0);
        printf("PROFILE_vec_pct : %.2f\n", sum > 0 ? 100.0 * vec / sum : 0.0);
        hbs_dev_free(ctx, dy);
    }
    printf("\n----------------------------------------------------------------------\n");

    if (hbs_d2h(ctx, hx.data(), dx, sizeof(double) * (size_t)n) != HBS_OK) return die("hbs_d2h");
    if (hbs_vector_write(args.output_x, n, hx.data()) != HBS_OK) return die("hbs_vector_write");
    hbs_solver_destroy(S);
    hbs_csr_destroy(A);
    hbs_dev_free(ctx, dx);
    hbs_dev_free(ctx, db);
    hbs_ctx_destroy(ctx);
    return EXIT_SUCCESS;
}
