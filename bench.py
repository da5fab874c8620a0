#!/usr/bin/env python3
"""Benchmark: FP64 BiCGStab time-to-1e-8 on one MI355X.

Workload (one "step" = one full solve, inputs resident in HBM): nonsymmetric
Markov-like matrix, n = 445,000 rows, per row 12..22 random columns in a band
of +-2% of n plus one far column, U[0,1] values, diagonal 1.05 x row sum;
x_target ~ U[-1,1], b = A x_target, x0 = 0, tol 1e-8 on ||r||/||r0||.

Printed on one JSON line: time per solve, iterations, the dominant kernel's
(CSR-stream SpMV) achieved bandwidth against the HBM roofline measured with
HIP events on the solver's stream, and the CPU oracle's time for the same
solve on one host core.  Extras (mixed precision vs FP64 on the circuit-like
3.5M-row matrix, one rank's local+halo SpMV) follow in the same line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "BiCGStab time-to-1e-8 residual & SpMV HBM GB/s (%peak): mixed vs FP64, 1 GPU"


def spmv_bytes(m, n, nnz, v=8):
    """Algorithmic bytes of one CSR SpMV: values+indices, row pointers, y
    written once, x read once."""
    return nnz * (v + 4) + 4 * (m + 1) + v * m + v * n


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per SpMV launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size)."""
    p = os.path.join(ROOT, "profiles", "spmv_markov445k_pmc.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(rp, c, v, b, budget_s=30.0):
    """The oracle's restatement of the reference CPU BiCGStab (single thread,
    FP64, no FMA) on the same solve, timed on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    A = orc.CSR(rp, c, v)
    t0 = time.perf_counter()
    x, it, rr, stop = orc.bicgstab(A, b, tol=1e-8, maxit=1000)
    dt = time.perf_counter() - t0
    return {"value": dt * 1e3, "unit": "ms per solve to 1e-8", "cores": 1, "kind": "port",
            "sample": f"one full FP64 solve of the same 445k-row system ({it} iterations, "
                      f"relres {rr:.2e}), oracle/hhp_oracle.c single thread",
            "iters": it}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=445000)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--check-every", type=int, default=4,
                    help="iterations per convergence poll of the FP64 solve")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="only the SpMV launches and the copy probe (for counter runs)")
    ap.add_argument("--circuit", action="store_true",
                    help="only the circuit-like 3.5M FP64 vs mixed solves and SpMVs (for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(one rank's share runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    from hethyppar_amd import hbs

    if args.circuit:
        ctx = hbs.Context(0)
        print(json.dumps({"circuit_3p5M": extras_circuit(hbs, ctx)}), flush=True)
        ctx.close()
        return
    n = args.n
    log(f"bench: generating Markov-like n={n}")
    rp, c, v = hbs.gen_markov(n, 12, 22, max(1, int(round(0.02 * n))), 1.05, 2)
    nnz = int(c.size)
    xt = hbs.gen_uniform(n, -1.0, 1.0, 12)
    b = hbs.host_spmv(rp, c, v, xt)

    ctx = hbs.Context(0)
    A = hbs.DeviceCSR(ctx, rp, c, v)
    if args.probe:
        probe(hbs, ctx, A, xt, n, nnz)
        return
    S = hbs.Solver(A)
    db = ctx.to_device(b)
    dx = ctx.empty(n)
    opts = hbs.default_opts(tol=1e-8, maxit=1000, use_graph=2, check_every=args.check_every)

    def solve():
        dx.zero()
        return S.bicgstab(db, dx, opts)

    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        solve()
    ctx.sync()
    log(f"bench: timing {args.steps} solves")
    times, infos = [], []
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        info = solve()
        times.append(time.perf_counter() - t0)
        infos.append(info)
    ctx.sync()
    t_all = time.perf_counter() - t_all0
    ms_per_step = t_all * 1e3 / args.steps
    dev_ms = float(np.mean([i.time_ms for i in infos]))
    it = infos[-1].iters
    relres = infos[-1].relres

    # dominant kernel: CSR-stream SpMV, one HIP event pair around 50 back-to-back launches
    log("bench: SpMV roofline")
    dxs, dys = ctx.to_device(xt), ctx.empty(n)
    spmv_ms = A.bench_spmv(dxs, dys, args.spmv_reps)
    B = spmv_bytes(n, n, nnz)
    achieved = B / (spmv_ms * 1e-3) / 1e9
    copy_n = 1 << 27   # 1 GiB per copy buffer, larger than the 256 MB Infinity Cache
    dsrc, ddst = ctx.empty(copy_n), ctx.empty(copy_n)
    dsrc.zero()
    import ctypes as C
    cms = C.c_double()
    hbs.check(hbs.lib().hbs_bench_copy(ctx.h, dsrc.ptr, ddst.ptr, copy_n, 10, C.byref(cms)))
    copy_gbs = 2 * 8 * copy_n / (cms.value * 1e-3) / 1e9
    dsrc.free()
    ddst.free()

    result = {
        "metric": METRIC,
        "value": round(ms_per_step, 4),
        "unit": "ms per FP64 solve to 1e-8 relative residual",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"markov_like_n{n}_fp64_bicgstab_tol1e-8", "n": n, "nnz": nnz,
                   "band": "+-2% of n, 12..22 per row + 1 far", "x0": "zeros", "tol": 1e-8},
        "iters": it,
        "relres": relres,
        "device_ms_per_solve": round(dev_ms, 4),
        "iters_per_s": round(it / (ms_per_step * 1e-3), 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(),
                     "kernel": "k_spmv<double, EPI_AXPBY>", "launch_ms": round(spmv_ms, 5),
                     "bytes_per_launch": B,
                     "note": "matrix+vectors (110 MB) fit the 256 MB Infinity Cache"},
        "copy_probe_gbs": round(copy_gbs, 1),
    }

    if not args.no_extras:
        result["extras"] = extras(hbs, ctx, rp, c, v, n)
        result["extras"].update(extras_batched(hbs, ctx))
        result["extras"]["fem_1p45M_adjoint"] = extras_adjoint(hbs, ctx)

    if not args.no_cpu:
        log("bench: CPU baseline (oracle, 1 thread)")
        result["cpu_baseline"] = cpu_baseline(rp, c, v, b)
    print(json.dumps(result), flush=True)
    ctx.close()


def probe(hbs, ctx, A, xt, n, nnz):
    """The launches the counter profiles read: 1 + 10 FP64 SpMV launches of the
    bench matrix, then 1 + 3 copies of 2^25 doubles (256 MiB read and 256 MiB
    written per launch), the scale of FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C
    dx, dy = ctx.to_device(xt), ctx.empty(n)
    ms = A.bench_spmv(dx, dy, 10)
    cn = 1 << 25
    s, d = ctx.empty(cn), ctx.empty(cn)
    s.zero()
    cms = C.c_double()
    hbs.check(hbs.lib().hbs_bench_copy(ctx.h, s.ptr, d.ptr, cn, 3, C.byref(cms)))
    print(json.dumps({"probe": "spmv+copy", "n": n, "nnz": nnz, "spmv_ms": ms,
                      "spmv_bytes": spmv_bytes(n, n, nnz), "copy_doubles": cn,
                      "copy_ms": cms.value}), flush=True)
    ctx.close()


def timed(ctx, fn, reps=1):
    """Device-synchronised wall time of fn() in ms (best of reps) and its
    last result."""
    best, res = None, None
    for _ in range(reps):
        ctx.sync()
        t0 = time.perf_counter()
        res = fn()
        ctx.sync()
        dt = (time.perf_counter() - t0) * 1e3
        best = dt if best is None else min(best, dt)
    return best, res


def extras_batched(hbs, ctx):
    """128^3 7-point convection-diffusion with 64 right-hand sides (one
    BiCGStab per column over a shared SpMM), and 256 uneven 2-D systems in
    the same launches."""
    out = {}
    log("bench: 128^3 x 64 right-hand sides")
    rp, c, v = hbs.gen_convdiff3d(128, 128, 128, 0.5)
    n, k = rp.size - 1, 64
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    S = hbs.Solver(A)
    dXt = ctx.to_device(hbs.gen_uniform(n * k, -1.0, 1.0, 4))
    dB, dX = ctx.empty(n * k), ctx.empty(n * k)
    A.spmm(dXt, dB, k)
    o = hbs.default_opts(tol=1e-8, maxit=2000, use_graph=1, check_every=8)

    def block():
        dX.zero()
        return S.bicgstab_block(dB, dX, k, o)
    block()
    ms, infos = timed(ctx, block)
    its = [i.iters for i in infos]
    nnz = int(c.size)
    spmm_b = nnz * 12 + 4 * (n + 1) + k * 8 * 2 * n
    ctx.timer_start()
    for _ in range(5):
        A.spmm(dXt, dB, k)
    sp_ms = ctx.timer_stop() / 5
    dXt32, dB32 = ctx.empty(n * k, np.float32), ctx.empty(n * k, np.float32)
    A.spmm(dXt32, dB32, k)
    ctx.timer_start()
    for _ in range(5):
        A.spmm(dXt32, dB32, k)
    sp32_ms = ctx.timer_stop() / 5
    spmm32_b = nnz * 8 + 4 * (n + 1) + k * 4 * 2 * n
    out["block_128cube_k64"] = {
        "n": n, "nnz": nnz, "k": k, "ms": round(ms, 2), "iters_min": min(its), "iters_max": max(its),
        "relres_max": max(i.relres for i in infos), "status_worst": min(i.status for i in infos),
        "spmm_fp64": {"ms": round(sp_ms, 4), "gbs": round(spmm_b / (sp_ms * 1e-3) / 1e9, 1)},
        "spmm_fp32": {"ms": round(sp32_ms, 4), "gbs": round(spmm32_b / (sp32_ms * 1e-3) / 1e9, 1)}}
    for d in (dXt, dB, dX, dXt32, dB32):
        d.free()
    S.close()
    A.close()

    log("bench: 256 uneven systems")
    sysl = hbs.gen_uneven_systems(256)
    rp, c, v, rows = hbs.concat_systems([(a, b, d) for a, b, d, _ in sysl])
    n = int(rows[-1])
    b = hbs.host_spmv(rp, c, v, hbs.gen_uniform(n, -1.0, 1.0, 5))
    A = hbs.DeviceCSR(ctx, rp, c, v, sys_rows=rows)
    S = hbs.Solver(A)
    db, dx = ctx.to_device(b), ctx.empty(n)
    o = hbs.default_opts(tol=1e-8, maxit=5000, use_graph=2, check_every=8)

    def systems():
        dx.zero()
        return S.bicgstab_systems(db, dx, o, raise_on_error=False)
    systems()
    ms, infos = timed(ctx, systems, reps=2)
    its = sorted(i.iters for i in infos)
    out["systems_256_uneven"] = {
        "n_total": n, "nnz": int(c.size), "ms": round(ms, 2), "iters_median": its[len(its) // 2],
        "iters_max": its[-1], "relres_max": max(i.relres for i in infos),
        "converged": sum(1 for i in infos if i.status == 0)}
    S.close()
    A.close()
    return out


def extras_adjoint(hbs, ctx):
    """FEM-like 3x3-block system, n = 1,450,002, 52.2M entries: FP64 solve,
    then the adjoint of L = 0.5 ||x - x_target||^2 (A^T solve + gradient on
    the pattern)."""
    log("bench: FEM 1.45M forward + adjoint")
    rp, c, v = hbs.gen_fem(483334, 12, 1.1, 6)
    n, nnz = rp.size - 1, int(c.size)
    b = hbs.host_spmv(rp, c, v, hbs.gen_uniform(n, -1.0, 1.0, 61))
    xt = hbs.gen_uniform(n, -1.0, 1.0, 62)
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_TRANSPOSE)
    S = hbs.Solver(A)
    db, dx, dxt = ctx.to_device(b), ctx.empty(n), ctx.to_device(xt)
    dg, dl, dgv = ctx.empty(n), ctx.empty(n), ctx.empty(nnz)
    o = hbs.default_opts(tol=1e-8, maxit=2000, use_graph=2, check_every=4)

    def forward():
        dx.zero()
        return S.bicgstab(db, dx, o)

    def adjoint():
        return S.adjoint(dg, dx, dl, dgv, o)
    forward()
    fms, finfo = timed(ctx, forward, reps=2)
    x = dx.download()
    dg.upload(x - xt)
    adjoint()
    ams, ainfo = timed(ctx, adjoint, reps=2)
    ys = ctx.empty(n)
    spmv_ms = A.bench_spmv(dx, ys, 20)
    ctx.timer_start()
    for _ in range(20):
        A.spmv_t(dx, ys)
    t_ms = ctx.timer_stop() / 20
    B = spmv_bytes(n, n, nnz)
    res = {"n": n, "nnz": nnz,
           "forward": {"ms": round(fms, 3), "iters": finfo.iters, "relres": finfo.relres},
           "adjoint": {"ms": round(ams, 3), "iters": ainfo.iters, "relres": ainfo.relres},
           "spmv": {"ms": round(spmv_ms, 4), "gbs": round(B / (spmv_ms * 1e-3) / 1e9, 1)},
           "spmv_transpose": {"ms": round(t_ms, 4), "gbs": round(B / (t_ms * 1e-3) / 1e9, 1)}}
    S.close()
    A.close()
    return res


def extras_circuit(hbs, ctx):
    """FP64 vs mixed BiCGStab and the SpMV on the 3.5M-row circuit-like matrix."""
    N3 = 3500000
    log("bench: circuit-like n=3.5M, FP64 vs mixed")
    rp3, c3, v3 = hbs.gen_circuit(N3, 4.2, 2.1, 200, 1000, 50000, 1.1, 3, equilibrate=True)
    xt3 = hbs.gen_uniform(N3, -1.0, 1.0, 33)
    b3 = hbs.host_spmv(rp3, c3, v3, xt3)
    A3 = hbs.DeviceCSR(ctx, rp3, c3, v3, flags=hbs.HBS_CSR_F32)
    S3 = hbs.Solver(A3)
    db3, dx3 = ctx.to_device(b3), ctx.empty(N3)
    res = {}
    for mode in ("fp64", "mixed"):
        fn = S3.bicgstab if mode == "fp64" else S3.bicgstab_mixed
        o = hbs.default_opts(tol=1e-8, maxit=5000, use_graph=2, check_every=4)
        dx3.zero()
        fn(db3, dx3, o, raise_on_error=False)
        ts = []
        for _ in range(3):
            dx3.zero()
            t0 = time.perf_counter()
            info = fn(db3, dx3, o, raise_on_error=False)
            ts.append(time.perf_counter() - t0)
        res[mode] = {"ms": round(1e3 * min(ts), 3), "iters": info.iters, "sweeps": info.sweeps,
                     "relres": info.relres, "status": info.status}
    dys = ctx.empty(N3)
    dxs = ctx.to_device(xt3)
    ms64 = A3.bench_spmv(dxs, dys, 20)
    dxs32 = ctx.to_device(xt3.astype(np.float32))
    dys32 = ctx.empty(N3, np.float32)
    ms32 = A3.bench_spmv(dxs32, dys32, 20)
    nnz3 = int(c3.size)
    res["ratio_mixed_over_fp64"] = round(res["mixed"]["ms"] / res["fp64"]["ms"], 3)
    res["nnz"] = nnz3
    res["spmv_fp64"] = {"ms": round(ms64, 5), "gbs": round(spmv_bytes(N3, N3, nnz3) / (ms64 * 1e-3) / 1e9, 1)}
    res["spmv_fp32"] = {"ms": round(ms32, 5), "gbs": round(spmv_bytes(N3, N3, nnz3, 4) / (ms32 * 1e-3) / 1e9, 1)}
    S3.close()
    A3.close()
    return res


def extras(hbs, ctx, rp, c, v, n):
    out = {}
    # one rank's share: part 0 of a contiguous 4-way split (labelled substitute
    # for the closed PaToH partition), local+halo SpMV with host halo
    log("bench: shard part 0 of 4")
    pv = (np.arange(n, dtype=np.int64) * 4 // n).astype(np.int32)
    hs = hbs.shard_build(rp, c, v, pv, 4, 0)
    sh = hbs.Shard(ctx, hs)
    Am = sh.csr()
    m, ns = sh.m_loc, sh.n_shr
    xe = ctx.to_device(np.random.default_rng(1).uniform(-1, 1, m + ns))
    ye = ctx.empty(m)
    ms = Am.bench_spmv(xe, ye, 50)
    Bs = spmv_bytes(m, m + ns, sh.nnz)
    out["shard_part0"] = {"m_loc": m, "n_shr": ns, "n_send": sh.n_send, "nnz": sh.nnz,
                          "spmv_ms": round(ms, 5), "spmv_gbs": round(Bs / (ms * 1e-3) / 1e9, 1)}
    sh.close()
    hbs.shard_free(hs)

    out["circuit_3p5M"] = extras_circuit(hbs, ctx)
    return out


if __name__ == "__main__":
    main()
