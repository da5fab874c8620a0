#!/usr/bin/env python3
"""Benchmark: BiCGStab time to a 1e-8 relative residual on one MI355X,
mixed precision (FP32 sweeps inside FP64 iterative refinement) against the
plain FP64 solve, on the 3.5M-row circuit-like system.

Workload (one "step" = one full mixed-precision solve, inputs resident in
HBM): n = 3,500,000 rows; off-diagonal counts per row from a truncated
Zipf(2.1) calibrated to 4.2 entries per row on average, plus 200 rows of
1k..50k entries; uniform random columns; values N(0,1) * 10^U(-3,3);
diagonal 1.1 x sum|a_ij|; then every row divided by its diagonal (row
equilibration: unpreconditioned BiCGStab stagnates near 1e-5 on the raw
matrix, in the CPU oracle too).  x_target ~ U[-1,1] (seed 33), b = A
x_target, x0 = 0, tol 1e-8 on ||r|| / ||r0||.

Printed on one JSON line: the mixed solve time (value), the FP64 solve time
and the ratio beside it, the mixed result's relative error against the FP64
CPU oracle, the dominant kernel's (the x-segmented FP32 SpMV) achieved
bandwidth against the HBM roofline measured with HIP events on the
library's stream, and the CPU oracle's FP64 solve of the same system on one
host core, with the all-core OpenMP SpMV beside it.  Extras: the 445k-row
Markov-like system (FP64 solve, its SpMV, which fits the Infinity Cache)
and one rank's share of it, the 128^3 x 64 block solve, 256 uneven
systems, and the 1.45M-row FEM forward + adjoint.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "BiCGStab time-to-1e-8 residual & SpMV HBM GB/s (%peak): mixed vs FP64, 1 GPU"
N3 = 3500000


def spmv_bytes(m, n, nnz, v=8):
    """Algorithmic bytes of one CSR SpMV: values+indices, row pointers, y
    written once, x read once."""
    return nnz * (v + 4) + 4 * (m + 1) + v * m + v * n


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_value(fname, key):
    """A 'key: value' line of a committed profile summary under profiles/."""
    p = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except (ValueError, IndexError):
                return None
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def circuit(hbs):
    rp, c, v = hbs.gen_circuit(N3, 4.2, 2.1, 200, 1000, 50000, 1.1, 3, equilibrate=True)
    xt = hbs.gen_uniform(N3, -1.0, 1.0, 33)
    b = hbs.host_spmv(rp, c, v, xt)
    return rp, c, v, xt, b


def cpu_baseline(rp, c, v, b, xs):
    """The oracle's restatement of the reference CPU BiCGStab (single
    thread, FP64, no FMA, the reference's stop test) on the same solve, timed
    on this host; beside it the oracle's OpenMP SpMV (CSR_spmxv_omp,
    hhp_cpu.c:459-468) on every host core.  Also the mixed result's relative
    error against the oracle's FP64 solution."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    A = orc.CSR(rp, c, v)
    t0 = time.perf_counter()
    x, it, rr, stop = orc.bicgstab(A, b, tol=1e-8, maxit=5000)
    dt = time.perf_counter() - t0
    n = rp.size - 1
    nnz = int(c.size)
    xv = np.random.default_rng(1).uniform(-1, 1, n)
    orc.spmv(A, xv, omp=True)
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        orc.spmv(A, xv, omp=True)
    omp_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    orc.spmv(A, xv)
    seq_ms = (time.perf_counter() - t0) * 1e3
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    errs = {k: float(np.linalg.norm(xk - x) / np.linalg.norm(x)) for k, xk in xs.items()}
    res = {"value": round(dt * 1e3, 1), "unit": "ms per FP64 solve to 1e-8", "cores": 1, "kind": "port",
           "sample": f"one full FP64 solve of the same 3.5M-row system ({it} iterations, relres {rr:.2e}), "
                     "oracle/hhp_oracle.c (bicgstab_cpu.c:172-255 + stop test), one thread",
           "iters": it, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "omp_spmv": {"ms": round(omp_ms, 3), "threads": threads,
                        "gbs": round(spmv_bytes(n, n, nnz) / (omp_ms * 1e-3) / 1e9, 1),
                        "what": "oracle CSR_spmxv_omp restatement, schedule(guided), FP64, 10 reps"},
           "seq_spmv_ms": round(seq_ms, 3)}
    return res, errs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spmv-reps", type=int, default=20)
    ap.add_argument("--check-every", type=int, default=8,
                    help="iterations per convergence poll of the circuit solves (measured on MI355X: 4, 6, 8, "
                         "12, 16 gave mixed 12.58, 12.55, 12.36, 12.57, 13.49 ms)")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", choices=["circuit", "markov", "fem", "block"],
                    help="only the SpMV launches and a copy of known size (for counter runs)")
    ap.add_argument("--circuit", action="store_true",
                    help="only the circuit solves and SpMVs (for kernel traces)")
    ap.add_argument("--markov", action="store_true",
                    help="only the 445k-row Markov-like solves, SpMV and shard (for kernel traces)")
    ap.add_argument("--batched", action="store_true",
                    help="only the 128^3 x 64 block solve and the 256 uneven systems (for kernel traces)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(one rank's share runs on one GPU, nothing here spans a second)", file=sys.stderr)
        sys.exit(2)

    from hethyppar_amd import hbs

    if args.probe:
        probe(hbs, args.probe)
        return
    if args.circuit:
        args.no_extras = args.no_cpu = True
    if args.markov:
        ctx = hbs.Context(0)
        print(json.dumps({"markov_445k_fp64_cache_resident": extras_markov(hbs, ctx, args)}), flush=True)
        ctx.close()
        return
    if args.batched:
        ctx = hbs.Context(0)
        print(json.dumps(extras_batched(hbs, ctx)), flush=True)
        ctx.close()
        return

    log(f"bench: generating circuit-like n={N3} (rows equilibrated)")
    rp, c, v, xt, b = circuit(hbs)
    nnz = int(c.size)
    ctx = hbs.Context(0)
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    S = hbs.Solver(A)
    db, dx = ctx.to_device(b), ctx.empty(N3)
    opts = hbs.default_opts(tol=1e-8, maxit=5000, use_graph=2, check_every=args.check_every)

    def solve(mixed):
        dx.zero()
        return (S.bicgstab_mixed if mixed else S.bicgstab)(db, dx, opts)

    log(f"bench: warmup {args.warmup} (mixed and FP64)")
    for _ in range(args.warmup):
        solve(True)
        solve(False)
    ctx.sync()
    out = {}
    xs = {}
    for mixed in (True, False):
        log(f"bench: timing {args.steps} {'mixed' if mixed else 'FP64'} solves")
        ctx.sync()
        t0 = time.perf_counter()
        infos = [solve(mixed) for _ in range(args.steps)]
        ctx.sync()
        out[mixed] = ((time.perf_counter() - t0) * 1e3 / args.steps, infos)
        xs["mixed" if mixed else "fp64"] = dx.download()
    ms_mixed, im = out[True]
    ms_fp64, i64 = out[False]

    # dominant kernel: the x-segmented FP32 SpMV (k_xseg + k_chunks + row
    # sums, the diagonal apart), one HIP event pair on the library's stream
    # around back-to-back launches
    log("bench: SpMV roofline")
    diag_split = A.info()["diag_split"]
    rows = np.repeat(np.arange(N3, dtype=c.dtype), np.diff(rp))
    nnz_diag = int(np.unique(rows[c == rows]).size)   # rows with a diagonal entry: one each goes apart
    x32, y32 = ctx.to_device(xt.astype(np.float32)), ctx.empty(N3, np.float32)
    spmv32_ms = A.bench_spmv(x32, y32, args.spmv_reps)
    x64, y64 = ctx.to_device(xt), ctx.empty(N3)
    spmv64_ms = A.bench_spmv(x64, y64, args.spmv_reps)
    B32, B64 = spmv_bytes(N3, N3, nnz, 4), spmv_bytes(N3, N3, nnz, 8)
    achieved = B32 / (spmv32_ms * 1e-3) / 1e9
    import ctypes as C
    copy_n = 1 << 27   # 1 GiB per copy buffer, larger than the 256 MB Infinity Cache
    dsrc, ddst = ctx.empty(copy_n), ctx.empty(copy_n)
    dsrc.zero()
    cms = C.c_double()
    hbs.check(hbs.lib().hbs_bench_copy(ctx.h, dsrc.ptr, ddst.ptr, copy_n, 10, C.byref(cms)))
    copy_gbs = 2 * 8 * copy_n / (cms.value * 1e-3) / 1e9
    for d in (dsrc, ddst, x32, y32, x64, y64):
        d.free()

    result = {
        "metric": METRIC,
        "value": round(ms_mixed, 4),
        "unit": "ms per mixed-precision solve to 1e-8 relative residual (FP32 sweeps, FP64 refinement)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "circuit_like_n3500000_row_equilibrated", "n": N3, "nnz": nnz,
                   "rows": "truncated Zipf(2.1) off-diagonal counts, mean 4.2 per row, 200 rows of 1k-50k; "
                           "uniform columns; N(0,1)*10^U(-3,3); diagonal 1.1*sum|a|; each row divided by "
                           "its diagonal",
                   "x0": "zeros", "tol": 1e-8, "check_every": args.check_every},
        "fp64_ms": round(ms_fp64, 4),
        "ratio_mixed_over_fp64": round(ms_mixed / ms_fp64, 4),
        "mixed": {"iters": im[-1].iters, "sweeps": im[-1].sweeps, "relres": im[-1].relres,
                  "device_ms": round(float(np.mean([i.time_ms for i in im])), 4)},
        "fp64": {"iters": i64[-1].iters, "relres": i64[-1].relres,
                 "device_ms": round(float(np.mean([i.time_ms for i in i64])), 4)},
        "iters_per_s_fp64": round(i64[-1].iters / (ms_fp64 * 1e-3), 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profile_value("spmv_circuit3p5M_pmc.txt", "traffic_bytes_per_spmv_fp32"),
                     "kernel": "x-segmented FP32 SpMV, diagonal apart (k_xseg<float,16> + k_chunks<float> + "
                               "k_spmv<float,EPI_AXPBY,16,512,true,DG>)",
                     "launch_ms": round(spmv32_ms, 5), "bytes_per_launch": B32,
                     "bytes_formula": "nnz*(4+4) + 4*(m+1) + 4*m + 4*n (CSR, every entry)",
                     "diag_split": diag_split,
                     "format_bytes": B32 - 8 * nnz_diag + (4 * N3 if diag_split == 2 else 0),
                     "format_formula": "(nnz - diagonal entries)*(4+4) + 4*(m+1) + 4*m + 4*n "
                                       "(+ 4*m diagonal values unless unit)"},
        "spmv_fp64": {"ms": round(spmv64_ms, 5), "bytes": B64,
                      "gbs": round(B64 / (spmv64_ms * 1e-3) / 1e9, 1),
                      "frac": round(B64 / (spmv64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "copy_probe_gbs": round(copy_gbs, 1),
    }
    S.close()
    A.close()

    if not args.no_extras:
        extras = {}
        extras["markov_445k_fp64_cache_resident"] = extras_markov(hbs, ctx, args)
        extras.update(extras_batched(hbs, ctx))
        extras["fem_1p45M_adjoint"] = extras_adjoint(hbs, ctx)
        result["extras"] = extras

    if not args.no_cpu:
        log("bench: CPU baseline (oracle FP64 solve on 1 thread; OpenMP SpMV on all cores)")
        cb, errs = cpu_baseline(rp, c, v, b, xs)
        result["cpu_baseline"] = cb
        result["rel_err_vs_fp64_oracle"] = {"mixed": errs["mixed"], "fp64": errs["fp64"]}
    print(json.dumps(result), flush=True)
    ctx.close()


def probe(hbs, which):
    """The launches the counter profiles read: 1 + 10 FP32 and 1 + 10 FP64
    SpMV launches of one workload, then 1 + 3 copies of 2^25 doubles (256 MiB
    read and 256 MiB written per launch), the scale of FETCH_SIZE /
    WRITE_SIZE."""
    import ctypes as C
    ctx = hbs.Context(0)
    if which == "block":
        probe_block(hbs, ctx)
        ctx.close()
        return
    flags = hbs.HBS_CSR_F32
    if which == "circuit":
        rp, c, v, xt, _ = circuit(hbs)
    elif which == "fem":
        rp, c, v = hbs.gen_fem(483334, 12, 1.1, 6)
        xt = hbs.gen_uniform(rp.size - 1, -1.0, 1.0, 62)
        flags |= hbs.HBS_CSR_TRANSPOSE
    else:
        rp, c, v = hbs.gen_markov(445000, 12, 22, 8900, 1.05, 2)
        xt = hbs.gen_uniform(445000, -1.0, 1.0, 12)
    n, nnz = rp.size - 1, int(c.size)
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=flags)
    x32, y32 = ctx.to_device(xt.astype(np.float32)), ctx.empty(n, np.float32)
    ms32 = A.bench_spmv(x32, y32, 10)
    x64, y64 = ctx.to_device(xt), ctx.empty(n)
    ms64 = A.bench_spmv(x64, y64, 10)
    extra = {}
    if flags & hbs.HBS_CSR_TRANSPOSE:
        # the adjoint's SpMV: A^T x, FP64, 1 + 10 launches
        A.spmv_t(x64, y64)
        ctx.timer_start()
        for _ in range(10):
            A.spmv_t(x64, y64)
        extra["spmv_t_fp64_ms"] = ctx.timer_stop() / 10
    cn = 1 << 25
    s, d = ctx.empty(cn), ctx.empty(cn)
    s.zero()
    cms = C.c_double()
    hbs.check(hbs.lib().hbs_bench_copy(ctx.h, s.ptr, d.ptr, cn, 3, C.byref(cms)))
    info = A.info()
    print(json.dumps({"probe": which, "n": n, "nnz": nnz, "xseg": info["xseg"], "block3": info["block3"],
                      "diag_split": info["diag_split"],
                      "spmv_fp32_ms": ms32, "spmv_fp32_bytes": spmv_bytes(n, n, nnz, 4),
                      "spmv_fp64_ms": ms64, "spmv_fp64_bytes": spmv_bytes(n, n, nnz, 8), **extra,
                      "copy_doubles": cn, "copy_ms": cms.value}), flush=True)
    ctx.close()


def probe_block(hbs, ctx):
    """The 64-column SpMM of the 128^3 operator: 1 + 10 FP64 and 1 + 10 FP32
    launches (one 64-column tile), then the copy of 2^25 doubles 1 + 3
    times."""
    import ctypes as C
    rp, c, v = hbs.gen_convdiff3d(128, 128, 128, 0.5)
    n, nnz, k = rp.size - 1, int(c.size), 64
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    X = hbs.gen_uniform(n * k, -1.0, 1.0, 4)
    res = {"probe": "block", "n": n, "nnz": nnz, "k": k, "spmm_runs": A.info()["spmm_runs"]}
    for name, dt, eb in (("fp64", np.float64, 8), ("fp32", np.float32, 4)):
        dX, dY = ctx.to_device(X.astype(dt)), ctx.empty(n * k, dt)
        A.spmm(dX, dY, k, 64)
        ctx.timer_start()
        for _ in range(10):
            A.spmm(dX, dY, k, 64)
        ms = ctx.timer_stop() / 10
        res[f"spmm_{name}_ms"] = ms
        res[f"spmm_{name}_bytes"] = nnz * (4 + eb) + 4 * (n + 1) + 2 * k * eb * n
        dX.free()
        dY.free()
    cn = 1 << 25
    s, d = ctx.empty(cn), ctx.empty(cn)
    s.zero()
    cms = C.c_double()
    hbs.check(hbs.lib().hbs_bench_copy(ctx.h, s.ptr, d.ptr, cn, 3, C.byref(cms)))
    res.update({"copy_doubles": cn, "copy_ms": cms.value})
    print(json.dumps(res), flush=True)


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


def extras_markov(hbs, ctx, args):
    """The 445k-row Markov-like system: FP64 solve and SpMV (matrix and
    vectors, 110 MB, fit the 256 MB Infinity Cache, so the SpMV rate is
    cache-resident, not HBM), and one rank's share of it: part 0 of the
    library's own 4-way partition, local+halo SpMV with the halo from the
    device and from a pinned host buffer (H2D + unpack + SpMV)."""
    n = 445000
    log(f"bench: Markov-like n={n}")
    rp, c, v = hbs.gen_markov(n, 12, 22, max(1, int(round(0.02 * n))), 1.05, 2)
    nnz = int(c.size)
    xt = hbs.gen_uniform(n, -1.0, 1.0, 12)
    b = hbs.host_spmv(rp, c, v, xt)
    A = hbs.DeviceCSR(ctx, rp, c, v)
    S = hbs.Solver(A)
    db, dx = ctx.to_device(b), ctx.empty(n)
    o = hbs.default_opts(tol=1e-8, maxit=1000, use_graph=2, check_every=4)

    def solve():
        dx.zero()
        return S.bicgstab(db, dx, o)
    for _ in range(3):
        solve()
    ctx.sync()
    t0 = time.perf_counter()
    infos = [solve() for _ in range(20)]
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / 20
    dxs, dys = ctx.to_device(xt), ctx.empty(n)
    spmv_ms = A.bench_spmv(dxs, dys, 50)
    B = spmv_bytes(n, n, nnz)
    res = {"n": n, "nnz": nnz, "solve_ms": round(ms, 4), "iters": infos[-1].iters, "relres": infos[-1].relres,
           "spmv": {"ms": round(spmv_ms, 5), "bytes": B, "gbs": round(B / (spmv_ms * 1e-3) / 1e9, 1),
                    "traffic": profile_value("spmv_markov445k_pmc.txt", "traffic_bytes_per_launch"),
                    "note": "cache-resident: 110 MB fit the 256 MB Infinity Cache"}}
    S.close()
    A.close()

    log("bench: shard part 0 of 4 (hbs_partition)")
    t0 = time.perf_counter()
    pv, cut = hbs.partition(rp, c, 4)
    part_s = time.perf_counter() - t0
    hs = hbs.shard_build(rp, c, v, pv, 4, 0)
    sh = hbs.Shard(ctx, hs)
    Am = sh.csr()
    m, ns = sh.m_loc, sh.n_shr
    xe = ctx.to_device(np.random.default_rng(1).uniform(-1, 1, m + ns))
    ye = ctx.empty(m)
    sms = Am.bench_spmv(xe, ye, 50)
    Bs = spmv_bytes(m, m + ns, sh.nnz)
    # halo from a pinned host buffer: H2D of n_shr values, unpack through
    # recv.J, merged SpMV; 50 back-to-back calls between one event pair, the
    # best of 3 such batches after 20 warm-up calls (the first batch after
    # the host-side shard build has run up to 20x slower on the box: a
    # transient on the host, gone by the next batch)
    import ctypes as C
    pin = hbs.PinnedArray(ctx, ns)
    pin.a[:] = np.random.default_rng(2).uniform(-1, 1, ns)
    hp = pin.a.ctypes.data_as(C.POINTER(C.c_double))
    for _ in range(20):
        hbs.check(hbs.lib().hbs_shard_spmv(sh.h, xe.ptr, hp, ye.ptr))
    ctx.sync()
    batches = []
    for _ in range(3):
        ctx.timer_start()
        for _ in range(50):
            hbs.check(hbs.lib().hbs_shard_spmv(sh.h, xe.ptr, hp, ye.ptr))
        batches.append(ctx.timer_stop() / 50)
    host_ms = min(batches)
    res["shard_part0"] = {"partition": "hbs_partition (CalcPartVec model), 4 parts, seed 42",
                          "partition_s": round(part_s, 2), "cut": int(cut),
                          "part_rows": np.bincount(pv, minlength=4).tolist(),
                          "m_loc": m, "n_shr": ns, "n_send": sh.n_send, "nnz": sh.nnz,
                          "spmv_device_halo_ms": round(sms, 5),
                          "spmv_device_halo_gbs": round(Bs / (sms * 1e-3) / 1e9, 1),
                          "spmv_host_halo_ms": round(host_ms, 5),
                          "spmv_host_halo_batches_ms": [round(b, 5) for b in batches],
                          "host_halo_bytes": ns * 8}
    pin.close()
    sh.close()
    hbs.shard_free(hs)
    return res


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
    res = {"n_total": n, "nnz": int(c.size)}
    for label, graph in (("graph", 1), ("eager", 0)):
        o = hbs.default_opts(tol=1e-8, maxit=5000, use_graph=graph, check_every=8)

        def systems():
            dx.zero()
            return S.bicgstab_systems(db, dx, o, raise_on_error=False)
        systems()
        ms, infos = timed(ctx, systems, reps=2)
        res[f"ms_{label}"] = round(ms, 2)
    its = sorted(i.iters for i in infos)
    res.update({"iters_median": its[len(its) // 2], "iters_max": its[-1],
                "relres_max": max(i.relres for i in infos),
                "converged": sum(1 for i in infos if i.status == 0)})
    out["systems_256_uneven"] = res
    S.close()
    A.close()
    return out


def extras_adjoint(hbs, ctx):
    """FEM-like 3x3-block system, n = 1,450,002, 52.2M entries: FP64 solve,
    then the adjoint of L = 0.5 ||x - x_target||^2 (A^T solve + gradient on
    the pattern).  655 MB per SpMV: HBM-resident."""
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
           "spmv": {"ms": round(spmv_ms, 4), "gbs": round(B / (spmv_ms * 1e-3) / 1e9, 1),
                    "frac": round(B / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "spmv_transpose": {"ms": round(t_ms, 4), "gbs": round(B / (t_ms * 1e-3) / 1e9, 1)}}
    S.close()
    A.close()
    return res


if __name__ == "__main__":
    main()
