"""Batched BiCGStab on the MI355X: many uneven systems in the same launches
(hbs_bicgstab_systems_f64) and k right-hand sides sharing each SpMM pass
(hbs_bicgstab_block_f64).  Neither has a reference counterpart; each system
or column is held to the single-system bar against the oracle's
restatement of the reference CPU solver (bicgstab_cpu.c:172-255 with the
stop test of bicgstab_gpu.c:148-150,231-236,289-293): iterations within
+-2, relative residual <= 1e-8 (1% slack), solution within 1e-6 relative.
"""
import numpy as np
import pytest

import pyoracle as orc
from conftest import golden
from hethyppar_amd import hbs
from parity import check as _check
from workloads import sample320

# Most systems or columns must meet the plain bar; at most this fraction may
# need the order-sensitive one (parity.check), and each that does is printed.
MAX_ORDER_SENSITIVE = 0.25

pytestmark = pytest.mark.gpu


def _systems(nsys, lo, hi, seed):
    sysl = hbs.gen_uneven_systems(nsys, lo, hi, seed)
    rp, c, v, rows = hbs.concat_systems([(a, b, d) for a, b, d, _ in sysl])
    return sysl, rp, c, v, rows


def test_uneven_systems_each_match_oracle(ctx):
    sysl, rp, c, v, rows = _systems(12, 1.5, 3.4, 5)
    assert np.any(np.diff(rows) % 4 != 0)          # systems start off the 16-byte grid
    n = int(rows[-1])
    xt = hbs.gen_uniform(n, -1.0, 1.0, 7)
    b = hbs.host_spmv(rp, c, v, xt)
    A = hbs.DeviceCSR(ctx, rp, c, v, sys_rows=rows)
    S = hbs.Solver(A)
    db, dx = ctx.to_device(b), ctx.empty(n)
    dx.zero()
    infos = S.bicgstab_systems(db, dx, tol=1e-8, maxit=2000)
    x = dx.download()
    its, loose = [], []
    for q, (srp, sc, sv, conv) in enumerate(sysl):
        r0, r1 = rows[q], rows[q + 1]
        Aq = orc.CSR(srp, sc, sv)
        xo, ito, _, _ = orc.bicgstab(Aq, b[r0:r1], tol=1e-8, maxit=2000)
        if _check(x[r0:r1], infos[q], xo, ito, f"system {q} n={r1 - r0} conv={conv:.2f}", Aq, b[r0:r1],
                  xt[r0:r1]) != "plain":
            loose.append(q)
        its.append(infos[q].iters)
    print(f"order-sensitive bar used by systems {loose} of {len(sysl)}", flush=True)
    assert len(loose) <= MAX_ORDER_SENSITIVE * len(sysl)
    assert len(set(its)) > 1                        # systems stopped at their own iteration


def test_uneven_systems_caps_and_zero_rhs(ctx):
    sysl, rp, c, v, rows = _systems(5, 1.5, 3.0, 11)
    n = int(rows[-1])
    xt = hbs.gen_uniform(n, -1.0, 1.0, 3)
    b = hbs.host_spmv(rp, c, v, xt)
    b[rows[2]:rows[3]] = 0.0                        # system 2: zero right-hand side
    A = hbs.DeviceCSR(ctx, rp, c, v, sys_rows=rows)
    S = hbs.Solver(A)
    # iteration cap between the smallest and the largest system's count
    full = []
    for q, (srp, sc, sv, _) in enumerate(sysl):
        _, ito, _, _ = orc.bicgstab(orc.CSR(srp, sc, sv), b[rows[q]:rows[q + 1]], tol=1e-8, maxit=2000)
        full.append(ito)
    cap = sorted(full)[2] + 3
    db, dx = ctx.to_device(b), ctx.empty(n)
    dx.zero()
    infos = S.bicgstab_systems(db, dx, tol=1e-8, maxit=cap, raise_on_error=False)
    x = dx.download()
    for q in range(5):
        print(f"system {q}: oracle {full[q]} gpu {infos[q].iters} stop {infos[q].stop} status {infos[q].status}",
              flush=True)
        if q == 2:
            assert infos[q].iters == 0 and infos[q].stop == 2 and infos[q].status == 0
            assert np.all(x[rows[2]:rows[3]] == 0.0)
        elif full[q] <= cap - 3:
            assert infos[q].status == 0 and abs(infos[q].iters - full[q]) <= 2
        elif full[q] > cap + 2:
            assert infos[q].status == hbs.HBS_ERR_NOCONV and infos[q].iters == cap
    # the aggregate call reports the worst system
    dx.zero()
    agg = S.bicgstab(db, dx, tol=1e-8, maxit=cap, raise_on_error=False)
    assert agg.status == (hbs.HBS_ERR_NOCONV if max(full) > cap + 2 else 0)
    assert agg.iters == max(i.iters for i in infos)


def test_systems_single_system_same_as_plain(ctx):
    rp, c, v = hbs.gen_convdiff2d(30, 30, 0.5)
    b = hbs.host_spmv(rp, c, v, hbs.gen_uniform(900, -1.0, 1.0, 1))
    A1 = hbs.DeviceCSR(ctx, rp, c, v)
    A2 = hbs.DeviceCSR(ctx, rp, c, v, sys_rows=[0, 900])
    xs = []
    for A in (A1, A2):
        S = hbs.Solver(A)
        db, dx = ctx.to_device(b), ctx.empty(900)
        dx.zero()
        info = S.bicgstab(db, dx, tol=1e-8)
        xs.append((dx.download(), info.iters))
    assert xs[0][1] == xs[1][1] and np.array_equal(xs[0][0], xs[1][0])


def test_systems_reject_cross_coupling(ctx):
    rp = np.array([0, 2, 3], np.int32)
    c = np.array([0, 1, 1], np.int32)
    v = np.ones(3)
    with pytest.raises(hbs.HBSError):
        hbs.DeviceCSR(ctx, rp, c, v, sys_rows=[0, 1, 2])


@pytest.mark.parametrize("k", [1, 8, 70])
def test_block_columns_match_oracle(ctx, k):
    rp, c, v = hbs.gen_convdiff2d(24, 20, 0.7)
    n = rp.size - 1
    Xt = hbs.gen_uniform(n * k, -1.0, 1.0, 40 + k).reshape(n, k)
    B = np.stack([hbs.host_spmv(rp, c, v, np.ascontiguousarray(Xt[:, j])) for j in range(k)], axis=1)
    B[:, k // 2] = 0.0                                     # one zero column
    A = hbs.DeviceCSR(ctx, rp, c, v)
    S = hbs.Solver(A)
    dB = ctx.to_device(np.ascontiguousarray(B).ravel())
    dX = ctx.to_device(np.zeros(n * k))
    infos = S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=1000)
    X = dX.download().reshape(n, k)
    loose = []
    for j in range(k):
        if j == k // 2:
            assert infos[j].iters == 0 and infos[j].stop == 2 and np.all(X[:, j] == 0.0)
            continue
        bj = np.ascontiguousarray(B[:, j])
        xo, ito, _, _ = orc.bicgstab(orc.CSR(rp, c, v), bj, tol=1e-8, maxit=1000)
        if _check(X[:, j], infos[j], xo, ito, f"column {j} of {k}", orc.CSR(rp, c, v), bj,
                  np.ascontiguousarray(Xt[:, j])) != "plain":
            loose.append(j)
    print(f"order-sensitive bar used by columns {loose} of {k}", flush=True)
    assert len(loose) <= MAX_ORDER_SENSITIVE * k


@pytest.mark.parametrize("tile,rb", [(64, 1), (64, 4), (32, 2), (16, 1), (8, 4)])
def test_block_forced_tiles_match_oracle(ctx, tile, rb):
    """Every column-group width and SpMM rows per lane step the autotuner
    can choose gives each column the oracle's solve."""
    rp, c, v = hbs.gen_convdiff2d(24, 20, 0.7)
    n, k = rp.size - 1, 40
    Xt = hbs.gen_uniform(n * k, -1.0, 1.0, 7).reshape(n, k)
    B = np.stack([hbs.host_spmv(rp, c, v, np.ascontiguousarray(Xt[:, j])) for j in range(k)], axis=1)
    A = hbs.DeviceCSR(ctx, rp, c, v)
    hbs.set_block_tuning(tile, rb)
    try:
        S = hbs.Solver(A)
        dB = ctx.to_device(np.ascontiguousarray(B).ravel())
        dX = ctx.to_device(np.zeros(n * k))
        infos = S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=1000)
        assert S.block_tuning(k) == (min(tile, 64), rb, (64 // tile) * rb)
    finally:
        hbs.set_block_tuning(0, 0)
    X = dX.download().reshape(n, k)
    Ao = orc.CSR(rp, c, v)
    loose = []
    for j in range(k):
        bj = np.ascontiguousarray(B[:, j])
        xo, ito, _, _ = orc.bicgstab(Ao, bj, tol=1e-8, maxit=1000)
        if _check(X[:, j], infos[j], xo, ito, f"tile {tile} rb {rb} column {j}", Ao, bj,
                  np.ascontiguousarray(Xt[:, j])) != "plain":
            loose.append(j)
    assert len(loose) <= MAX_ORDER_SENSITIVE * k


@pytest.mark.parametrize("k", [64, 40, 100, 80])
@pytest.mark.parametrize("wide", [1, 0])
def test_block_wide_gathers_match_oracle(ctx, k, wide):
    """64-column groups with X gathered in 16-byte pieces (two columns per
    lane, two rows per step; k = 40: lanes past the group's columns; 100:
    a second group of 36 columns; 80: a 16-column remainder group in the
    one-column-per-lane form) and, with wide loads off, the one-column
    form: each column gets the oracle's solve.  The columns stop at
    different iterations, and a zero column is done at once, so a lane's
    two columns are not always both active."""
    rp, c, v = hbs.gen_convdiff2d(40, 36, 0.6)
    n = rp.size - 1
    Xt = hbs.gen_uniform(n * k, -1.0, 1.0, 11 + k).reshape(n, k)
    B = np.stack([hbs.host_spmv(rp, c, v, np.ascontiguousarray(Xt[:, j])) for j in range(k)], axis=1)
    B[:, 5] = 0.0
    A = hbs.DeviceCSR(ctx, rp, c, v)
    hbs.set_block_tuning(64, 0)
    hbs.set_spmm_tuning(0, 0, 16, wide)
    try:
        S = hbs.Solver(A)
        dB = ctx.to_device(np.ascontiguousarray(B).ravel())
        dX = ctx.to_device(np.zeros(n * k))
        infos = S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=1000)
        tile, rb, rpw = S.block_tuning(k)
        assert tile == 64
        if wide:
            assert (rb, rpw) == (2, 2)
        else:
            assert rb in (1, 2, 4) and rpw == rb
    finally:
        hbs.set_block_tuning(0, 0)
        hbs.set_spmm_tuning()
    X = dX.download().reshape(n, k)
    Ao = orc.CSR(rp, c, v)
    assert infos[5].iters == 0 and infos[5].stop == 2 and np.all(X[:, 5] == 0.0)
    loose = []
    for j in range(k):
        if j == 5:
            continue
        bj = np.ascontiguousarray(B[:, j])
        xo, ito, _, _ = orc.bicgstab(Ao, bj, tol=1e-8, maxit=1000)
        if _check(X[:, j], infos[j], xo, ito, f"k {k} wide {wide} column {j}", Ao, bj,
                  np.ascontiguousarray(Xt[:, j])) != "plain":
            loose.append(j)
    print(f"order-sensitive bar used by columns {loose} of {k}", flush=True)
    assert len(loose) <= MAX_ORDER_SENSITIVE * k


def test_block_wide_gathers_unaligned_rhs(ctx):
    """A caller's B that is only 8-byte aligned (one double into an
    allocation): the 16-byte SpMM form cannot read it, so the residual
    SpMMs against it take the one-column-per-lane form; the solve still
    meets the oracle's bar on every column."""
    import types
    rp, c, v = hbs.gen_convdiff2d(40, 36, 0.6)
    n, k = rp.size - 1, 64
    Xt = hbs.gen_uniform(n * k, -1.0, 1.0, 77).reshape(n, k)
    B = np.stack([hbs.host_spmv(rp, c, v, np.ascontiguousarray(Xt[:, j])) for j in range(k)], axis=1)
    A = hbs.DeviceCSR(ctx, rp, c, v)
    dBbuf = ctx.to_device(np.concatenate([[0.0], np.ascontiguousarray(B).ravel()]))
    dB = types.SimpleNamespace(ptr=dBbuf.offset(1))
    dX = ctx.to_device(np.zeros(n * k))
    hbs.set_block_tuning(64, 0)
    try:
        S = hbs.Solver(A)
        infos = S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=1000)
        assert S.block_tuning(k) == (64, 2, 2)
    finally:
        hbs.set_block_tuning(0, 0)
    X = dX.download().reshape(n, k)
    Ao = orc.CSR(rp, c, v)
    loose = []
    for j in range(k):
        bj = np.ascontiguousarray(B[:, j])
        xo, ito, _, _ = orc.bicgstab(Ao, bj, tol=1e-8, maxit=1000)
        if _check(X[:, j], infos[j], xo, ito, f"unaligned B column {j}", Ao, bj,
                  np.ascontiguousarray(Xt[:, j])) != "plain":
            loose.append(j)
    assert len(loose) <= MAX_ORDER_SENSITIVE * k


@pytest.mark.parametrize("sgroups", [512, 1024])
def test_block_spmm_groups_forced_match_oracle(ctx, sgroups):
    """The block solve's SpMMs on fewer workgroups than its streaming
    kernels (their partials reduced over their own count) give each column
    the oracle's solve."""
    rp, c, v = hbs.gen_convdiff2d(128, 128, 0.6)
    n, k = rp.size - 1, 64
    Xt = hbs.gen_uniform(n * k, -1.0, 1.0, 9).reshape(n, k)
    B = np.stack([hbs.host_spmv(rp, c, v, np.ascontiguousarray(Xt[:, j])) for j in range(k)], axis=1)
    A = hbs.DeviceCSR(ctx, rp, c, v)
    hbs.set_block_tuning(64, 4)
    hbs.set_block_spmm_groups(sgroups)
    try:
        S = hbs.Solver(A)
        dB = ctx.to_device(np.ascontiguousarray(B).ravel())
        dX = ctx.to_device(np.zeros(n * k))
        infos = S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=2000)
        assert S.block_spmm_groups(k) == sgroups
    finally:
        hbs.set_block_tuning(0, 0)
        hbs.set_block_spmm_groups(0)
    X = dX.download().reshape(n, k)
    Ao = orc.CSR(rp, c, v)
    loose = []
    for j in range(0, k, 4):
        bj = np.ascontiguousarray(B[:, j])
        xo, ito, _, _ = orc.bicgstab(Ao, bj, tol=1e-8, maxit=2000)
        if _check(X[:, j], infos[j], xo, ito, f"spmm groups {sgroups} column {j}", Ao, bj,
                  np.ascontiguousarray(Xt[:, j])) != "plain":
            loose.append(j)
    assert len(loose) <= MAX_ORDER_SENSITIVE * (k // 4)
    for j in range(k):
        assert infos[j].status == 0 and infos[j].relres <= 1.01e-8


def test_block_fixed_dealing_reruns_bit_identical(ctx):
    """With the 64-column sweep's steps at a fixed stride per wave (no step
    counters), each workgroup sums the same rows every time, so the
    per-column dots -- and the whole block solve -- repeat bit for bit; with
    steps dealt from counters the rows a workgroup sums vary run to run."""
    rp, c, v = hbs.gen_convdiff3d(128, 128, 4, 0.5)
    n, k = rp.size - 1, 64
    B = hbs.gen_uniform(n * k, -1.0, 1.0, 21)
    hbs.set_spmm_tuning(0, 0, 0, 1)
    try:
        A = hbs.DeviceCSR(ctx, rp, c, v)
    finally:
        hbs.set_spmm_tuning()
    assert A.info()["spmm_runs"] > 0
    S = hbs.Solver(A)
    out = []
    for _ in range(2):
        dB, dX = ctx.to_device(B), ctx.to_device(np.zeros(n * k))
        infos = S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=2000)
        assert all(i.status == 0 and i.relres <= 1.01e-8 for i in infos)
        out.append((dX.download(), [i.iters for i in infos]))
    assert out[0][1] == out[1][1] and np.array_equal(out[0][0], out[1][0])


def test_block_autotuner_choice(ctx):
    """The tuner's choice is one of its candidates and fixed per workspace."""
    rp, c, v = hbs.gen_convdiff2d(64, 64, 0.5)
    n, k = rp.size - 1, 24
    A = hbs.DeviceCSR(ctx, rp, c, v)
    S = hbs.Solver(A)
    dB = ctx.to_device(hbs.gen_uniform(n * k, -1.0, 1.0, 3))
    dX = ctx.to_device(np.zeros(n * k))
    S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=1000)
    tile, rb, rpw = S.block_tuning(k)
    assert tile in (32, 16, 8) and rb in (1, 2, 4) and rpw == (64 // tile) * rb
    g = S.block_spmm_groups(k)
    assert g > 0 and g % 8 == 0
    dX.zero()
    S.bicgstab_block(dB, dX, k, tol=1e-8, maxit=1000)
    assert S.block_tuning(k) == (tile, rb, rpw)
    with pytest.raises(hbs.HBSError):
        S.block_tuning(k + 1)


def test_block_fixed_iterations_golden(ctx):
    rp, c, v = sample320()
    b = hbs.vector_read(golden("sample320", "B.txt"), 320)
    x0 = hbs.vector_read(golden("sample320", "X_init.txt"), 320)
    k = 3
    A = hbs.DeviceCSR(ctx, rp, c, v)
    S = hbs.Solver(A)
    dB = ctx.to_device(np.repeat(b[:, None], k, axis=1).ravel())
    dX = ctx.to_device(np.repeat(x0[:, None], k, axis=1).ravel())
    infos = S.bicgstab_block(dB, dX, k, maxit=20, fixed_iters=1)
    X = dX.download().reshape(320, k)
    xo, _, _, _ = orc.bicgstab(orc.CSR(rp, c, v), b, x0, maxit=20, fixed=True)
    gold = hbs.vector_read(golden("sample320", "out_X_seq.txt"), 320)
    for j in range(k):
        assert infos[j].iters == 20
        assert np.max(np.abs(X[:, j] - xo)) < 1e-9
        assert np.max(np.abs(X[:, j] - gold)) <= 1e-6 + 1e-9


def test_block_graph_and_eager_agree(ctx):
    rp, c, v = hbs.gen_convdiff2d(16, 16, 0.3)
    n, k = 256, 5
    B = hbs.gen_uniform(n * k, -1.0, 1.0, 9)
    A = hbs.DeviceCSR(ctx, rp, c, v)
    S = hbs.Solver(A)
    out = []
    for graph, every in ((1, 4), (0, 1), (1, 8)):
        dB, dX = ctx.to_device(B), ctx.to_device(np.zeros(n * k))
        infos = S.bicgstab_block(dB, dX, k, tol=1e-10, use_graph=graph, check_every=every)
        out.append((dX.download(), [i.iters for i in infos]))
    for x, its in out[1:]:
        assert its == out[0][1] and np.array_equal(x, out[0][0])
