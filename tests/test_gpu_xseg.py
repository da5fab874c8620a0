"""x-segmented SpMV (k_xseg + k_spmv<PRE>) on the MI355X against the oracle.

Matrices whose columns spread over all of x are multiplied in two passes:
products formed per 8192-column x segment staged in LDS, then summed per
row in CSR order.  Bars, as for the CSR-stream SpMV: rows of serial row
blocks (every row of at most 256 entries here) equal the oracle's
CSR_spmxv_seq bit for bit; chunked long rows lie within the summation bound;
FP32 within its bound.  The solver on such a matrix meets the BiCGStab bar
(iterations +-2, relres <= 1e-8, solution within 1e-6).
"""
import numpy as np
import pytest

import pyoracle as orc
from hethyppar_amd import hbs
from workloads import markov, random_csr, sample320, seq_bound

pytestmark = pytest.mark.gpu


def _forced(ctx, rp, c, v, ncols=None, flags=0, split=1):
    hbs.set_spmv_xseg(1)
    hbs.set_xseg_split(split)
    try:
        return hbs.DeviceCSR(ctx, rp, c, v, ncols=ncols, flags=flags)
    finally:
        hbs.set_spmv_xseg()
        hbs.set_xseg_split()


def _check_f64(rp, c, v, x, y, serial_max=256):
    ref = orc.spmv(orc.CSR(rp, c, v, ncols=x.size), x)
    k = np.diff(rp)
    serial = k <= serial_max
    assert np.array_equal(y[serial], ref[serial])
    assert np.all(np.abs(y - ref) <= seq_bound(rp, c, v, x) + 1e-300)


def _segment_counts_csr(counts, m, seed):
    """m-row CSR whose column segment s (columns [8192 s, 8192 (s + 1)))
    holds exactly counts[s] entries, at distinct random positions."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for s, k in enumerate(counts):
        flat = rng.choice(m * 8192, size=k, replace=False)
        rows.append(flat // 8192)
        cols.append(s * 8192 + flat % 8192)
    r, c = np.concatenate(rows), np.concatenate(cols)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    rp = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=m))]).astype(np.int32)
    return rp, c.astype(np.int32), rng.standard_normal(c.size)


def _cases():
    out = []
    # segments of exactly one k_xseg pass (12 x 512 FP64, 16 x 512 FP32
    # entries), one more, one fewer, two passes, a single entry and none
    counts = [6144, 8192, 6145, 8191, 12288, 1, 0]
    out.append(("segment_passes", *_segment_counts_csr(counts, 3000, 17), 8192 * len(counts)))
    out.append(("circuit60k", *hbs.gen_circuit(60000, nheavy=12, heavy_lo=300, heavy_hi=30000, seed=11), None))
    out.append(("markov50k", *markov(50000), None))
    out.append(("random_rect", *random_csr(3000, 20001, 0.002, 5), 20001))   # empty rows, n % 8192 != 0
    out.append(("sample320", *sample320(), None))                            # one segment
    return out


@pytest.mark.parametrize("split", [1, 3], ids=lambda s: f"split{s}")
@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_xseg_spmv_bit_exact(ctx, case, split):
    """split: workgroups per segment in k_xseg (shares of a segment's
    entries that are not whole passes, a share without entries)."""
    name, rp, c, v, ncols = case
    n = ncols or rp.size - 1
    m = rp.size - 1
    x = np.random.default_rng(3).uniform(-1, 1, n)
    A = _forced(ctx, rp, c, v, ncols=ncols, flags=hbs.HBS_CSR_F32, split=split)
    info = A.info()
    print(name, info, flush=True)
    assert info["xseg"] == (n + 8191) // 8192
    dx, dy = ctx.to_device(x), ctx.empty(m)
    A.spmv(dx, dy)
    _check_f64(rp, c, v, x, dy.download())
    # again: the products are formed afresh each time (new x)
    x2 = x[::-1].copy()
    dx.upload(x2)
    A.spmv(dx, dy)
    _check_f64(rp, c, v, x2, dy.download())
    # FP32 copy
    x32 = x.astype(np.float32)
    dx32, dy32 = ctx.to_device(x32), ctx.empty(m, np.float32)
    A.spmv(dx32, dy32)
    y32 = dy32.download().astype(np.float64)
    ref32 = orc.spmv(orc.CSR(rp, c, v.astype(np.float32).astype(np.float64), ncols=n), x32.astype(np.float64))
    u = np.finfo(np.float32).eps / 2
    k = np.diff(rp)
    bound = 2 * (2 * k * u / (1 - k * u)) * orc.spmv(orc.CSR(rp, c, np.abs(v), ncols=n), np.abs(x))
    assert np.all(np.abs(y32 - ref32) <= bound * 1.01 + 1e-30)


def test_xseg_two_pass_blocks_in_row_order(ctx):
    """Row blocks of 16384 entries take two product passes each; the rows of
    power-law blocks are summed in their rperm order, which the block's
    first pass loads and its last pass uses: FP64 bit-exact on serial rows,
    FP32 (diagonal apart) within its bound."""
    rp, c, v = hbs.gen_circuit(60000, nheavy=12, heavy_lo=300, heavy_hi=30000, seed=11)
    n = rp.size - 1
    x = np.random.default_rng(23).uniform(-1, 1, n)
    hbs.set_spmv_plan(block_nnz=16384)
    try:
        A = _forced(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    finally:
        hbs.set_spmv_plan()
    assert A.info()["xseg"] > 0
    dx, dy = ctx.to_device(x), ctx.empty(n)
    A.spmv(dx, dy)
    _check_f64(rp, c, v, x, dy.download())
    x32 = x.astype(np.float32)
    dx32, dy32 = ctx.to_device(x32), ctx.empty(n, np.float32)
    A.spmv(dx32, dy32)
    ref32 = orc.spmv(orc.CSR(rp, c, v.astype(np.float32).astype(np.float64)), x32.astype(np.float64))
    u = np.finfo(np.float32).eps / 2
    k = np.diff(rp)
    bound = 2 * (2 * k * u / (1 - k * u)) * orc.spmv(orc.CSR(rp, c, np.abs(v)), np.abs(x))
    assert np.all(np.abs(dy32.download().astype(np.float64) - ref32) <= bound * 1.01 + 1e-30)


def test_xseg_alpha_beta(ctx):
    rp, c, v = random_csr(5000, 5000, 0.003, 21)
    x = np.random.default_rng(4).standard_normal(5000)
    y0 = np.random.default_rng(5).standard_normal(5000)
    A = _forced(ctx, rp, c, v)
    dx, dy = ctx.to_device(x), ctx.to_device(y0)
    A.spmv(dx, dy, alpha=-2.0, beta=0.5)
    ref = orc.spmv(orc.CSR(rp, c, v), x)
    assert np.allclose(dy.download(), -2.0 * ref + 0.5 * y0, rtol=1e-14, atol=1e-14)


def test_xseg_auto_takes_scattered_columns_only(ctx):
    """Default mode: the 1.2M-row circuit matrix (uniform random columns) is
    x-segmented, the banded Markov matrix of the same size is not; both give
    the oracle's product."""
    rp, c, v = hbs.gen_circuit(1200000, nheavy=40, heavy_lo=1000, heavy_hi=20000, seed=5)
    n = rp.size - 1
    x = np.random.default_rng(6).uniform(-1, 1, n)
    A = hbs.DeviceCSR(ctx, rp, c, v)
    assert A.info()["xseg"] == (n + 8191) // 8192
    dx, dy = ctx.to_device(x), ctx.empty(n)
    A.spmv(dx, dy)
    _check_f64(rp, c, v, x, dy.download())
    A.close()
    rp, c, v = markov(1200000)
    B = hbs.DeviceCSR(ctx, rp, c, v)
    assert B.info()["xseg"] == 0
    B.close()


def test_xseg_set_values_and_transpose(ctx):
    """New values reach the segment-major copies (FP64 and FP32) of A and A^T."""
    rp, c, v = random_csr(4000, 4000, 0.004, 8, empty_rows=False)
    n = rp.size - 1
    A = _forced(ctx, rp, c, v, flags=hbs.HBS_CSR_F32 | hbs.HBS_CSR_TRANSPOSE)
    v2 = v * 1.5 + 0.25
    A.set_values(ctx.to_device(v2))
    x = np.random.default_rng(9).uniform(-1, 1, n)
    dx, dy = ctx.to_device(x), ctx.empty(n)
    A.spmv(dx, dy)
    assert np.array_equal(dy.download(), orc.spmv(orc.CSR(rp, c, v2), x))
    A.spmv_t(dx, dy)
    tp, tc, tv = orc.csr_to_csc(orc.CSR(rp, c, v2))
    assert np.array_equal(dy.download(), orc.spmv(orc.CSR(tp, tc, tv), x))
    dx32, dy32 = ctx.to_device(x.astype(np.float32)), ctx.empty(n, np.float32)
    A.spmv(dx32, dy32)
    ref = orc.spmv(orc.CSR(rp, c, v2), x.astype(np.float32).astype(np.float64))
    assert np.allclose(dy32.download(), ref, rtol=1e-5, atol=1e-5)


def _diag_cases():
    rp, c, v = hbs.gen_circuit(60000, nheavy=12, heavy_lo=300, heavy_hi=30000, seed=7)
    out = [("unit", rp, c, v, 1)]
    rp, c, v = hbs.gen_circuit(60000, nheavy=12, heavy_lo=300, heavy_hi=30000, seed=7, equilibrate=False)
    out.append(("values", rp, c, v, 2))
    # rows without a diagonal entry, and rows with two (the first is the
    # diagonal, the second stays an off-diagonal entry)
    rp2, c2, v2 = [0], [], []
    rng = np.random.default_rng(12)
    for r in range(rp.size - 1):
        cols, vals = list(c[rp[r]:rp[r + 1]]), list(v[rp[r]:rp[r + 1]])
        if r % 7 == 3:
            keep = [k for k in range(len(cols)) if cols[k] != r]
            cols, vals = [cols[k] for k in keep], [vals[k] for k in keep]
        elif r % 11 == 5:
            cols.append(r)
            vals.append(rng.standard_normal())
        c2 += cols
        v2 += vals
        rp2.append(len(c2))
    out.append(("missing_and_twice", np.array(rp2, np.int32), np.array(c2, np.int32), np.array(v2), 2))
    return out


@pytest.mark.parametrize("case", _diag_cases(), ids=lambda c: c[0])
def test_xseg_fp32_diagonal_apart(ctx, case):
    """FP32 SpMVs of a square x-segmented matrix keep the diagonal apart:
    the off-diagonal sum plus d_i x_i (x_i for a unit diagonal) lies within
    the FP32 summation bound of the oracle's product of the FP32 values; the
    FP64 SpMV of the same matrix stays bit-exact; new values (a diagonal no
    longer 1) reach both parts."""
    name, rp, c, v, mode = case
    n = rp.size - 1
    x = np.random.default_rng(31).uniform(-1, 1, n)
    A = _forced(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    info = A.info()
    print(name, info, flush=True)
    assert info["xseg"] > 0 and info["diag_split"] == mode
    dx, dy = ctx.to_device(x), ctx.empty(n)
    A.spmv(dx, dy)
    _check_f64(rp, c, v, x, dy.download())
    u = np.finfo(np.float32).eps / 2
    k = np.diff(rp)

    def check32(vals):
        x32 = x.astype(np.float32)
        dx32, dy32 = ctx.to_device(x32), ctx.empty(n, np.float32)
        A.spmv(dx32, dy32)
        y32 = dy32.download().astype(np.float64)
        ref = orc.spmv(orc.CSR(rp, c, vals.astype(np.float32).astype(np.float64)), x32.astype(np.float64))
        bound = 2 * (2 * k * u / (1 - k * u)) * orc.spmv(orc.CSR(rp, c, np.abs(vals)), np.abs(x))
        assert np.all(np.abs(y32 - ref) <= bound * 1.01 + 1e-30)

    check32(v)
    v2 = v * 0.5 - 0.125
    A.set_values(ctx.to_device(v2))
    assert A.info()["diag_split"] == 2
    check32(v2)
    A.spmv(dx, dy)
    assert np.array_equal(dy.download()[k <= 256], orc.spmv(orc.CSR(rp, c, v2), x)[k <= 256])


def test_xseg_fp32_diagonal_apart_off_matches_on(ctx):
    """The same mixed solve with the diagonal apart and among the entries:
    both meet the bar (relres <= 1e-8, forward error no worse than twice the
    oracle's); the split takes no more FP32 iterations than two over the
    other's."""
    rp, c, v = hbs.gen_circuit(150000, nheavy=10, heavy_lo=1000, heavy_hi=20000, seed=3)
    n = rp.size - 1
    xt = hbs.gen_uniform(n, -1.0, 1.0, 13)
    b = hbs.host_spmv(rp, c, v, xt)
    xo, ito, _, _ = orc.bicgstab(orc.CSR(rp, c, v), b, tol=1e-8, maxit=2000)
    e_o = np.linalg.norm(xo - xt) / np.linalg.norm(xt)
    res = {}
    for split in (1, 0):
        hbs.set_spmv_diag_split(split)
        try:
            A = _forced(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
        finally:
            hbs.set_spmv_diag_split()
        assert A.info()["diag_split"] == split
        S = hbs.Solver(A)
        db, dx = ctx.to_device(b), ctx.empty(n)
        dx.zero()
        im = S.bicgstab_mixed(db, dx, tol=1e-8, maxit=4000)
        e_m = np.linalg.norm(dx.download() - xt) / np.linalg.norm(xt)
        print(f"split {split}: iters {im.iters} sweeps {im.sweeps} relres {im.relres:.3e} err {e_m:.2e} "
              f"(oracle {ito}, {e_o:.2e})", flush=True)
        assert im.status == 0 and im.relres <= 1.01e-8
        assert e_m <= 2 * max(e_o, 1e-7)
        res[split] = im.iters
        S.close()
        A.close()
    assert res[1] <= res[0] + 2


@pytest.mark.parametrize("graph", [0, 1])
def test_xseg_fused_updates_same_iterates(ctx, graph):
    """With P and S updated inside the x-segmented product passes the FP64
    and the mixed solves take the same iterates as with the separate vector
    passes (the same arithmetic; only ||s||^2, for the ||s|| test, is summed
    over other partitions): bit-identical x, same iteration counts."""
    rp, c, v = hbs.gen_circuit(150000, nheavy=10, heavy_lo=1000, heavy_hi=20000, seed=5)
    n = rp.size - 1
    xt = hbs.gen_uniform(n, -1.0, 1.0, 17)
    b = hbs.host_spmv(rp, c, v, xt)
    A = _forced(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    out = {}
    for fuse in (1, 0):
        hbs.set_xseg_fuse(fuse)
        try:
            S = hbs.Solver(A)
        finally:
            hbs.set_xseg_fuse()
        db, dx = ctx.to_device(b), ctx.empty(n)
        dx.zero()
        i64 = S.bicgstab(db, dx, tol=1e-8, maxit=2000, use_graph=graph)
        x64 = dx.download()
        dx.zero()
        im = S.bicgstab_mixed(db, dx, tol=1e-8, maxit=4000, use_graph=graph)
        out[fuse] = (x64, i64.iters, i64.relres, dx.download(), im.iters, im.sweeps)
        print(f"fuse {fuse}: fp64 {i64.iters} its relres {i64.relres:.3e}; mixed {im.iters} its "
              f"{im.sweeps} sweeps relres {im.relres:.3e}", flush=True)
        assert i64.status == 0 and im.status == 0
        S.close()
    f, u = out[1], out[0]
    assert f[1:3] == u[1:3] and np.array_equal(f[0], u[0])
    assert f[4:] == u[4:] and np.array_equal(f[3], u[3])


@pytest.mark.parametrize("graph", [0, 1])
def test_xseg_bicgstab_fp64_and_mixed(ctx, graph):
    rp, c, v = hbs.gen_circuit(150000, nheavy=10, heavy_lo=1000, heavy_hi=20000, seed=3)
    n = rp.size - 1
    xt = hbs.gen_uniform(n, -1.0, 1.0, 13)
    b = hbs.host_spmv(rp, c, v, xt)
    xo, ito, _, _ = orc.bicgstab(orc.CSR(rp, c, v), b, tol=1e-8, maxit=2000)
    A = _forced(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    assert A.info()["xseg"] > 0
    S = hbs.Solver(A)
    db = ctx.to_device(b)
    dx = ctx.empty(n)
    dx.zero()
    info = S.bicgstab(db, dx, tol=1e-8, maxit=2000, use_graph=graph)
    x = dx.download()
    print(f"xseg gpu iters {info.iters} relres {info.relres:.3e} | oracle iters {ito}", flush=True)
    assert info.status == 0 and abs(info.iters - ito) <= 2 and info.relres <= 1.01e-8
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6
    dx.zero()
    im = S.bicgstab_mixed(db, dx, tol=1e-8, maxit=4000, use_graph=graph)
    xm = dx.download()
    e_m = np.linalg.norm(xm - xt) / np.linalg.norm(xt)
    e_o = np.linalg.norm(xo - xt) / np.linalg.norm(xt)
    print(f"xseg mixed iters {im.iters} sweeps {im.sweeps} relres {im.relres:.3e} err {e_m:.2e} ({e_o:.2e})",
          flush=True)
    assert im.status == 0 and im.relres <= 1.01e-8
    assert e_m <= 2 * max(e_o, 1e-7)
