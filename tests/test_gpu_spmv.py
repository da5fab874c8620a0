"""SpMV / SpMM on the MI355X through the C-ABI against the oracle.

Bar: rows of order-exact (serial) row blocks must equal the oracle's
CSR_spmxv_seq bit for bit; rows summed by several lanes or split over chunks
must lie within the summation bound 2*gamma_k*sum|a_ij x_j| of the oracle.
FP32: |y32 - y64| <= 2*gamma32_k * sum|a x| (+ input rounding of a and x).
"""
import numpy as np
import pytest

import pyoracle as orc
from conftest import golden
from hethyppar_amd import hbs
from workloads import convdiff64, markov, random_csr, sample, sample320, seq_bound

pytestmark = pytest.mark.gpu


def _run_spmv(ctx, rp, c, v, x, ncols=None, flags=0):
    A = hbs.DeviceCSR(ctx, rp, c, v, ncols=ncols, flags=flags)
    dx = ctx.to_device(x)
    dy = ctx.empty(rp.size - 1)
    A.spmv(dx, dy)
    y = dy.download()
    return A, y


def _cases():
    rng = np.random.default_rng(7)
    out = [("sample", *sample()), ("sample320", *sample320())]
    rp, c, v, xt, b = convdiff64()
    out.append(("convdiff64", rp, c, v))
    out.append(("markov50k", *markov(50000)))
    out.append(("random", *random_csr(3000, 3000, 0.004, 3)))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_spmv_bit_exact_serial_blocks(ctx, case):
    name, rp, c, v = case
    n = rp.size - 1
    x = np.random.default_rng(1).uniform(-1, 1, n)
    A, y = _run_spmv(ctx, rp, c, v, x)
    info = A.info()
    ref = orc.spmv(orc.CSR(rp, c, v), x)
    print(name, info, flush=True)
    assert info["serial_rows"] == n
    assert np.array_equal(y, ref)


def test_spmv_kat_x_mul(ctx):
    rp, c, v = sample()
    x = hbs.vector_read(golden("sample", "X_mul.txt"), 16)
    _, y = _run_spmv(ctx, rp, c, v, x)
    assert np.max(np.abs(y - np.arange(1, 17))) < 1e-13


@pytest.mark.parametrize("block_nnz,serial", [(256, 0), (512, 4), (2048, 64), (4096, 1000)])
def test_spmv_plans_power_law(ctx, block_nnz, serial):
    """Power-law rows: chunked long rows, lane-parallel blocks, serial blocks."""
    rp, c, v = hbs.gen_circuit(60000, nheavy=12, heavy_lo=300, heavy_hi=9000, seed=11)
    x = np.random.default_rng(2).uniform(-1, 1, rp.size - 1)
    hbs.set_spmv_plan(block_nnz, serial)
    try:
        A, y = _run_spmv(ctx, rp, c, v, x)
    finally:
        hbs.set_spmv_plan()
    ref = orc.spmv(orc.CSR(rp, c, v), x)
    bound = seq_bound(rp, c, v, x)
    assert np.all(np.abs(y - ref) <= bound + 1e-300)


@pytest.mark.parametrize("block_nnz,which", [(b, w) for b in (6144, 8192, 16384) for w in ("markov50k", "circuit60k")]
                         + [(16384, "markov600k")])
def test_spmv_wide_blocks_bit_exact(ctx, block_nnz, which):
    """512-thread row blocks, several passes of products into LDS (column-sorted
    where it pays): serial rows stay bit-exact in FP64 and within the FP32
    bound, long rows within the summation bound.  markov600k's column (20
    bits) and in-block position (14 bits) do not fit one word, so its sorted
    blocks read the separate position stream; the others pack them."""
    if which == "markov50k":
        rp, c, v = markov(50000)
    elif which == "markov600k":
        rp, c, v = markov(600000)
    else:
        rp, c, v = hbs.gen_circuit(60000, nheavy=12, heavy_lo=300, heavy_hi=30000, seed=11)
    n = rp.size - 1
    x = np.random.default_rng(8).uniform(-1, 1, n)
    hbs.set_spmv_plan(block_nnz, 1 << 20)   # every block that fits sums serially
    try:
        A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
        dx, dy = ctx.to_device(x), ctx.empty(n)
        A.spmv(dx, dy)
        y = dy.download()
        x32 = x.astype(np.float32)
        dx32, dy32 = ctx.to_device(x32), ctx.empty(n, np.float32)
        A.spmv(dx32, dy32)
        y32 = dy32.download().astype(np.float64)
        info = A.info()
    finally:
        hbs.set_spmv_plan()
    print(which, block_nnz, info, flush=True)
    assert info["block_nnz"] == block_nnz
    ref = orc.spmv(orc.CSR(rp, c, v), x)
    k = np.diff(rp)
    serial = k <= min(block_nnz, 256)   # longer rows are summed in chunks
    assert np.array_equal(y[serial], ref[serial])
    assert np.all(np.abs(y - ref) <= seq_bound(rp, c, v, x) + 1e-300)
    ref32 = orc.spmv(orc.CSR(rp, c, v.astype(np.float32).astype(np.float64)), x32.astype(np.float64))
    u = np.finfo(np.float32).eps / 2
    bound = 2 * (2 * k * u / (1 - k * u)) * orc.spmv(orc.CSR(rp, c, np.abs(v)), np.abs(x))
    assert np.all(np.abs(y32 - ref32) <= bound * 1.01 + 1e-30)


def test_spmv_lds_limit_kept_for_earlier_plans(ctx):
    """A plan with 16384-entry blocks (132 KiB of LDS per workgroup), then
    one with 8192: the first matrix's SpMV still launches (the kernels' LDS
    limits only grow)."""
    rp, c, v = markov(50000)
    n = rp.size - 1
    x = np.random.default_rng(9).uniform(-1, 1, n)
    try:
        hbs.set_spmv_plan(16384, 1 << 20)
        A1 = hbs.DeviceCSR(ctx, rp, c, v)
        hbs.set_spmv_plan(8192, 1 << 20)
        A2 = hbs.DeviceCSR(ctx, rp, c, v)
    finally:
        hbs.set_spmv_plan()
    ref = orc.spmv(orc.CSR(rp, c, v), x)
    dx, dy = ctx.to_device(x), ctx.empty(n)
    for A in (A1, A2, A1):
        A.spmv(dx, dy)
        assert np.array_equal(dy.download(), ref)


def test_spmv_alpha_beta_and_empty_rows(ctx):
    rp, c, v = random_csr(500, 400, 0.01, 9)   # has empty rows
    assert (np.diff(rp) == 0).any()
    x = np.random.default_rng(3).standard_normal(400)
    y0 = np.random.default_rng(4).standard_normal(500)
    A = hbs.DeviceCSR(ctx, rp, c, v, ncols=400)
    dx, dy = ctx.to_device(x), ctx.to_device(y0)
    A.spmv(dx, dy, alpha=-2.0, beta=0.5)
    ref = orc.spmv(orc.CSR(rp, c, v, ncols=400), x)
    assert np.allclose(dy.download(), -2.0 * ref + 0.5 * y0, rtol=1e-14, atol=1e-14)


def test_spmv_f32(ctx):
    rp, c, v = markov(30000)
    x = np.random.default_rng(5).uniform(-1, 1, rp.size - 1)
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    x32 = x.astype(np.float32)
    dx, dy = ctx.to_device(x32), ctx.empty(rp.size - 1, np.float32)
    A.spmv(dx, dy)
    y32 = dy.download().astype(np.float64)
    ref = orc.spmv(orc.CSR(rp, c, v.astype(np.float32).astype(np.float64)), x32.astype(np.float64))
    u = np.finfo(np.float32).eps / 2
    k = np.diff(rp)
    bound = 2 * (2 * k * u / (1 - k * u)) * orc.spmv(orc.CSR(rp, c, np.abs(v)), np.abs(x))
    assert np.all(np.abs(y32 - ref) <= bound * 1.01 + 1e-30)


def test_spmv_transpose(ctx):
    rp, c, v = random_csr(700, 700, 0.01, 12, empty_rows=False)
    x = np.random.default_rng(6).standard_normal(700)
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_TRANSPOSE)
    dx, dy = ctx.to_device(x), ctx.empty(700)
    A.spmv_t(dx, dy)
    import scipy.sparse as sp
    AT = sp.csr_matrix((v, c, rp), shape=(700, 700)).T.tocsr()
    AT.sort_indices()
    ref = orc.spmv(orc.CSR(AT.indptr, AT.indices, AT.data), x)
    assert np.array_equal(dy.download(), ref)


@pytest.mark.parametrize("k,tile", [(64, 0), (64, 64), (64, 16), (8, 0), (5, 4), (130, 64)])
def test_spmm_columns_bit_exact(ctx, k, tile):
    rp, c, v, xt, b = convdiff64()
    n = rp.size - 1
    X = np.random.default_rng(k).uniform(-1, 1, (n, k))
    A = hbs.DeviceCSR(ctx, rp, c, v)
    dX, dY = ctx.to_device(X.ravel()), ctx.empty(n * k)
    Ao = orc.CSR(rp, c, v)
    # twice: the second launch starts from the step counters the first left
    for rep in range(2):
        dY.zero()
        A.spmm(dX, dY, k, tile)
        Y = dY.download().reshape(n, k)
        for j in range(k):
            assert np.array_equal(Y[:, j], orc.spmv(Ao, X[:, j])), f"launch {rep} column {j}"


def test_spmm_tiny_matrix_few_waves(ctx):
    """A 25-row matrix: the 64-column SpMM runs 8 workgroups, 4 waves per
    XCD, fewer than an XCD's step counters; every row is still formed,
    launch after launch."""
    rp, c, v = hbs.gen_convdiff2d(5, 5, 0.5)
    n, k = rp.size - 1, 64
    X = np.random.default_rng(3).uniform(-1, 1, (n, k))
    A = hbs.DeviceCSR(ctx, rp, c, v)
    dX, dY = ctx.to_device(X.ravel()), ctx.empty(n * k)
    Ao = orc.CSR(rp, c, v)
    for rep in range(3):
        dY.zero()
        A.spmm(dX, dY, k, 64)
        Y = dY.download().reshape(n, k)
        for j in range(k):
            assert np.array_equal(Y[:, j], orc.spmv(Ao, X[:, j])), f"launch {rep} column {j}"


@pytest.mark.parametrize("k,tile", [(64, 64), (40, 64), (64, 16), (24, 8)])
def test_spmm_grid_runs_bit_exact(ctx, k, tile):
    """A 3-D grid operator whose plane outgrows an XCD's L2 is swept in
    y-slab runs: every column still equals the oracle's serial SpMV, and the
    FP32 SpMM equals the one over rows in order."""
    rp, c, v = hbs.gen_convdiff3d(128, 128, 6, 0.5)
    n = rp.size - 1
    X = np.random.default_rng(k + tile).uniform(-1, 1, (n, k))
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    assert A.info()["spmm_runs"] == 128 * 6 and A.info()["spmm_run_len"] == 128
    dX, dY = ctx.to_device(X.ravel()), ctx.empty(n * k)
    A.spmm(dX, dY, k, tile)
    Y = dY.download().reshape(n, k)
    Ao = orc.CSR(rp, c, v)
    for j in range(k):
        assert np.array_equal(Y[:, j], orc.spmv(Ao, X[:, j])), f"column {j}"
    X32 = X.astype(np.float32)
    dX32, dY32 = ctx.to_device(X32.ravel()), ctx.empty(n * k, np.float32)
    A.spmm(dX32, dY32, k, tile)
    hbs.set_spmm_order(0)
    try:
        B = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    finally:
        hbs.set_spmm_order(1)
    assert B.info()["spmm_runs"] == 0
    dZ32 = ctx.empty(n * k, np.float32)
    B.spmm(dX32, dZ32, k, tile)
    assert np.array_equal(dY32.download(), dZ32.download())


@pytest.mark.parametrize("k", [64, 40])
def test_spmm_long_rows_bit_exact(ctx, k):
    """Power-law rows (up to 3000 entries, rows in order, n not a multiple
    of a step's rows): the 64-column SpMM's lanes hold two rows each, a
    step's rows of very different lengths, every column in CSR order."""
    rp, c, v = hbs.gen_circuit(20003, nheavy=8, heavy_lo=300, heavy_hi=3000, seed=19)
    n = rp.size - 1
    X = np.random.default_rng(k + 5).uniform(-1, 1, (n, k))
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    Ao = orc.CSR(rp, c, v)
    dX, dY = ctx.to_device(X.ravel()), ctx.empty(n * k)
    A.spmm(dX, dY, k, 64)
    Y = dY.download().reshape(n, k)
    for j in range(k):
        assert np.array_equal(Y[:, j], orc.spmv(Ao, X[:, j])), f"column {j}"
    X32 = X.astype(np.float32)
    dX32, dY32 = ctx.to_device(X32.ravel()), ctx.empty(n * k, np.float32)
    A.spmm(dX32, dY32, k, 64)
    Y32 = dY32.download().reshape(n, k).astype(np.float64)
    kk = np.diff(rp)
    u = np.finfo(np.float32).eps / 2
    bound = (2 * kk * u / (1 - kk * u)) * orc.spmv(orc.CSR(rp, c, np.abs(v)), np.ones(n))
    v32 = v.astype(np.float32).astype(np.float64)
    for j in (0, k // 2, k - 1):
        ref = orc.spmv(orc.CSR(rp, c, v32), X32[:, j].astype(np.float64))
        assert np.all(np.abs(Y32[:, j] - ref) <= 1.01 * bound + 1e-30), f"FP32 column {j}"


def test_spmm_f32(ctx):
    rp, c, v, xt, b = convdiff64()
    n, k = rp.size - 1, 32
    X = np.random.default_rng(1).uniform(-1, 1, (n, k)).astype(np.float32)
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    dX, dY = ctx.to_device(X.ravel()), ctx.empty(n * k, np.float32)
    A.spmm(dX, dY, k, 0)
    Y = dY.download().reshape(n, k).astype(np.float64)
    ref = (orc.spmv(orc.CSR(rp, c, v), X[:, 3].astype(np.float64)))
    assert np.allclose(Y[:, 3], ref, rtol=1e-5, atol=1e-5)


def test_dot_and_blas1(ctx):
    x = np.random.default_rng(1).standard_normal(100003)
    y = np.random.default_rng(2).standard_normal(100003)
    dx, dy = ctx.to_device(x), ctx.to_device(y)
    d = ctx.dot(dx, dy)
    assert abs(d - float(x @ y)) <= 1e-12 * float(np.abs(x) @ np.abs(y))
    assert ctx.dot(dx, dy) == d   # deterministic reduction order


def test_set_values_updates_every_copy(ctx):
    """New values reach the SpMV (including its column-sorted copy), the FP32
    copy and the transpose."""
    rp, c, v, xt, b = convdiff64()
    n = rp.size - 1
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32 | hbs.HBS_CSR_TRANSPOSE)
    v2 = v * 1.5 + 0.25
    dv2 = ctx.to_device(v2)
    A.set_values(dv2)
    x = np.random.default_rng(5).uniform(-1, 1, n)
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


@pytest.mark.parametrize("block_nnz", [2304, 4608, 9216, 16384])
def test_spmv_block3_bit_exact(ctx, block_nnz):
    """3x3-block matrices (FEM-like: 12 blocks per block row, one far) read
    their values block-major with one column per block (512-thread row
    blocks, above 2048 entries): every row equals the oracle's serial CSR
    sum, FP32 within rounding, for every row block size; the same matrix
    with the plain streams gives the same y."""
    rp, c, v = hbs.gen_fem(4001, 12, 1.1, 9)
    n = rp.size - 1
    x = np.random.default_rng(block_nnz).uniform(-1, 1, n)
    ref = orc.spmv(orc.CSR(rp, c, v), x)
    hbs.set_spmv_plan(block_nnz=block_nnz)
    try:
        A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32 | hbs.HBS_CSR_TRANSPOSE)
        hbs.set_spmv_block3(0)
        try:
            B = hbs.DeviceCSR(ctx, rp, c, v)
        finally:
            hbs.set_spmv_block3(1)
    finally:
        hbs.set_spmv_plan()
    assert A.info()["block3"] == 1 and B.info()["block3"] == 0
    dx, dy, dz = ctx.to_device(x), ctx.empty(n), ctx.empty(n)
    A.spmv(dx, dy)
    B.spmv(dx, dz)
    assert np.array_equal(dy.download(), ref)
    assert np.array_equal(dz.download(), ref)
    dy.upload(np.full(n, 0.5))
    A.spmv(dx, dy, alpha=2.0, beta=-1.0)
    assert np.array_equal(dy.download(), 2.0 * ref - 0.5)
    A.spmv_t(dx, dy)
    tp, tc, tv = orc.csr_to_csc(orc.CSR(rp, c, v))
    assert np.array_equal(dy.download(), orc.spmv(orc.CSR(tp, tc, tv), x))
    dx32, dy32 = ctx.to_device(x.astype(np.float32)), ctx.empty(n, np.float32)
    A.spmv(dx32, dy32)
    ref32 = orc.spmv(orc.CSR(rp, c, v), x.astype(np.float32).astype(np.float64))
    assert np.allclose(dy32.download(), ref32, rtol=1e-5, atol=1e-5 * np.abs(ref32).max())
    # new values reach the block-major copies (FP64, FP32, transpose)
    v2 = v * 0.75 - 0.125
    A.set_values(ctx.to_device(v2))
    A.spmv(dx, dy)
    assert np.array_equal(dy.download(), orc.spmv(orc.CSR(rp, c, v2), x))
    A.spmv_t(dx, dy)
    tp, tc, tv = orc.csr_to_csc(orc.CSR(rp, c, v2))
    assert np.array_equal(dy.download(), orc.spmv(orc.CSR(tp, tc, tv), x))
    A.spmv(dx32, dy32)
    ref32 = orc.spmv(orc.CSR(rp, c, v2), x.astype(np.float32).astype(np.float64))
    assert np.allclose(dy32.download(), ref32, rtol=1e-5, atol=1e-5 * np.abs(ref32).max())


@pytest.mark.parametrize("block_nnz", [2304, 4608])
def test_spmv_block3_mixed_lengths_bit_exact(ctx, block_nnz):
    """A 3x3-block matrix whose block rows differ in length (the transpose
    of the FEM-like matrix: its far blocks land on random block rows) keeps
    whole row blocks of mixed lengths, each row's blocks located through its
    row pointers: every row equals the oracle's serial CSR sum, with every
    epilogue's alpha/beta form, and FP32 within rounding."""
    rp0, c0, v0 = hbs.gen_fem(4001, 12, 1.1, 11)
    rp, c, v = orc.csr_to_csc(orc.CSR(rp0, c0, v0))
    rp, c, v = np.asarray(rp, np.int32), np.asarray(c, np.int32), np.asarray(v)
    n = rp.size - 1
    lens = np.diff(rp)
    assert lens.min() < lens.max()                 # block rows of several lengths
    assert hbs.detect_block3(rp, c)
    x = np.random.default_rng(block_nnz + 1).uniform(-1, 1, n)
    ref = orc.spmv(orc.CSR(rp, c, v), x)
    hbs.set_spmv_plan(block_nnz=block_nnz)
    try:
        A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_F32)
    finally:
        hbs.set_spmv_plan()
    assert A.info()["block3"] == 2
    dx, dy = ctx.to_device(x), ctx.empty(n)
    A.spmv(dx, dy)
    assert np.array_equal(dy.download(), ref)
    dy.upload(np.full(n, 0.5))
    A.spmv(dx, dy, alpha=2.0, beta=-1.0)
    assert np.array_equal(dy.download(), 2.0 * ref - 0.5)
    dx32, dy32 = ctx.to_device(x.astype(np.float32)), ctx.empty(n, np.float32)
    A.spmv(dx32, dy32)
    ref32 = orc.spmv(orc.CSR(rp, c, v), x.astype(np.float32).astype(np.float64))
    assert np.allclose(dy32.download(), ref32, rtol=1e-5, atol=1e-5 * np.abs(ref32).max())


def test_block3_solve_matches_oracle(ctx):
    """The FP64 solve on a 3x3-block matrix (kind-3 row blocks) meets the
    single-system bar."""
    from parity import check
    rp, c, v = hbs.gen_fem(6000, 12, 1.1, 4)
    n = rp.size - 1
    xt = hbs.gen_uniform(n, -1.0, 1.0, 8)
    b = orc.spmv(orc.CSR(rp, c, v), xt)
    hbs.set_spmv_plan(block_nnz=4608)
    try:
        A = hbs.DeviceCSR(ctx, rp, c, v)
    finally:
        hbs.set_spmv_plan()
    assert A.info()["block3"] == 1
    S = hbs.Solver(A)
    db, dx = ctx.to_device(b), ctx.empty(n)
    dx.zero()
    info = S.bicgstab(db, dx, tol=1e-8, maxit=2000)
    xo, ito, _, _ = orc.bicgstab(orc.CSR(rp, c, v), b, tol=1e-8, maxit=2000)
    check(dx.download(), info, xo, ito, "fem 18k block3", orc.CSR(rp, c, v), b, xt)
