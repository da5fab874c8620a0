"""Pin the CPU oracle against the reference's fixtures and the mathematics."""
import copy

import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as orc
from conftest import golden, read_ints
from halo_checks import check_send, check_shard
from workloads import PARTITIONS, convdiff64, dense, markov, partition, random_csr


def _file_triplets(path):
    """Independent parse of a .mtx file: (row, col, val) in file order."""
    with open(path) as f:
        lines = [l for l in f if not l.startswith("%")]
    m, n, nz = (int(t) for t in lines[0].split())
    trip = [(int(a) - 1, int(b) - 1, float(c)) for a, b, c in (l.split() for l in lines[1:1 + nz])]
    return m, n, trip


@pytest.mark.parametrize("name,m,nnz", [("sample", 16, 111), ("sample320", 320, 830)])
def test_mm_read_fixture(name, m, nnz):
    M, N, rp, c, v = orc.mm_read(golden(name, "A.mtx"))
    assert (M, N, rp[-1]) == (m, m, nnz)
    _, _, trip = _file_triplets(golden(name, "A.mtx"))
    # per-row FILE order (hhp_wrap_bora.c:78-84 fills backwards with --ia[r])
    for i in range(M):
        exp = [(cc, vv) for r, cc, vv in trip if r == i]
        got = list(zip(c[rp[i]:rp[i + 1]].tolist(), v[rp[i]:rp[i + 1]].tolist()))
        assert got == exp


def test_mm_read_keeps_duplicates_and_file_order(tmp_path):
    p = tmp_path / "d.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% c\n3 3 6\n"
                 "2 3 1.5\n1 1 2\n2 1 -1\n2 3 0.25\n3 2 7\n1 3 4\n")
    M, N, rp, c, v = orc.mm_read(str(p))
    assert rp.tolist() == [0, 2, 5, 6]
    assert c.tolist() == [0, 2, 2, 0, 2, 1]
    assert v.tolist() == [2.0, 4.0, 1.5, -1.0, 0.25, 7.0]


def test_mm_read_rejects_symmetric(tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n1 1 1\n")
    with pytest.raises(RuntimeError):
        orc.mm_read(str(p))


def test_spmv_kat_sample():
    # data/sample/in/X_mul.txt is built so that A * X_mul = [1..16]
    _, _, rp, c, v = orc.mm_read(golden("sample", "A.mtx"))
    A = orc.CSR(rp, c, v)
    x = orc.vector_read(golden("sample", "X_mul.txt"), 16)
    y = orc.spmv(A, x)
    assert np.max(np.abs(y - np.arange(1, 17))) < 1e-13


@pytest.mark.parametrize("name,n,tol", [("sample", 16, 1e-12), ("sample320", 320, 1e-12)])
def test_spmv_kat_b(name, n, tol):
    # B.txt = A * [1..n] (tools/scripts/process_matrixi.m)
    _, _, rp, c, v = orc.mm_read(golden(name, "A.mtx"))
    A = orc.CSR(rp, c, v)
    b = orc.vector_read(golden(name, "B.txt"), n)
    y = orc.spmv(A, np.arange(1, n + 1, dtype=float))
    assert np.max(np.abs(y - b) / np.maximum(1, np.abs(b))) < tol


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spmv_dense_brute_force(seed):
    rp, c, v = random_csr(57, 43, 0.2, seed)
    A = orc.CSR(rp, c, v, ncols=43)
    x = np.random.default_rng(seed + 10).standard_normal(43)
    assert np.allclose(orc.spmv(A, x), dense(rp, c, v, 43) @ x, rtol=1e-13, atol=1e-13)
    assert np.array_equal(orc.spmv(A, x), orc.spmv(A, x, omp=True))


def test_csr_to_csc_matches_scipy():
    rp, c, v = random_csr(40, 35, 0.15, 5, empty_rows=False)
    A = orc.CSR(rp, c, v, ncols=35)
    I, J, val = orc.csr_to_csc(A)
    S = sp.csr_matrix((v, c, rp), shape=(40, 35)).tocsc()
    S.sort_indices()
    assert np.array_equal(I, S.indptr) and np.array_equal(J, S.indices) and np.array_equal(val, S.data)
    R, Cc, Vv = orc.shard_csc_to_csr(I, J, val, 40, 35)
    assert np.array_equal(R, rp) and np.array_equal(Cc, c) and np.array_equal(Vv, v)


def test_bicgstab_sample320_golden():
    # reference outputs of bicgstab-cpu / -mpi / -mpi-gpu: X = [1..320] to %lf
    _, _, rp, c, v = orc.mm_read(golden("sample320", "A.mtx"))
    A = orc.CSR(rp, c, v)
    b = orc.vector_read(golden("sample320", "B.txt"), 320)
    x0 = orc.vector_read(golden("sample320", "X_init.txt"), 320)
    x, it, rr, stop = orc.bicgstab(A, b, x0, maxit=20, fixed=True)
    assert it == 20
    for f in ("out_X_seq.txt", "out_X1.txt", "out_X_mpi_gpu.txt"):
        ref = orc.vector_read(golden("sample320", f), 320)
        assert np.max(np.abs(x - ref)) <= 1e-6 + 1e-9
    assert rr < 1e-8


def test_bicgstab_convdiff_recovers_target():
    rp, c, v, xt, b = convdiff64()
    A = orc.CSR(rp, c, v)
    x, it, rr, stop = orc.bicgstab(A, b, tol=1e-8, maxit=1000)
    assert stop in (1, 2) and rr <= 1.01e-8
    # forward error <= cond(A) * relative residual; cond of this operator ~1e3
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-5
    # the number of iterations grows as the tolerance shrinks
    its = [orc.bicgstab(A, b, tol=t, maxit=1000)[1] for t in (1e-2, 1e-4, 1e-6, 1e-8, 1e-10)]
    assert its == sorted(its) and its[0] < its[-1]


def test_bicgstab_vs_scipy_iterations():
    from scipy.sparse.linalg import bicgstab as sp_bicgstab
    rp, c, v, xt, b = convdiff64()
    A = orc.CSR(rp, c, v)
    x, it, rr, stop = orc.bicgstab(A, b, tol=1e-8, maxit=1000)
    S = sp.csr_matrix((v, c, rp), shape=(rp.size - 1,) * 2)
    count = [0]
    xs, info = sp_bicgstab(S, b, rtol=1e-8, atol=0.0, maxiter=1000,
                           callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert info == 0
    assert abs(count[0] - it) <= 2
    assert np.linalg.norm(xs - x) / np.linalg.norm(x) < 1e-6


def test_bicgstab_fixed_matches_first_iterations():
    # fixed mode runs exactly maxit iterations with no exit (bicgstab_cpu.c)
    rp, c, v, xt, b = convdiff64()
    A = orc.CSR(rp, c, v)
    _, it, _, stop = orc.bicgstab(A, b, maxit=7, fixed=True)
    assert it == 7 and stop == 0


@pytest.mark.parametrize("name,fname,nparts", PARTITIONS)
def test_halo_lists_reference_partitions(name, fname, nparts):
    _, _, rp, c, v = orc.mm_read(golden(name, "A.mtx"))
    pv = partition(name, fname)
    A = orc.CSR(rp, c, v)
    shards = orc.build_shards(A, pv, nparts)
    for r in range(nparts):
        check_shard(rp, c, v, pv, nparts, r, shards[r])
    check_send(shards, pv, nparts)


def _global_vs_shards(rp, c, v, pv, nparts, shards, x):
    y_global = orc.spmv(orc.CSR(rp, c, v), x)
    for r in range(nparts):
        s = shards[r]
        x_loc = orc.vector_read_parted(x, pv, r)
        recv_buf = x[s["gind"][s["loc"]["n"]:][s["recv"]["J"]]]
        y = orc.shard_spmv(s, x_loc, recv_buf)
        owned = np.flatnonzero(pv == r)
        assert np.allclose(y, y_global[owned], rtol=1e-12, atol=1e-12)
        # pack then deliver: what rank q receives from r is x at q's recv block
        packed = orc.shard_pack(s, x_loc)
        for k, q in enumerate(s["send"]["ranks"]):
            sq = shards[q]
            kk = int(np.flatnonzero(sq["recv"]["ranks"] == r)[0])
            blk = sq["recv"]["J"][sq["recv"]["I"][kk]:sq["recv"]["I"][kk + 1]]
            exp = x[sq["gind"][sq["loc"]["n"] + blk]]
            got = packed[s["send"]["I"][k]:s["send"]["I"][k + 1]]
            assert np.array_equal(got, exp)


@pytest.mark.parametrize("nparts,seed", [(2, 0), (4, 1), (7, 2)])
def test_halo_random_partitions_and_spmv(nparts, seed):
    rp, c, v, xt, b = convdiff64()
    pv = np.random.default_rng(seed).integers(0, nparts, rp.size - 1).astype(np.int32)
    shards = orc.build_shards(orc.CSR(rp, c, v), pv, nparts)
    for r in range(nparts):
        check_shard(rp, c, v, pv, nparts, r, shards[r])
    check_send(shards, pv, nparts)
    _global_vs_shards(rp, c, v, pv, nparts, shards, xt)


def test_halo_markov_contiguous_4_parts():
    rp, c, v = markov(4000)
    n = rp.size - 1
    pv = (np.arange(n) * 4 // n).astype(np.int32)
    shards = orc.build_shards(orc.CSR(rp, c, v), pv, 4)
    check_send(shards, pv, 4)
    check_shard(rp, c, v, pv, 4, 0, shards[0])
    x = np.random.default_rng(3).uniform(-1, 1, n)
    _global_vs_shards(rp, c, v, pv, 4, shards, x)


def test_halo_checks_catch_planted_mistakes():
    _, _, rp, c, v = orc.mm_read(golden("sample320", "A.mtx"))
    pv = partition("sample320", "row_part_16")
    shards = orc.build_shards(orc.CSR(rp, c, v), pv, 16)
    # recv.J not grouped by owner
    bad = copy.deepcopy(shards[0])
    if bad["recv"]["J"].size >= 2:
        bad["recv"]["J"][[0, -1]] = bad["recv"]["J"][[-1, 0]]
        with pytest.raises(AssertionError):
            check_shard(rp, c, v, pv, 16, 0, bad)
    # send list permuted within a peer block
    worse = copy.deepcopy(shards)
    s = worse[3]["send"]
    j = int(np.argmax(np.diff(s["I"])))
    a, b_ = s["I"][j], s["I"][j + 1]
    assert b_ - a >= 2
    s["J"][a:b_] = s["J"][a:b_][::-1]
    with pytest.raises(AssertionError):
        check_send(worse, pv, 16)
    # an off-by-one in the shr column map
    off = copy.deepcopy(shards[5])
    off["gind"][-1] += 1
    with pytest.raises(AssertionError):
        check_shard(rp, c, v, pv, 16, 5, off)


def test_vector_read_parted():
    x = np.arange(10, dtype=float)
    pv = np.array([1, 0, 1, 1, 0, 2, 0, 1, 2, 0], dtype=np.int32)
    assert orc.vector_read_parted(x, pv, 0).tolist() == [1, 4, 6, 9]
    assert orc.vector_read_parted(x, pv, 1).tolist() == [0, 2, 3, 7]


def test_vector_gather_brute_force_and_round_trip():
    """MPI_vector_gather (hhp_cpu.c:199-250) restated: brute force on a small
    case, then slice -> gather is the identity over every rank of the
    reference partitions (gind's owned prefix is the slice's row order), and
    a planted swap of two gind entries breaks the round trip."""
    big = orc.vector_gather([np.array([5.0, 6.0]), np.array([7.0])], [np.array([2, 0]), np.array([1])], 4,
                            big=np.full(4, -1.0))
    assert big.tolist() == [6.0, 7.0, 5.0, -1.0]
    for name, fname, nparts in PARTITIONS:
        _, _, rp, c, v = orc.mm_read(golden(name, "A.mtx"))
        n = rp.size - 1
        pv = partition(name, fname)
        x = np.random.default_rng(3).standard_normal(n)
        shards = orc.build_shards(orc.CSR(rp, c, v), pv, nparts)
        locs = [orc.vector_read_parted(x, pv, r) for r in range(nparts)]
        ginds = [sh["gind"][:sh["loc"]["n"]] for sh in shards]
        assert np.array_equal(orc.vector_gather(locs, ginds, n), x)
    bad = [g.copy() for g in ginds]
    k = next(r for r in range(nparts) if bad[r].size >= 2)
    bad[k][[0, 1]] = bad[k][[1, 0]]
    assert not np.array_equal(orc.vector_gather(locs, bad, n), x)


def test_blocked_dots_variant_reduces_to_reference():
    """The re-associated variant (order-sensitivity probe only) is the
    reference loop when one block (or one interleaved running sum) holds the
    whole vector, or blk == 0."""
    rp, c, v, xt, b = convdiff64()
    A = orc.CSR(rp, c, v)
    x0, i0, r0, s0 = orc.bicgstab(A, b, tol=1e-8, maxit=1000)
    for blk in (0, rp.size - 1, 10 ** 6, -1):
        x1, i1, r1, s1 = orc.bicgstab_blocked(A, b, blk, tol=1e-8, maxit=1000)
        assert (i1, s1) == (i0, s0) and np.array_equal(x1, x0) and r1 == r0
    x2, i2, _, _ = orc.bicgstab_blocked(A, b, 64, tol=1e-8, maxit=1000)
    assert abs(i2 - i0) <= 2 and np.linalg.norm(x2 - xt) / np.linalg.norm(xt) < 1e-5
