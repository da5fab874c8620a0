"""One rank's share on the MI355X: halo pack/unpack bit-exact against the
index lists, local+halo SpMV against the oracle's restatement of
MPI_SHARD_CSC_mpi_spmxv_seq with the halo supplied from a host buffer."""
import numpy as np
import pytest

import pyoracle as orc
from conftest import golden
from hethyppar_amd import hbs
from workloads import PARTITIONS, markov, partition

pytestmark = pytest.mark.gpu


def _rank_check(ctx, rp, c, v, pv, nparts, r, x, exact=True):
    hs = hbs.shard_build(rp, c, v, pv, nparts, r)
    sh = hbs.Shard(ctx, hs)
    m, ns = hs["loc"]["m"], hs["shr"]["n"]
    x_loc = x[hs["gind"][:m]]
    recv_buf = x[hs["gind"][m:][hs["recv"]["J"]]]   # what the peers would send
    # pack
    dxl = ctx.to_device(x_loc)
    nsend = sh.n_send
    if nsend:
        dsend = ctx.empty(nsend)
        sh.pack(dxl, dsend)
        assert np.array_equal(dsend.download(), orc.shard_pack(hs, x_loc))
    # unpack from device buffer
    dext = ctx.to_device(np.concatenate([x_loc, np.full(ns, np.nan)]))
    if ns:
        drecv = ctx.to_device(recv_buf)
        sh.unpack(drecv, dext)
        assert np.array_equal(dext.download()[m:], x[hs["gind"][m:]])
    # SpMV with the halo straight from a host buffer
    dext2 = ctx.to_device(np.concatenate([x_loc, np.zeros(ns)]))
    dy = ctx.empty(m)
    sh.spmv(dext2, dy, recv_host=recv_buf)
    y = dy.download()
    ref = orc.shard_spmv(hs, x_loc, recv_buf)
    if exact:
        assert np.array_equal(y, ref)
    else:
        assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)
    # partial dot (the reference Allreduces it)
    d = sh.dot_partial(dy, dy)
    assert abs(d - float(ref @ ref)) <= 1e-12 * float(ref @ ref)
    sh.close()
    hbs.shard_free(hs)


@pytest.mark.parametrize("name,fname,nparts", PARTITIONS)
def test_shard_reference_partitions(ctx, name, fname, nparts):
    _, _, rp, c, v = hbs.mm_read(golden(name, "A.mtx"))
    pv = partition(name, fname)
    x = np.random.default_rng(0).uniform(-1, 1, rp.size - 1)
    for r in range(nparts):
        _rank_check(ctx, rp, c, v, pv, nparts, r, x)


def test_shard_markov_part0_of_4(ctx):
    """The 445k-row system split 4 ways by the library's own partitioner
    (hbs_partition, the CalcPartVec model of hhp_matrix.c:225-300), and by a
    contiguous split."""
    rp, c, v = markov(445000)
    n = rp.size - 1
    pv, cut = hbs.partition(rp, c, 4)
    print(f"hbs_partition: 4 parts of {np.bincount(pv, minlength=4).tolist()} rows, cut {cut}", flush=True)
    x = hbs.gen_uniform(n, -1.0, 1.0, 7)
    for r in (0, 3):
        _rank_check(ctx, rp, c, v, pv, 4, r, x)
    pv = (np.arange(n) * 4 // n).astype(np.int32)
    _rank_check(ctx, rp, c, v, pv, 4, 0, x)


def test_shard_random_owner_order(ctx):
    # owners not monotone in global column order: recv.J is not the identity
    # (the reference GPU path would mis-assign here, hhp_cuda.c:323-329)
    rp, c, v = markov(20000)
    pv = np.random.default_rng(5).integers(0, 6, rp.size - 1).astype(np.int32)
    x = np.random.default_rng(6).uniform(-1, 1, rp.size - 1)
    hs = hbs.shard_build(rp, c, v, pv, 6, 2)
    assert not np.array_equal(hs["recv"]["J"], np.arange(hs["shr"]["n"]))
    hbs.shard_free(hs)
    for r in range(6):
        _rank_check(ctx, rp, c, v, pv, 6, r, x)
