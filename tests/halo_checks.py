"""Invariants that pin the halo index lists (loc/shr split, send/recv lists).

The reference holds no golden halo files, so the lists are pinned by what
the algorithm of src/util/hhp_matrix.c:820-1247 fixes:
  * gind = owned rows ascending, then every non-owned column the owned rows
    reference, ascending;
  * loc / shr hold exactly the owned rows' entries, split by column owner;
  * recv.J is a permutation of [0, n_shr) grouped by owner, owners ascending,
    columns ascending inside an owner block;
  * send(p -> q) = locmap_p(global columns of q's recv block for p), in order;
  * the shard SpMV summed over ranks equals the global SpMV.
"""
import numpy as np


def check_shard(rp, col, val, partvec, nparts, rank, s):
    """Raise AssertionError if shard s of `rank` violates the invariants."""
    n = rp.size - 1
    owned = np.flatnonzero(partvec == rank).astype(np.int32)
    m = owned.size
    assert s["loc"]["m"] == m and s["loc"]["n"] == m and s["shr"]["m"] == m
    refs = set()
    for r in owned:
        for j in range(rp[r], rp[r + 1]):
            if partvec[col[j]] != rank:
                refs.add(int(col[j]))
    shr_cols = np.array(sorted(refs), dtype=np.int32)
    assert np.array_equal(s["gind"], np.concatenate([owned, shr_cols]))
    assert s["shr"]["n"] == shr_cols.size
    # loc / shr entries == owned rows' entries (as multisets per row/col pair)
    loc = {}
    shr = {}
    for lr, r in enumerate(owned):
        for j in range(rp[r], rp[r + 1]):
            key = (lr, int(col[j]))
            (loc if partvec[col[j]] == rank else shr).setdefault(key, []).append(val[j])
    got_loc, got_shr = {}, {}
    for c in range(m):
        for j in range(s["loc"]["I"][c], s["loc"]["I"][c + 1]):
            got_loc.setdefault((int(s["loc"]["J"][j]), int(owned[c])), []).append(s["loc"]["val"][j])
    for c in range(shr_cols.size):
        for j in range(s["shr"]["I"][c], s["shr"]["I"][c + 1]):
            got_shr.setdefault((int(s["shr"]["J"][j]), int(shr_cols[c])), []).append(s["shr"]["val"][j])
    assert got_loc == loc and got_shr == shr
    # row indices ascend inside each CSC column
    for blk in ("loc", "shr"):
        I, J = s[blk]["I"], s[blk]["J"]
        for c in range(s[blk]["n"]):
            assert np.all(np.diff(J[I[c]:I[c + 1]]) >= 0)
    # recv: grouping by owner
    recv = s["recv"]
    owners = partvec[shr_cols]
    exp_ranks = np.unique(owners)
    assert np.array_equal(recv["ranks"], exp_ranks)
    exp_J = np.concatenate([np.flatnonzero(owners == q) for q in exp_ranks]).astype(np.int32) \
        if shr_cols.size else np.zeros(0, np.int32)
    assert np.array_equal(recv["J"], exp_J)
    exp_I = np.concatenate([[0], np.cumsum([np.sum(owners == q) for q in exp_ranks])]).astype(np.int32)
    assert np.array_equal(recv["I"], exp_I)


def check_send(shards, partvec, nparts):
    """send(p -> q) == locmap_p(q's recv block for p)."""
    for p in range(nparts):
        sp = shards[p]
        locmap = {int(g): i for i, g in enumerate(sp["gind"][: sp["loc"]["n"]])}
        exp_ranks, exp_J, exp_I = [], [], [0]
        for q in range(nparts):
            if q == p:
                continue
            rq = shards[q]["recv"]
            hit = np.flatnonzero(rq["ranks"] == p)
            if hit.size == 0:
                continue
            k = int(hit[0])
            blk = rq["J"][rq["I"][k]:rq["I"][k + 1]]
            glob = shards[q]["gind"][shards[q]["loc"]["n"] + blk]
            exp_ranks.append(q)
            exp_J.extend(locmap[int(g)] for g in glob)
            exp_I.append(len(exp_J))
        send = sp["send"]
        assert np.array_equal(send["ranks"], np.array(exp_ranks, dtype=np.int32))
        assert np.array_equal(send["J"], np.array(exp_J, dtype=np.int32))
        assert np.array_equal(send["I"], np.array(exp_I, dtype=np.int32))


def assert_same_shard(a, b):
    for blk in ("loc", "shr"):
        for k in ("I", "J", "val"):
            assert np.array_equal(a[blk][k], b[blk][k]), (blk, k)
        for k in ("nnz", "m", "n"):
            assert a[blk][k] == b[blk][k], (blk, k)
    assert np.array_equal(a["gind"], b["gind"])
    for c in ("send", "recv"):
        assert a[c]["num"] == b[c]["num"], c
        for k in ("I", "J", "ranks"):
            assert np.array_equal(a[c][k], b[c][k]), (c, k)
