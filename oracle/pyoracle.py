"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use it to check results / as the CPU baseline.  It is
never a fallback for the HIP path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class OrcCSR(C.Structure):
    _fields_ = [("I", _ip), ("J", _ip), ("val", _dp), ("nnz", C.c_int), ("m", C.c_int), ("n", C.c_int)]


class OrcComm(C.Structure):
    _fields_ = [("val", _dp), ("J", _ip), ("I", _ip), ("ranks", _ip), ("num", C.c_int)]


class OrcShard(C.Structure):
    _fields_ = [("loc", OrcCSR), ("shr", OrcCSR), ("gind", _ip), ("ngind", C.c_int),
                ("send", OrcComm), ("recv", OrcComm)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_mm_read.argtypes = [C.c_char_p, P(OrcCSR)]
        L.orc_vector_read.argtypes = [C.c_char_p, C.c_int, _dp]
        L.orc_csr_free.argtypes = [P(OrcCSR)]
        for f in ("orc_csr_spmv_seq", "orc_csr_spmv_omp", "orc_csr_spmv_seq_acc", "orc_csc_spmv_seq_acc"):
            getattr(L, f).argtypes = [P(OrcCSR), _dp, _dp]
            getattr(L, f).restype = None
        L.orc_vector_dot_seq.argtypes = [_dp, _dp, C.c_int]
        L.orc_vector_dot_seq.restype = C.c_double
        L.orc_bicgstab.argtypes = [P(OrcCSR), _dp, _dp, C.c_double, C.c_int, C.c_int, _ip, _dp, _ip]
        L.orc_bicgstab_blocked.argtypes = [P(OrcCSR), _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_int, _ip, _dp,
                                           _ip]
        L.orc_csr_to_csc.argtypes = [P(OrcCSR), P(OrcCSR)]
        L.orc_shard_csc_to_csr.argtypes = [P(OrcCSR), P(OrcCSR)]
        L.orc_split_row.argtypes = [P(OrcCSR), _ip, C.c_int, C.c_int, P(OrcCSR)]
        L.orc_build_shards.argtypes = [P(OrcCSR), _ip, C.c_int, P(OrcShard)]
        L.orc_shard_free.argtypes = [P(OrcShard)]
        L.orc_vector_read_parted.argtypes = [_dp, _ip, C.c_int, C.c_int, _dp]
        L.orc_vector_gather.argtypes = [C.c_int, _ip, C.POINTER(_dp), C.POINTER(_ip), C.c_int, _dp]
        L.orc_vector_gather.restype = None
        L.orc_shard_spmv.argtypes = [P(OrcShard), _dp, _dp, _dp]
        L.orc_shard_pack.argtypes = [P(OrcShard), _dp, _dp]
        _lib = L
    return _lib


def _dp_of(a):
    return a.ctypes.data_as(_dp)


def _ip_of(a):
    return a.ctypes.data_as(_ip)


def _np(ptr, n, dtype):
    if n <= 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class CSR:
    """Holds numpy arrays and an OrcCSR view of them."""

    def __init__(self, rowptr, col, val, ncols=None):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.m = self.rowptr.size - 1
        self.n = self.m if ncols is None else ncols
        self.s = OrcCSR(_ip_of(self.rowptr), _ip_of(self.col), _dp_of(self.val),
                        int(self.col.size), self.m, self.n)


def mm_read(path):
    s = OrcCSR()
    rc = lib().orc_mm_read(os.fsencode(path), C.byref(s))
    if rc != 0:
        raise RuntimeError(f"orc_mm_read failed: {rc}")
    out = (s.m, s.n, _np(s.I, s.m + 1, np.int32), _np(s.J, s.nnz, np.int32), _np(s.val, s.nnz, np.float64))
    lib().orc_csr_free(C.byref(s))
    return out


def vector_read(path, size):
    out = np.zeros(size, dtype=np.float64)
    rc = lib().orc_vector_read(os.fsencode(path), size, _dp_of(out))
    if rc < 0:
        raise RuntimeError(f"orc_vector_read failed: {rc}")
    return out


def spmv(A, x, omp=False):
    y = np.zeros(A.m, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = lib().orc_csr_spmv_omp if omp else lib().orc_csr_spmv_seq
    f(C.byref(A.s), _dp_of(x), _dp_of(y))
    return y


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_vector_dot_seq(_dp_of(a), _dp_of(b), a.size)


def bicgstab(A, b, x0=None, tol=1e-8, maxit=1000, fixed=False):
    """Returns (x, iters, relres, stop)."""
    x = np.zeros(A.m) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    it, st = C.c_int(), C.c_int()
    rr = C.c_double()
    rc = lib().orc_bicgstab(C.byref(A.s), _dp_of(b), _dp_of(x), tol, maxit, 1 if fixed else 0,
                            C.byref(it), C.byref(rr), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_bicgstab failed: {rc}")
    return x, it.value, rr.value, st.value


def bicgstab_blocked(A, b, blk, x0=None, tol=1e-8, maxit=1000, fixed=False):
    """bicgstab with block-summed dots (blk elements per block); returns
    (x, iters, relres, stop).  Measures order sensitivity only."""
    x = np.zeros(A.m) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    it, st = C.c_int(), C.c_int()
    rr = C.c_double()
    rc = lib().orc_bicgstab_blocked(C.byref(A.s), _dp_of(b), _dp_of(x), tol, maxit, 1 if fixed else 0, blk,
                                    C.byref(it), C.byref(rr), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_bicgstab_blocked failed: {rc}")
    return x, it.value, rr.value, st.value


def iteration_spread(A, b, tol=1e-8, maxit=1000, blocks=(512, 64, -256, -4096), perturb=4, threads=8,
                     seed0=0, sequential=True):
    """How far rounding-level changes move the oracle's solve: the reference
    loop with its dot sums re-associated (blocks > 0: consecutive blocks;
    < 0: interleaved running sums), and the reference loop on b perturbed by
    1e-15 relative (seeds seed0..seed0+perturb-1).  Returns (counts,
    solutions), the sequential run first when `sequential`.  Runs in threads
    (the oracle's calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    b = np.ascontiguousarray(b, dtype=np.float64)
    jobs = [lambda: bicgstab(A, b, tol=tol, maxit=maxit)] if sequential else []
    jobs += [(lambda k=k: bicgstab_blocked(A, b, k, tol=tol, maxit=maxit)) for k in blocks]
    for seed in range(seed0, seed0 + perturb):
        bb = b * (1.0 + 1e-15 * np.random.default_rng(seed).standard_normal(b.size))
        jobs.append(lambda bb=bb: bicgstab(A, bb, tol=tol, maxit=maxit))
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        res = list(ex.map(lambda f: f(), jobs))
    return [r[1] for r in res], [r[0] for r in res]


def csr_to_csc(A):
    out = OrcCSR()
    rc = lib().orc_csr_to_csc(C.byref(A.s), C.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_csr_to_csc failed: {rc}")
    res = (_np(out.I, out.n + 1, np.int32), _np(out.J, out.nnz, np.int32), _np(out.val, out.nnz, np.float64))
    lib().orc_csr_free(C.byref(out))
    return res


def shard_csc_to_csr(colptr, rowidx, val, m, n):
    colptr = np.ascontiguousarray(colptr, dtype=np.int32)
    rowidx = np.ascontiguousarray(rowidx, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    src = OrcCSR(_ip_of(colptr), _ip_of(rowidx), _dp_of(val), int(rowidx.size), m, n)
    out = OrcCSR()
    rc = lib().orc_shard_csc_to_csr(C.byref(src), C.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_shard_csc_to_csr failed: {rc}")
    res = (_np(out.I, out.m + 1, np.int32), _np(out.J, out.nnz, np.int32), _np(out.val, out.nnz, np.float64))
    lib().orc_csr_free(C.byref(out))
    return res


def _comm(c):
    I = _np(c.I, c.num + 1, np.int32)
    cnt = int(I[-1]) if c.num > 0 else 0
    return {"I": I, "J": _np(c.J, cnt, np.int32), "ranks": _np(c.ranks, c.num, np.int32), "num": c.num}


def _csc(c):
    return {"I": _np(c.I, c.n + 1, np.int32), "J": _np(c.J, c.nnz, np.int32),
            "val": _np(c.val, c.nnz, np.float64), "nnz": c.nnz, "m": c.m, "n": c.n}


def build_shards(A, partvec, nparts):
    """All ranks' SHARD_CSC in one process -> list of dicts (numpy copies)."""
    pv = np.ascontiguousarray(partvec, dtype=np.int32)
    arr = (OrcShard * nparts)()
    rc = lib().orc_build_shards(C.byref(A.s), _ip_of(pv), nparts, arr)
    if rc != 0:
        for k in range(nparts):
            lib().orc_shard_free(C.byref(arr[k]))
        raise RuntimeError(f"orc_build_shards failed: {rc}")
    out = []
    for k in range(nparts):
        s = arr[k]
        out.append({"loc": _csc(s.loc), "shr": _csc(s.shr), "gind": _np(s.gind, s.ngind, np.int32),
                    "send": _comm(s.send), "recv": _comm(s.recv)})
        lib().orc_shard_free(C.byref(arr[k]))
    return out


def _shard_struct(d):
    keep = []

    def arr(a, t):
        a = np.ascontiguousarray(a, dtype=t)
        keep.append(a)
        return a.ctypes.data_as(_ip if t == np.int32 else _dp)

    def csc(c):
        return OrcCSR(arr(c["I"], np.int32), arr(c["J"], np.int32), arr(c["val"], np.float64),
                      c["nnz"], c["m"], c["n"])

    def comm(c):
        n = int(c["I"][-1]) if c["num"] > 0 else 0
        return OrcComm(arr(np.zeros(max(n, 1)), np.float64), arr(c["J"] if n else np.zeros(1, np.int32), np.int32),
                       arr(c["I"], np.int32), arr(c["ranks"] if c["num"] else np.zeros(1, np.int32), np.int32),
                       c["num"])

    s = OrcShard(csc(d["loc"]), csc(d["shr"]), arr(d["gind"], np.int32), int(d["gind"].size),
                 comm(d["send"]), comm(d["recv"]))
    return s, keep


def shard_spmv(d, x_loc, recv_buf):
    s, keep = _shard_struct(d)
    y = np.zeros(d["loc"]["m"], dtype=np.float64)
    x_loc = np.ascontiguousarray(x_loc, dtype=np.float64)
    rb = np.ascontiguousarray(recv_buf if len(recv_buf) else np.zeros(1), dtype=np.float64)
    lib().orc_shard_spmv(C.byref(s), _dp_of(x_loc), _dp_of(rb), _dp_of(y))
    return y


def shard_pack(d, x_loc):
    s, keep = _shard_struct(d)
    n = int(d["send"]["I"][-1]) if d["send"]["num"] > 0 else 0
    out = np.zeros(max(n, 1), dtype=np.float64)
    x_loc = np.ascontiguousarray(x_loc, dtype=np.float64)
    lib().orc_shard_pack(C.byref(s), _dp_of(x_loc), _dp_of(out))
    return out[:n]


def vector_read_parted(big, partvec, rank):
    big = np.ascontiguousarray(big, dtype=np.float64)
    pv = np.ascontiguousarray(partvec, dtype=np.int32)
    out = np.zeros(int((pv == rank).sum()), dtype=np.float64)
    lib().orc_vector_read_parted(_dp_of(big), _ip_of(pv), big.size, rank, _dp_of(out))
    return out


def vector_gather(locs, ginds, size, big=None):
    """Every rank's (local values, gind prefix) in rank order -> the global
    vector (MPI_vector_gather at the root); entries no rank owns keep big's."""
    big = np.zeros(size) if big is None else np.array(big, dtype=np.float64, copy=True)
    locs = [np.ascontiguousarray(a, dtype=np.float64) for a in locs]
    ginds = [np.ascontiguousarray(g, dtype=np.int32) for g in ginds]
    counts = np.array([a.size for a in locs], dtype=np.int32)
    lp = (_dp * len(locs))(*[_dp_of(a) for a in locs])
    gp = (_ip * len(ginds))(*[_ip_of(g) for g in ginds])
    lib().orc_vector_gather(len(locs), _ip_of(counts), lp, gp, size, _dp_of(big))
    return big
