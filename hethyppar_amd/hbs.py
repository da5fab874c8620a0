"""ctypes binding of libhbs.so (include/hbs.h).

The library is the product; this module only marshals numpy arrays and
handles.  It raises if libhbs.so is missing: there is no Python or CPU
fallback for any device operation.
"""
import ctypes as C
import os
import weakref

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhbs.so")

HBS_OK = 0
ERRORS = {
    -1: "HBS_ERR_ARG", -2: "HBS_ERR_HIP", -3: "HBS_ERR_ALLOC", -4: "HBS_ERR_IO",
    -5: "HBS_ERR_FORMAT", -6: "HBS_ERR_PARTITION", -7: "HBS_ERR_BREAKDOWN",
    -8: "HBS_ERR_NOCONV", -9: "HBS_ERR_NODEVICE",
}
HBS_ERR_BREAKDOWN = -7
HBS_ERR_NOCONV = -8
HBS_CSR_F32 = 1
HBS_CSR_TRANSPOSE = 2


class HBSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class HostCSR(C.Structure):        # CSR/CSC of hhp_common.h:43-58
    _fields_ = [("I", C.POINTER(C.c_int)), ("J", C.POINTER(C.c_int)),
                ("val", C.POINTER(C.c_double)), ("nnz", C.c_int), ("m", C.c_int), ("n", C.c_int)]


class HostIVec(C.Structure):       # iVector
    _fields_ = [("vals", C.POINTER(C.c_int)), ("nvals", C.c_uint)]


class HostComm(C.Structure):       # COMM
    _fields_ = [("val", C.POINTER(C.c_double)), ("J", C.POINTER(C.c_int)),
                ("I", C.POINTER(C.c_int)), ("ranks", C.POINTER(C.c_int)), ("num", C.c_int)]


class HostShard(C.Structure):      # SHARD_CSC
    _fields_ = [("loc", HostCSR), ("shr", HostCSR), ("gind", HostIVec),
                ("send", HostComm), ("recv", HostComm)]


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int), ("fixed_iters", C.c_int),
                ("check_every", C.c_int), ("use_graph", C.c_int), ("inner_maxit", C.c_int),
                ("inner_tol", C.c_double), ("max_sweeps", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("stop", C.c_int), ("sweeps", C.c_int), ("status", C.c_int),
                ("relres", C.c_double), ("rel_r0", C.c_double), ("time_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_ipp = C.POINTER(_ip)
_dpp = C.POINTER(_dp)

_PROTOS = {
    "hbs_last_error": (C.c_char_p, []),
    "hbs_version": (C.c_char_p, []),
    "hbs_device_count": (_i, [_ip]),
    "hbs_ctx_create": (_i, [_i, _vp, C.POINTER(_vp)]),
    "hbs_ctx_destroy": (_i, [_vp]),
    "hbs_ctx_sync": (_i, [_vp]),
    "hbs_ctx_stream": (_vp, [_vp]),
    "hbs_dev_alloc": (_i, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "hbs_dev_free": (_i, [_vp, _vp]),
    "hbs_h2d": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "hbs_d2h": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "hbs_d2d": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "hbs_dev_zero": (_i, [_vp, _vp, C.c_size_t]),
    "hbs_pinned_alloc": (_i, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "hbs_pinned_free": (_i, [_vp, _vp]),
    "hbs_csr_create": (_i, [_vp, _i, _i, _i, _ip, _ip, _dp, _i, C.POINTER(_vp)]),
    "hbs_csr_create_systems": (_i, [_vp, _i, _ip, _i, _ip, _ip, _dp, _i, C.POINTER(_vp)]),
    "hbs_csr_systems": (_i, [_vp, _ip]),
    "hbs_csr_set_values": (_i, [_vp, _vp]),
    "hbs_csr_destroy": (_i, [_vp]),
    "hbs_csr_info": (_i, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "hbs_csr_device_arrays": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "hbs_csr_spmv_f64": (_i, [_vp, _d, _vp, _d, _vp]),
    "hbs_csr_spmv_f32": (_i, [_vp, C.c_float, _vp, C.c_float, _vp]),
    "hbs_csr_spmv_t_f64": (_i, [_vp, _vp, _vp]),
    "hbs_csr_spmm_f64": (_i, [_vp, _i, _vp, _vp, _i]),
    "hbs_csr_spmm_f32": (_i, [_vp, _i, _vp, _vp, _i]),
    "hbs_dot_f64": (_i, [_vp, _i, _vp, _vp, _dp]),
    "hbs_dot_f64_noalloc": (_i, [_vp, _i, _vp, _vp, _dp]),
    "hbs_axpy_f64": (_i, [_vp, _i, _d, _vp, _vp]),
    "hbs_scal_f64": (_i, [_vp, _i, _d, _vp]),
    "hbs_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "hbs_solver_create": (_i, [_vp, C.POINTER(_vp)]),
    "hbs_solver_destroy": (_i, [_vp]),
    "hbs_bicgstab_f64": (_i, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "hbs_bicgstab_mixed": (_i, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "hbs_bicgstab_systems_f64": (_i, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "hbs_bicgstab_block_f64": (_i, [_vp, _i, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "hbs_adjoint_f64": (_i, [_vp, _vp, _vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
    "hbs_host_free": (None, [_vp]),
    "hbs_mm_read": (_i, [C.c_char_p, _ip, _ip, _ip, _ipp, _ipp, _dpp]),
    "hbs_vector_read": (_i, [C.c_char_p, _i, _dp]),
    "hbs_vector_write": (_i, [C.c_char_p, _i, _dp]),
    "hbs_gen_convdiff2d": (_i, [_i, _i, _d, _ip, _ip, _ipp, _ipp, _dpp]),
    "hbs_gen_convdiff3d": (_i, [_i, _i, _i, _d, _ip, _ip, _ipp, _ipp, _dpp]),
    "hbs_gen_markov": (_i, [_i, _i, _i, _i, _d, C.c_uint64, _ip, _ipp, _ipp, _dpp]),
    "hbs_gen_circuit": (_i, [_i, _d, _d, _i, _i, _i, _d, _i, C.c_uint64, _ip, _ipp, _ipp, _dpp]),
    "hbs_gen_fem": (_i, [_i, _i, _d, C.c_uint64, _ip, _ipp, _ipp, _dpp]),
    "hbs_gen_uniform": (_i, [_i, _d, _d, C.c_uint64, _dp]),
    "hbs_host_csr_spmv": (_i, [_i, _ip, _ip, _dp, _dp, _dp]),
    "hbs_shard_build_host": (_i, [_i, _ip, _ip, _dp, _ip, _i, _i, C.POINTER(HostShard)]),
    "hbs_partition": (_i, [_i, _ip, _ip, _i, _dp, C.c_double, _i, _ip, C.POINTER(C.c_longlong)]),
    "hbs_host_shard_free": (None, [C.POINTER(HostShard)]),
    "hbs_shard_create": (_i, [_vp, C.POINTER(HostShard), C.POINTER(_vp)]),
    "hbs_shard_destroy": (_i, [_vp]),
    "hbs_shard_info": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "hbs_shard_pack": (_i, [_vp, _vp, _vp]),
    "hbs_shard_unpack": (_i, [_vp, _vp, _vp]),
    "hbs_shard_spmv": (_i, [_vp, _vp, _dp, _vp]),
    "hbs_shard_dot_partial": (_i, [_vp, _vp, _vp, _dp]),
    "hbs_shard_csr": (_i, [_vp, C.POINTER(_vp)]),
    "hbs_timer_start": (_i, [_vp]),
    "hbs_timer_stop": (_i, [_vp, _dp]),
    "hbs_bench_spmv_f64": (_i, [_vp, _vp, _vp, _i, _dp]),
    "hbs_bench_spmv_f32": (_i, [_vp, _vp, _vp, _i, _dp]),
    "hbs_bench_copy": (_i, [_vp, _vp, _vp, C.c_size_t, _i, _dp]),
    "hbs_set_spmv_plan": (_i, [_i, _i]),
    "hbs_set_spmv_groups": (_i, [_i]),
    "hbs_set_spmv_stream_nt": (_i, [_i]),
    "hbs_set_spmv_pass": (_i, [_i]),
    "hbs_set_vec_blocks": (_i, [_i]),
    "hbs_set_spmv_sort": (_i, [_i]),
    "hbs_csr_plan_info": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "hbs_set_spmv_xseg": (_i, [_i]),
    "hbs_set_xseg_split": (_i, [_i]),
    "hbs_set_spmm_order": (_i, [_i]),
    "hbs_set_spmv_block3": (_i, [_i]),
    "hbs_csr_block3": (_i, [_vp, _ip]),
    "hbs_detect_block3": (_i, [_i, _i, _ip, _ip, _ip]),
    "hbs_set_spmm_tuning": (_i, [_i, _i, _i, _i]),
    "hbs_set_block_tuning": (_i, [_i, _i]),
    "hbs_block_tuning": (_i, [_vp, _i, _ip, _ip, _ip]),
    "hbs_block_spmm_groups": (_i, [_vp, _i, _ip]),
    "hbs_set_block_spmm_groups": (_i, [_i]),
    "hbs_plan_grid_runs": (_i, [_i, _ip, _ip, _ip, _i, _ip, _ip]),
    "hbs_csr_spmm_runs": (_i, [_vp, _ip, _ip]),
    "hbs_csr_xseg": (_i, [_vp, _ip]),
    "hbs_set_spmv_diag_split": (_i, [_i]),
    "hbs_set_xseg_fuse": (_i, [_i]),
    "hbs_csr_diag_split": (_i, [_vp, _ip]),
    "hbs_shard_pack_host": (_i, [_vp, _vp, _dp]),
    "hbs_vector_part": (_i, [_i, _dp, _ip, _i, _i, _dp, _ip]),
    "hbs_vector_read_parted": (_i, [C.c_char_p, _i, _ip, _i, _i, _dpp, _ip]),
    "hbs_vector_gather": (_i, [_i, _dp, _ip, _i, _dp]),
    "hbs_set_wait_limit": (_i, [_d]),
    "hbs_shard_spmv2": (_i, [_vp, _vp, _dp, _vp, _d, _d, _vp]),
}

_lib = None


def lib():
    """Load libhbs.so once; raise if it has not been built.

    torch, where installed, is loaded first: it brings its own copy of the
    HIP runtime, and libhbs.so binds to that copy only if it is already in
    the process (the loader matches it by soname).  Loaded the other way
    round, the process holds two runtimes, and whichever opens the GPU
    second finds no device."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m hethyppar_amd.build`")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != HBS_OK:
        raise HBSError(rc, lib().hbs_last_error().decode())
    return rc


def _ptr(a):
    """Device address of a DeviceArray, a torch tensor or a plain integer."""
    if isinstance(a, (DeviceArray, DevPtr)):
        return a.ptr
    if hasattr(a, "data_ptr"):
        return _vp(a.data_ptr())
    return _vp(int(a))


class DevPtr:
    """A borrowed device address with the .ptr attribute the wrappers use
    (e.g. the storage of a torch tensor)."""

    def __init__(self, a):
        self.ptr = _ptr(a)


def _ip_of(a):
    return a.ctypes.data_as(_ip)


def _dp_of(a):
    return a.ctypes.data_as(_dp)


def _take(ptr, n, dtype):
    """Copy a library-malloc'd host array into numpy and free it."""
    if n == 0:
        out = np.zeros(0, dtype=dtype)
    else:
        out = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    lib().hbs_host_free(C.cast(ptr, _vp))
    return out


def _csr_out(call, *args):
    m = C.c_int()
    nnz = C.c_int()
    rp, cc, vv = _ip(), _ip(), _dp()
    check(call(*args, C.byref(m), C.byref(nnz), C.byref(rp), C.byref(cc), C.byref(vv)))
    return (_take(rp, m.value + 1, np.int32), _take(cc, nnz.value, np.int32),
            _take(vv, nnz.value, np.float64))


# ---- host-side helpers ------------------------------------------------------

def mm_read(path):
    """Matrix Market file -> (m, n, rowptr, col, val), per-row file order."""
    m, n, nnz = C.c_int(), C.c_int(), C.c_int()
    rp, cc, vv = _ip(), _ip(), _dp()
    check(lib().hbs_mm_read(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz),
                            C.byref(rp), C.byref(cc), C.byref(vv)))
    return (m.value, n.value, _take(rp, m.value + 1, np.int32), _take(cc, nnz.value, np.int32),
            _take(vv, nnz.value, np.float64))


def vector_read(path, size):
    out = np.zeros(size, dtype=np.float64)
    check(lib().hbs_vector_read(os.fsencode(path), size, _dp_of(out)))
    return out


def vector_write(path, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    check(lib().hbs_vector_write(os.fsencode(path), v.size, _dp_of(v)))


def gen_convdiff2d(nx, ny, conv):
    m = C.c_int()
    rp, cc, vv = _ip(), _ip(), _dp()
    nnz = C.c_int()
    check(lib().hbs_gen_convdiff2d(nx, ny, conv, C.byref(m), C.byref(nnz), C.byref(rp), C.byref(cc),
                                   C.byref(vv)))
    return (_take(rp, m.value + 1, np.int32), _take(cc, nnz.value, np.int32),
            _take(vv, nnz.value, np.float64))


def gen_convdiff3d(nx, ny, nz, conv):
    m = C.c_int()
    rp, cc, vv = _ip(), _ip(), _dp()
    nnz = C.c_int()
    check(lib().hbs_gen_convdiff3d(nx, ny, nz, conv, C.byref(m), C.byref(nnz), C.byref(rp),
                                   C.byref(cc), C.byref(vv)))
    return (_take(rp, m.value + 1, np.int32), _take(cc, nnz.value, np.int32),
            _take(vv, nnz.value, np.float64))


def gen_markov(n, kmin=12, kmax=22, band=None, dom=1.05, seed=2):
    if band is None:
        band = max(1, int(round(0.02 * n)))
    nnz = C.c_int()
    rp, cc, vv = _ip(), _ip(), _dp()
    check(lib().hbs_gen_markov(n, kmin, kmax, band, dom, seed, C.byref(nnz), C.byref(rp),
                               C.byref(cc), C.byref(vv)))
    return (_take(rp, n + 1, np.int32), _take(cc, nnz.value, np.int32),
            _take(vv, nnz.value, np.float64))


def gen_circuit(n, mean_nnz=4.2, zipf_s=2.1, nheavy=200, heavy_lo=1000, heavy_hi=50000,
                dom=1.1, seed=3, equilibrate=True):
    nnz = C.c_int()
    rp, cc, vv = _ip(), _ip(), _dp()
    check(lib().hbs_gen_circuit(n, mean_nnz, zipf_s, nheavy, heavy_lo, heavy_hi, dom,
                                1 if equilibrate else 0, seed,
                                C.byref(nnz), C.byref(rp), C.byref(cc), C.byref(vv)))
    return (_take(rp, n + 1, np.int32), _take(cc, nnz.value, np.int32),
            _take(vv, nnz.value, np.float64))


def gen_fem(nblockrows, nblk_per_row=12, dom=1.1, seed=6):
    nnz = C.c_int()
    rp, cc, vv = _ip(), _ip(), _dp()
    check(lib().hbs_gen_fem(nblockrows, nblk_per_row, dom, seed, C.byref(nnz), C.byref(rp),
                            C.byref(cc), C.byref(vv)))
    n = 3 * nblockrows
    return (_take(rp, n + 1, np.int32), _take(cc, nnz.value, np.int32),
            _take(vv, nnz.value, np.float64))


def gen_uniform(n, lo=-1.0, hi=1.0, seed=1):
    out = np.zeros(n, dtype=np.float64)
    check(lib().hbs_gen_uniform(n, lo, hi, seed, _dp_of(out)))
    return out


def gen_uneven_systems(nsys, lo_exp=3.0, hi_exp=5.0, seed=5):
    """Independent 2-D convection-diffusion systems of uneven size: system q
    has n_q = 10^U(lo_exp, hi_exp) rounded to a square grid (side >= 2) and
    convection c_q ~ U[0, 1] (diagonal 4 + c_q, west -1 - c_q, others -1).
    The uniforms come from the SplitMix64 stream of gen_uniform(seed).
    Returns [(rowptr, col, val, conv), ...]."""
    u = gen_uniform(2 * nsys, 0.0, 1.0, seed)
    out = []
    for q in range(nsys):
        side = max(2, int(round((10.0 ** (lo_exp + (hi_exp - lo_exp) * u[2 * q])) ** 0.5)))
        rp, c, v = gen_convdiff2d(side, side, float(u[2 * q + 1]))
        out.append((rp, c, v, float(u[2 * q + 1])))
    return out


def vector_part(big, partvec, nparts, rank):
    """Rank `rank`'s slice of a global vector (entries of its rows in
    ascending row order), as MPI_vector_read_parted hands it out."""
    big = np.ascontiguousarray(big, dtype=np.float64)
    pv = np.ascontiguousarray(partvec, dtype=np.int32)
    k = C.c_int()
    check(lib().hbs_vector_part(big.size, _dp_of(big), _ip_of(pv), nparts, rank, None, C.byref(k)))
    out = np.zeros(k.value, dtype=np.float64)
    check(lib().hbs_vector_part(big.size, _dp_of(big), _ip_of(pv), nparts, rank, _dp_of(out), C.byref(k)))
    return out


def vector_read_parted(path, size, partvec, nparts, rank):
    pv = np.ascontiguousarray(partvec, dtype=np.int32)
    out, k = _dp(), C.c_int()
    check(lib().hbs_vector_read_parted(os.fsencode(path), size, _ip_of(pv), nparts, rank, C.byref(out),
                                       C.byref(k)))
    return _take(out, k.value, np.float64)


def vector_gather(loc, gind, big):
    """big[gind[k]] = loc[k] for this rank's entries (in place; returns big)."""
    loc = np.ascontiguousarray(loc, dtype=np.float64)
    gind = np.ascontiguousarray(gind, dtype=np.int32)
    assert big.dtype == np.float64 and big.flags.c_contiguous
    check(lib().hbs_vector_gather(loc.size, _dp_of(loc), _ip_of(gind[:loc.size]), big.size, _dp_of(big)))
    return big


def set_wait_limit(seconds):
    check(lib().hbs_set_wait_limit(float(seconds)))


def host_spmv(rowptr, col, val, x):
    """y = A x summed in CSR order (host; used to form b = A x_target)."""
    m = rowptr.size - 1
    y = np.zeros(m, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    check(lib().hbs_host_csr_spmv(m, _ip_of(rowptr), _ip_of(col), _dp_of(val), _dp_of(x), _dp_of(y)))
    return y


def partition(rowptr, col, nparts, target=None, final_imbal=0.03, seed=42):
    """Row partition vector of a square CSR for the shard build
    (hbs_partition: the CalcPartVec model, this library's own partitioner).
    Returns (partvec int32[n], connectivity - 1 cut)."""
    n = rowptr.size - 1
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    pv = np.zeros(n, dtype=np.int32)
    cut = C.c_longlong()
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float64)
    if t is not None and t.size != nparts:
        raise ValueError("target needs one weight per part")
    check(lib().hbs_partition(n, _ip_of(rowptr), _ip_of(col), nparts, None if t is None else _dp_of(t),
                              float(final_imbal), int(seed), _ip_of(pv), C.byref(cut)))
    return pv, cut.value


def shard_build(rowptr, col, val, partvec, nparts, rank):
    """One rank's SHARD_CSC (hbs_shard_build_host) as a dict of numpy arrays
    plus the raw structure (kept alive in the dict under 'raw')."""
    n = rowptr.size - 1
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    partvec = np.ascontiguousarray(partvec, dtype=np.int32)
    raw = HostShard()
    check(lib().hbs_shard_build_host(n, _ip_of(rowptr), _ip_of(col), _dp_of(val), _ip_of(partvec),
                                     nparts, rank, C.byref(raw)))
    return shard_to_numpy(raw)


def _np(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def shard_to_numpy(raw):
    def csc(c):
        return {"I": _np(c.I, c.n + 1, np.int32), "J": _np(c.J, c.nnz, np.int32),
                "val": _np(c.val, c.nnz, np.float64), "nnz": c.nnz, "m": c.m, "n": c.n}

    def comm(c):
        tot = _np(c.I, c.num + 1, np.int32)
        cnt = int(tot[-1]) if c.num > 0 else 0
        return {"I": tot, "J": _np(c.J, cnt, np.int32), "ranks": _np(c.ranks, c.num, np.int32),
                "num": c.num}

    d = {"loc": csc(raw.loc), "shr": csc(raw.shr), "gind": _np(raw.gind.vals, raw.gind.nvals, np.int32),
         "send": comm(raw.send), "recv": comm(raw.recv), "raw": raw}
    return d


def shard_free(d):
    raw = d.get("raw")
    if raw is not None:
        lib().hbs_host_shard_free(C.byref(raw))
        d["raw"] = None


def device_count():
    n = C.c_int()
    check(lib().hbs_device_count(C.byref(n)))
    return n.value


# ---- device objects -----------------------------------------------------------

class Context:
    """One device + one stream (hbs_ctx)."""

    def __init__(self, device=0, stream=None):
        h = _vp()
        check(lib().hbs_ctx_create(device, stream, C.byref(h)))
        self.h = h
        self._children = []

    def _adopt(self, obj):
        self._children.append(weakref.ref(obj))

    def close(self):
        """Release every object created on this context (newest first), then
        the context itself."""
        if self.h:
            for r in reversed(self._children):
                o = r()
                if o is not None:
                    o.close()
            self._children = []
            lib().hbs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().hbs_ctx_sync(self.h))

    def empty(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.size, a.dtype)
        d.upload(a)
        return d

    def dot(self, x, y):
        out = C.c_double()
        check(lib().hbs_dot_f64(self.h, x.n, x.ptr, y.ptr, C.byref(out)))
        return out.value

    def timer_start(self):
        check(lib().hbs_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_double()
        check(lib().hbs_timer_stop(self.h, C.byref(ms)))
        return ms.value


class PinnedArray:
    """Page-locked host array (hbs_pinned_alloc) seen as a numpy array."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        p = _vp()
        check(lib().hbs_pinned_alloc(ctx.h, max(int(n), 1) * self.dtype.itemsize, C.byref(p)))
        self.ptr = p
        buf = (C.c_char * (max(int(n), 1) * self.dtype.itemsize)).from_address(p.value)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=int(n))
        ctx._adopt(self)

    def close(self):
        if self.ptr and self.ctx.h:
            lib().hbs_pinned_free(self.ctx.h, self.ptr)
        self.ptr = None
        self.a = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = _vp()
        check(lib().hbs_dev_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p
        ctx._adopt(self)

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().hbs_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(_vp), self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().hbs_d2h(self.ctx.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def zero(self):
        check(lib().hbs_dev_zero(self.ctx.h, self.ptr, self.nbytes))

    def offset(self, k):
        """Raw device pointer k elements in."""
        return _vp(self.ptr.value + k * self.dtype.itemsize)

    def free(self):
        if self.ptr and self.ctx.h:
            lib().hbs_dev_free(self.ctx.h, self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def concat_systems(systems):
    """Block-diagonal CSR of independent square systems [(rowptr, col, val),
    ...] with each system's columns shifted to its rows; returns (rowptr,
    col, val, sys_rows) for DeviceCSR(..., sys_rows=sys_rows)."""
    rps, cols, vals, sys_rows = [np.zeros(1, np.int64)], [], [], [0]
    nz = 0
    for rp, c, v in systems:
        rp = np.asarray(rp, dtype=np.int64)
        rps.append(rp[1:] + nz)
        cols.append(np.asarray(c, dtype=np.int64) + sys_rows[-1])
        vals.append(np.asarray(v, dtype=np.float64))
        nz += int(rp[-1])
        sys_rows.append(sys_rows[-1] + rp.size - 1)
    return (np.concatenate(rps).astype(np.int32), np.concatenate(cols).astype(np.int32),
            np.concatenate(vals), np.asarray(sys_rows, dtype=np.int32))


class DeviceCSR:
    def __init__(self, ctx, rowptr, col, val, ncols=None, flags=0, sys_rows=None):
        self.ctx = ctx
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        m = rowptr.size - 1
        n = m if ncols is None else ncols
        h = _vp()
        if sys_rows is None:
            check(lib().hbs_csr_create(ctx.h, m, n, col.size, _ip_of(rowptr), _ip_of(col), _dp_of(val),
                                       flags, C.byref(h)))
            self.nsys = 1
        else:
            sys_rows = np.ascontiguousarray(sys_rows, dtype=np.int32)
            check(lib().hbs_csr_create_systems(ctx.h, sys_rows.size - 1, _ip_of(sys_rows), col.size,
                                               _ip_of(rowptr), _ip_of(col), _dp_of(val), flags,
                                               C.byref(h)))
            self.nsys = sys_rows.size - 1
        self.h = h
        self.m, self.n, self.nnz = m, n, int(col.size)
        ctx._adopt(self)

    @classmethod
    def wrap(cls, ctx, handle, m, n, nnz):
        obj = cls.__new__(cls)
        obj.ctx, obj.h, obj.m, obj.n, obj.nnz = ctx, handle, m, n, nnz
        obj._borrowed = True
        return obj

    def info(self):
        v = [C.c_int() for _ in range(5)]
        check(lib().hbs_csr_info(self.h, *[C.byref(x) for x in v]))
        d = dict(zip(["m", "n", "nnz", "nblocks", "serial_rows"], [x.value for x in v]))
        w = [C.c_int() for _ in range(4)]
        check(lib().hbs_csr_plan_info(self.h, *[C.byref(x) for x in w]))
        d.update(zip(["ngroups", "nchunks", "block_nnz", "max_row"], [x.value for x in w]))
        ns = C.c_int()
        check(lib().hbs_csr_xseg(self.h, C.byref(ns)))
        d["xseg"] = ns.value
        b3 = C.c_int()
        check(lib().hbs_csr_block3(self.h, C.byref(b3)))
        d["block3"] = b3.value
        ds = C.c_int()
        check(lib().hbs_csr_diag_split(self.h, C.byref(ds)))
        d["diag_split"] = ds.value
        nr, ln = C.c_int(), C.c_int()
        check(lib().hbs_csr_spmm_runs(self.h, C.byref(nr), C.byref(ln)))
        d["spmm_runs"], d["spmm_run_len"] = nr.value, ln.value
        return d

    def spmv(self, x, y, alpha=1.0, beta=0.0):
        if x.dtype == np.float32:
            check(lib().hbs_csr_spmv_f32(self.h, alpha, x.ptr, beta, y.ptr))
        else:
            check(lib().hbs_csr_spmv_f64(self.h, alpha, x.ptr, beta, y.ptr))

    def set_values(self, val_dev):
        """New values (same pattern) from a device array of nnz doubles."""
        check(lib().hbs_csr_set_values(self.h, _ptr(val_dev)))

    def spmv_t(self, x, y):
        check(lib().hbs_csr_spmv_t_f64(self.h, x.ptr, y.ptr))

    def spmm(self, X, Y, k, tile=0):
        if X.dtype == np.float32:
            check(lib().hbs_csr_spmm_f32(self.h, k, X.ptr, Y.ptr, tile))
        else:
            check(lib().hbs_csr_spmm_f64(self.h, k, X.ptr, Y.ptr, tile))

    def bench_spmv(self, x, y, reps):
        ms = C.c_double()
        if x.dtype == np.float32:
            check(lib().hbs_bench_spmv_f32(self.h, x.ptr, y.ptr, reps, C.byref(ms)))
        else:
            check(lib().hbs_bench_spmv_f64(self.h, x.ptr, y.ptr, reps, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h and not getattr(self, "_borrowed", False) and self.ctx.h:
            lib().hbs_csr_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_opts(**kw):
    o = SolveOpts()
    lib().hbs_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Solver:
    def __init__(self, A):
        self.A = A
        h = _vp()
        check(lib().hbs_solver_create(A.h, C.byref(h)))
        self.h = h
        A.ctx._adopt(self)

    def _run(self, fn, b, x, opts, raise_on_error):
        info = SolveInfo()
        rc = fn(self.h, b.ptr, x.ptr, C.byref(opts), C.byref(info))
        if rc != HBS_OK and raise_on_error:
            check(rc)
        return info

    def bicgstab(self, b, x, opts=None, raise_on_error=True, **kw):
        return self._run(lib().hbs_bicgstab_f64, b, x, opts or default_opts(**kw), raise_on_error)

    def bicgstab_mixed(self, b, x, opts=None, raise_on_error=True, **kw):
        return self._run(lib().hbs_bicgstab_mixed, b, x, opts or default_opts(**kw), raise_on_error)

    def bicgstab_systems(self, b, x, opts=None, raise_on_error=True, **kw):
        """Every system of a DeviceCSR(..., sys_rows=...) at once; one
        SolveInfo per system."""
        infos = (SolveInfo * self.A.nsys)()
        o = opts or default_opts(**kw)
        rc = lib().hbs_bicgstab_systems_f64(self.h, b.ptr, x.ptr, C.byref(o), infos)
        if rc != HBS_OK and raise_on_error:
            check(rc)
        return list(infos)

    def bicgstab_block(self, B, X, k, opts=None, raise_on_error=True, **kw):
        """k right-hand sides, B and X row-major n x k on the device; one
        independent BiCGStab per column sharing every SpMM pass over A."""
        infos = (SolveInfo * k)()
        o = opts or default_opts(**kw)
        rc = lib().hbs_bicgstab_block_f64(self.h, k, B.ptr, X.ptr, C.byref(o), infos)
        if rc != HBS_OK and raise_on_error:
            check(rc)
        return list(infos)

    def block_tuning(self, k):
        """(tile, rows per lane step, rows per wavefront) the block solve's
        autotuner chose for k right-hand sides."""
        t, r, w = C.c_int(), C.c_int(), C.c_int()
        check(lib().hbs_block_tuning(self.h, k, C.byref(t), C.byref(r), C.byref(w)))
        return t.value, r.value, w.value

    def block_spmm_groups(self, k):
        """Workgroups of the block solve's SpMMs for k right-hand sides."""
        g = C.c_int()
        check(lib().hbs_block_spmm_groups(self.h, k, C.byref(g)))
        return g.value

    def adjoint(self, g, x, lam, grad_val, opts=None, raise_on_error=True, **kw):
        info = SolveInfo()
        o = opts or default_opts(**kw)
        rc = lib().hbs_adjoint_f64(self.h, g.ptr, x.ptr, lam.ptr,
                                   grad_val.ptr if grad_val is not None else None,
                                   C.byref(o), C.byref(info))
        if rc != HBS_OK and raise_on_error:
            check(rc)
        return info

    def close(self):
        if self.h and self.A.h and self.A.ctx.h:
            lib().hbs_solver_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Shard:
    """One rank's share on the device (hbs_shard)."""

    def __init__(self, ctx, host_shard):
        self.ctx = ctx
        h = _vp()
        check(lib().hbs_shard_create(ctx.h, C.byref(host_shard["raw"]), C.byref(h)))
        self.h = h
        v = [C.c_int() for _ in range(4)]
        check(lib().hbs_shard_info(h, *[C.byref(x) for x in v]))
        self.m_loc, self.n_shr, self.n_send, self.nnz = [x.value for x in v]
        ctx._adopt(self)

    def pack(self, x_loc, send):
        check(lib().hbs_shard_pack(self.h, x_loc.ptr, send.ptr))

    def unpack(self, recv, x_ext):
        check(lib().hbs_shard_unpack(self.h, recv.ptr, x_ext.ptr))

    def spmv(self, x_ext, y, recv_host=None):
        rh = None
        if recv_host is not None:
            recv_host = np.ascontiguousarray(recv_host, dtype=np.float64)
            self._keep = recv_host
            rh = _dp_of(recv_host)
        check(lib().hbs_shard_spmv(self.h, x_ext.ptr, rh, y.ptr))

    def dot_partial(self, a, b):
        out = C.c_double()
        check(lib().hbs_shard_dot_partial(self.h, a.ptr, b.ptr, C.byref(out)))
        return out.value

    def csr(self):
        h = _vp()
        check(lib().hbs_shard_csr(self.h, C.byref(h)))
        return DeviceCSR.wrap(self.ctx, h, self.m_loc, self.m_loc + self.n_shr, self.nnz)

    def close(self):
        if self.h and self.ctx.h:
            lib().hbs_shard_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_spmv_plan(block_nnz=0, serial_max_avg=64, max_groups=0, stream_nt=0, npt=0):
    """Process-wide plan knobs used by DeviceCSR objects created afterwards."""
    check(lib().hbs_set_spmv_plan(block_nnz, serial_max_avg))
    check(lib().hbs_set_spmv_groups(max_groups))
    check(lib().hbs_set_spmv_stream_nt(stream_nt))
    check(lib().hbs_set_spmv_pass(npt))


def set_xseg_split(split=1):
    check(lib().hbs_set_xseg_split(split))


def set_spmv_block3(mode=1):
    """Process-wide: 3x3-block streams for 3x3-block matrices created
    afterwards (1) or the plain CSR streams (0)."""
    check(lib().hbs_set_spmv_block3(mode))


def detect_block3(rowptr, col, n=None):
    """Whether a host CSR is a 3x3-block matrix (kind-3 streams)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    c = np.ascontiguousarray(col, dtype=np.int32)
    m = rp.size - 1
    y = C.c_int()
    check(lib().hbs_detect_block3(m, m if n is None else n, rp.ctypes.data_as(_ip), c.ctypes.data_as(_ip),
                                  C.byref(y)))
    return bool(y.value)


def set_spmm_order(mode=1):
    """Process-wide: multi-column SpMM traversal for DeviceCSR objects created
    afterwards (1 y-slab runs on a 3-D grid operator, 0 rows in order)."""
    check(lib().hbs_set_spmm_order(mode))


def set_block_tuning(tile=0, rows_per_step=0):
    """Process-wide: force the block solve's column-group width and SpMM rows
    per lane step for workspaces created afterwards (0: tuned)."""
    check(lib().hbs_set_block_tuning(tile, rows_per_step))


def set_block_spmm_groups(groups=0):
    """Process-wide: workgroups of the block solve's SpMMs for workspaces
    created afterwards (0: tuned; else a multiple of 8)."""
    check(lib().hbs_set_block_spmm_groups(groups))


def set_spmm_tuning(slab_lines=0, groups=0, dynamic=16, wide_loads=1):
    """Process-wide multi-column SpMM tuning: slab height of the grid runs
    and step dealing (counters per XCD, 0 for a fixed stride; DeviceCSR
    objects created afterwards), workgroups per launch, 16-byte X gathers."""
    check(lib().hbs_set_spmm_tuning(slab_lines, groups, dynamic, wide_loads))


def plan_grid_runs(rowptr, col):
    """(runs, len) the SpMM traversal plans for a host CSR; runs is empty
    when the matrix is not a 3-D grid operator that needs it."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    c = np.ascontiguousarray(col, dtype=np.int32)
    m = rp.size - 1
    out = np.empty(max(m, 1), dtype=np.int32)
    nr, ln = C.c_int(), C.c_int()
    check(lib().hbs_plan_grid_runs(m, rp.ctypes.data_as(_ip), c.ctypes.data_as(_ip), out.ctypes.data_as(_ip),
                                   int(out.size), C.byref(nr), C.byref(ln)))
    return out[:nr.value].copy(), ln.value


def set_xseg_fuse(mode=1):
    """Process-wide: solvers created afterwards update P and S inside the
    x-segmented product passes (1) or in their own vector passes (0)."""
    check(lib().hbs_set_xseg_fuse(mode))


def set_spmv_diag_split(mode=1):
    """Process-wide: FP32 SpMVs of square x-segmented DeviceCSR objects
    created afterwards keep the diagonal apart (1) or in the entries (0)."""
    check(lib().hbs_set_spmv_diag_split(mode))


def set_spmv_xseg(mode=2):
    """Process-wide: x-segmented SpMV for DeviceCSR objects created
    afterwards (0 never, 1 every one-system matrix, 2 scattered columns)."""
    check(lib().hbs_set_spmv_xseg(mode))


def set_vec_blocks(target=512):
    """Process-wide: vector workgroups a Solver created afterwards aims at."""
    check(lib().hbs_set_vec_blocks(target))


def set_spmv_sort(max_line_pct=70):
    """Process-wide: column-sorted gathers for row blocks created afterwards
    whose sorted entries touch at most this percentage of lines (0: off)."""
    check(lib().hbs_set_spmv_sort(max_line_pct))
