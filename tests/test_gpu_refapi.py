"""The reference-shaped device API (include/hhp_hip.h, include/hhp_dp_kernels.h)
called the way the reference's drivers call it (structs by value), checked
against the oracle and against numpy's element-wise arithmetic."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as orc
from conftest import golden
from hethyppar_amd import hbs
from workloads import convdiff64, partition

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


class Vector(C.Structure):
    _fields_ = [("vals", _dp), ("nvals", C.c_uint)]


class DeviceVector(C.Structure):
    _fields_ = [("desc", C.c_void_p), ("vals", _dp), ("nvals", C.c_uint)]


class DeviceCSR(C.Structure):
    _fields_ = [("desc", C.c_void_p), ("data", hbs.HostCSR)]


class DeviceShard(C.Structure):
    _fields_ = [("loc", DeviceCSR), ("shr", DeviceCSR), ("gind", hbs.HostIVec),
                ("send", hbs.HostComm), ("recv", hbs.HostComm), ("dev", C.c_void_p)]


def L():
    lib = hbs.lib()
    lib.device_csr_create.argtypes = [hbs.HostCSR, C.POINTER(DeviceCSR)]
    lib.device_csr_destroy.argtypes = [C.POINTER(DeviceCSR)]
    lib.device_vector_init.argtypes = [C.c_size_t, C.POINTER(DeviceVector)]
    lib.device_vector_destroy.argtypes = [C.POINTER(DeviceVector)]
    lib.device_vector_toGPU.argtypes = [Vector, DeviceVector]
    lib.device_vector_toCPU.argtypes = [DeviceVector, Vector]
    lib.device_csr_spmv.argtypes = [C.c_void_p, DeviceCSR, DeviceVector, DeviceVector, C.c_double,
                                    C.c_double, C.c_void_p]
    lib.device_vector_dot.argtypes = [C.c_void_p, DeviceVector, DeviceVector, _dp]
    lib.device_vector_axpy.argtypes = [C.c_void_p, DeviceVector, C.c_double, DeviceVector]
    lib.device_vector_scale.argtypes = [C.c_void_p, C.c_double, DeviceVector]
    lib.device_SHARD_CSC_create.argtypes = [hbs.HostShard, C.POINTER(DeviceShard)]
    lib.device_SHARD_CSC_destroy.argtypes = [C.POINTER(DeviceShard)]
    lib.device_SHARD_CSC_spmxv.argtypes = [DeviceShard, Vector, DeviceVector, DeviceVector, DeviceVector,
                                           C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_void_p]
    lib.hhp_default_handle.restype = C.c_void_p
    for f in ("hhp_dp_vecop_P", "hhp_dp_vecop_S", "hhp_dp_vecop_XR", "hhp_dp_update_beta",
              "hhp_dp_update_alpha", "hhp_dp_update_omega", "hhp_dp_gather", "hhp_dp_add"):
        getattr(lib, f).restype = None
    return lib


def host_vec(a):
    return Vector(a.ctypes.data_as(_dp), a.size)


def dvec(lib, a):
    v = DeviceVector()
    assert lib.device_vector_init(a.size, C.byref(v)) == 0
    assert lib.device_vector_toGPU(host_vec(a), v) == 0
    return v


def down(lib, v):
    out = np.zeros(v.nvals)
    assert lib.device_vector_toCPU(v, host_vec(out)) == 0
    return out


def test_device_csr_spmv_and_blas1():
    lib = L()
    rp, c, v, xt, b = convdiff64()
    csr = hbs.HostCSR(rp.ctypes.data_as(C.POINTER(C.c_int)), c.ctypes.data_as(C.POINTER(C.c_int)),
                      v.ctypes.data_as(_dp), c.size, rp.size - 1, rp.size - 1)
    dA = DeviceCSR()
    assert lib.device_csr_create(csr, C.byref(dA)) == 0
    assert dA.data.nnz == c.size and dA.data.m == rp.size - 1
    dx = dvec(lib, xt)
    dy = dvec(lib, np.zeros(rp.size - 1))
    assert lib.device_csr_spmv(None, dA, dx, dy, 1.0, 0.0, None) == 0
    assert np.array_equal(down(lib, dy), orc.spmv(orc.CSR(rp, c, v), xt))
    d = C.c_double()
    assert lib.device_vector_dot(None, dy, dx, C.byref(d)) == 0
    y = down(lib, dy)
    assert abs(d.value - y @ xt) <= 1e-12 * (np.abs(y) @ np.abs(xt))
    assert lib.device_vector_axpy(None, dx, -0.5, dy) == 0        # dy += -0.5 dx
    assert lib.device_vector_scale(None, 3.0, dy) == 0            # dy *= 3
    assert np.allclose(down(lib, dy), 3.0 * (y - 0.5 * xt), rtol=1e-15, atol=1e-15)
    for vv in (dx, dy):
        assert lib.device_vector_destroy(C.byref(vv)) == 0
    assert lib.device_csr_destroy(C.byref(dA)) == 0


def test_dp_kernels_match_elementwise_arithmetic():
    lib = L()
    n = 10007
    rng = np.random.default_rng(1)
    P, V, R, S, T, X = (rng.standard_normal(n) for _ in range(6))
    sc = np.array([0.7, -1.3, 2.5, 0.4, 1.9, -0.2, 0.0, 0.0, 0.0, 0.0])
    dP, dV, dR, dS, dT, dX = (dvec(lib, a) for a in (P, V, R, S, T, X))
    dsc = dvec(lib, sc)
    st = C.c_void_p(hbs.lib().hbs_ctx_stream(lib.hhp_default_handle()))   # same stream as the copies
    p = lambda k: C.cast(C.c_void_p(C.cast(dsc.vals, C.c_void_p).value + 8 * k), _dp)
    om, be, na = p(0), p(1), p(2)
    lib.hhp_dp_vecop_P(dP.vals, dV.vals, dR.vals, om, be, n, st)
    lib.hhp_dp_vecop_S(dS.vals, dR.vals, dV.vals, na, n, st)
    al, om2, nom = p(3), p(4), p(5)
    lib.hhp_dp_vecop_XR(dX.vals, dR.vals, dP.vals, dS.vals, dT.vals, al, om2, nom, n, st)
    Pn = (P - sc[0] * V) * sc[1] + R
    Sn = R + sc[2] * V
    Xn = X + (sc[3] * Pn + sc[4] * Sn)
    Rn = Sn + sc[5] * T
    assert np.array_equal(down(lib, dP), Pn)
    assert np.array_equal(down(lib, dS), Sn)
    assert np.array_equal(down(lib, dX), Xn)
    assert np.array_equal(down(lib, dR), Rn)
    # beta = (rho_new/rho_old)*(alpha/omega); rho_old <- rho_new
    s2 = np.array([0.0, 3.0, 6.0, 0.5, 0.25, 0.0, 0.0])
    d2 = dvec(lib, s2)
    q = lambda k: C.cast(C.c_void_p(C.cast(d2.vals, C.c_void_p).value + 8 * k), _dp)
    lib.hhp_dp_update_beta(q(0), q(1), q(2), q(3), q(4), st)
    lib.hhp_dp_update_alpha(q(5), q(6), q(1), q(4), st)
    out = down(lib, d2)
    assert out[0] == (6.0 / 3.0) * (0.5 / 0.25) and out[1] == 6.0
    assert out[5] == 6.0 / 0.25 and out[6] == -(6.0 / 0.25)
    idx = rng.integers(0, n, 999).astype(np.int32)
    didx = dvec(lib, idx.astype(np.float64))   # room for 999 ints
    hbs.check(hbs.lib().hbs_h2d(lib.hhp_default_handle(), C.cast(didx.vals, C.c_void_p),
                                idx.ctypes.data_as(C.c_void_p), idx.nbytes))
    dout = dvec(lib, np.zeros(999))
    lib.hhp_dp_gather(dout.vals, dR.vals, C.cast(didx.vals, C.POINTER(C.c_int)), 999, st)
    assert np.array_equal(down(lib, dout), Rn[idx])


def test_device_shard_spmxv_host_halo():
    lib = L()
    _, _, rp, c, v = hbs.mm_read(golden("sample320", "A.mtx"))
    pv = partition("sample320", "row_part_16")
    x = np.random.default_rng(2).uniform(-1, 1, 320)
    for r in (0, 7, 15):
        hs = hbs.shard_build(rp, c, v, pv, 16, r)
        raw = hs["raw"]
        dA = DeviceShard()
        assert lib.device_SHARD_CSC_create(raw, C.byref(dA)) == 0
        m, ns = hs["loc"]["m"], hs["shr"]["n"]
        x_loc = x[hs["gind"][:m]]
        recv_buf = x[hs["gind"][m:][hs["recv"]["J"]]]
        for k in range(ns):      # the halo as MPI would have delivered it
            raw.recv.val[k] = recv_buf[k]
        dX = dvec(lib, x_loc)
        dXs = dvec(lib, np.zeros(max(ns, 1)))
        dY = dvec(lib, np.zeros(m))
        assert lib.device_SHARD_CSC_spmxv(dA, Vector(), dX, dXs, dY, None, 1.0, 0.0, None, None, None) == 0
        assert np.array_equal(down(lib, dY), orc.shard_spmv(hs, x_loc, recv_buf))
        nsend = int(hs["send"]["I"][-1]) if hs["send"]["num"] else 0
        sent = np.array([raw.send.val[k] for k in range(nsend)])
        assert np.array_equal(sent, orc.shard_pack(hs, x_loc))
        if ns:
            assert np.array_equal(down(lib, dXs)[:ns], x[hs["gind"][m:]])
        assert lib.device_SHARD_CSC_destroy(C.byref(dA)) == 0
        hbs.shard_free(hs)


def _csc_struct(d):
    """hbs.HostCSR view of a SHARD-convention CSC dict (arrays kept alive in d)."""
    ip = C.POINTER(C.c_int)
    for k in ("I", "J", "val"):
        d[k] = np.ascontiguousarray(d[k])
    return hbs.HostCSR(d["I"].ctypes.data_as(ip), d["J"].ctypes.data_as(ip), d["val"].ctypes.data_as(_dp),
                       d["nnz"], d["m"], d["n"])


def _L2():
    lib = L()
    lib.device_csc_create.argtypes = [hbs.HostCSR, C.POINTER(DeviceCSR)]
    lib.device_csc_destroy.argtypes = [C.POINTER(DeviceCSR)]
    lib.device_csc_spmv.argtypes = [C.c_void_p, DeviceCSR, DeviceVector, DeviceVector, C.c_double, C.c_double,
                                    C.c_void_p]
    lib.device_csr_spmv_async.argtypes = lib.device_csc_spmv.argtypes
    lib.device_vector_zero.argtypes = [DeviceVector]
    lib.device_vector_GPUtoGPU.argtypes = [DeviceVector, DeviceVector]
    return lib


@pytest.mark.parametrize("rank", [0, 7, 15])
def test_device_csc_blocks_and_reference_built_shard(rank):
    """device_csc_create / device_csc_spmv on the loc and shr blocks of a
    SHARD_CSC, and MPI_device_SHARD_CSC_mpi_spmxv on a shard built the way
    bicgstab_mpi_gpu.c:248-257 builds it (gind/send/recv borrowed, loc and
    shr from device_csc_create).  Each block's SpMV sums a row over its
    columns in ascending order (SHARD_loc_CSC_to_CSR, hhp_matrix.c:
    1441-1474), so it equals the oracle's CSR SpMV of that block bit for bit;
    the shard SpMV follows hhp_cuda.c:259-352: Y = 0, Y = alpha*loc*X +
    beta*Y, Y = alpha*shr*x_shr + alpha*Y."""
    lib = _L2()
    _, _, rp, c, v = hbs.mm_read(golden("sample320", "A.mtx"))
    pv = partition("sample320", "row_part_16")
    x = np.random.default_rng(4 + rank).uniform(-1, 1, 320)
    hs = hbs.shard_build(rp, c, v, pv, 16, rank)
    raw = hs["raw"]
    m, ns = hs["loc"]["m"], hs["shr"]["n"]
    x_loc = x[hs["gind"][:m]]
    x_shr = x[hs["gind"][m:]]
    recv_buf = x_shr[hs["recv"]["J"]]
    dloc = DeviceCSR()
    assert lib.device_csc_create(raw.loc, C.byref(dloc)) == 0
    assert (dloc.data.m, dloc.data.n, dloc.data.nnz) == (m, hs["loc"]["n"], hs["loc"]["nnz"])
    loc_csr = orc.CSR(*orc.shard_csc_to_csr(hs["loc"]["I"], hs["loc"]["J"], hs["loc"]["val"], m, m))
    dX, dY = dvec(lib, x_loc), dvec(lib, np.full(m, 3.0))
    assert lib.device_csc_spmv(None, dloc, dX, dY, 1.0, 0.0, None) == 0
    y_loc = orc.spmv(loc_csr, x_loc)
    assert np.array_equal(down(lib, dY), y_loc)
    dshr = DeviceCSR()
    if ns:
        assert lib.device_csc_create(raw.shr, C.byref(dshr)) == 0
        shr_csr = orc.CSR(*orc.shard_csc_to_csr(hs["shr"]["I"], hs["shr"]["J"], hs["shr"]["val"], m, ns), ncols=ns)
        dXs = dvec(lib, x_shr)
        dYs = dvec(lib, np.zeros(m))
        assert lib.device_csc_spmv(None, dshr, dXs, dYs, 1.0, 0.0, None) == 0
        y_shr = orc.spmv(shr_csr, x_shr)
        assert np.array_equal(down(lib, dYs), y_shr)
    else:
        y_shr = np.zeros(m)
        dXs = dvec(lib, np.zeros(1))
    # the shard as the reference driver builds it: no merged form (dev NULL)
    dA = DeviceShard()
    dA.loc, dA.shr, dA.gind, dA.send, dA.recv, dA.dev = dloc, dshr, raw.gind, raw.send, raw.recv, None
    for k in range(ns):
        raw.recv.val[k] = recv_buf[k]
    for alpha, beta in ((1.0, 0.0), (2.0, 0.5)):
        dZ = dvec(lib, np.full(m, 7.0))           # zeroed by the call, as in the reference
        dXs2 = dvec(lib, np.zeros(max(ns, 1)))
        assert lib.device_SHARD_CSC_spmxv(dA, Vector(), dX, dXs2, dZ, None, alpha, beta, None, None, None) == 0
        want = alpha * y_shr + alpha * (alpha * y_loc)
        assert np.array_equal(down(lib, dZ), want)
        if ns:
            assert np.array_equal(down(lib, dXs2)[:ns], x_shr)
        nsend = int(hs["send"]["I"][-1]) if hs["send"]["num"] else 0
        assert np.array_equal(np.array([raw.send.val[k] for k in range(nsend)]), orc.shard_pack(hs, x_loc))
        for vv in (dZ, dXs2):
            assert lib.device_vector_destroy(C.byref(vv)) == 0
    # with alpha = 1 and beta = 0 the two-SpMV form is within rounding of the
    # CPU halo SpMV (loc part and shr part summed separately, then added)
    ys = orc.shard_spmv(hs, x_loc, recv_buf)
    assert np.allclose(y_shr + y_loc, ys, rtol=1e-14, atol=1e-14)
    assert lib.device_csc_destroy(C.byref(dloc)) == 0
    if ns:
        assert lib.device_csc_destroy(C.byref(dshr)) == 0
    hbs.shard_free(hs)


def test_merged_shard_spmxv_alpha_beta():
    """The merged shard (device_SHARD_CSC_create): Y = alpha*(loc*X +
    shr*x_shr) + beta*Y on the caller's Y, the row sum bit-identical to the
    CPU halo SpMV (hhp_cpu.c:514-601)."""
    lib = L()
    _, _, rp, c, v = hbs.mm_read(golden("sample320", "A.mtx"))
    pv = partition("sample320", "row_part_16")
    x = np.random.default_rng(9).uniform(-1, 1, 320)
    hs = hbs.shard_build(rp, c, v, pv, 16, 3)
    raw = hs["raw"]
    m, ns = hs["loc"]["m"], hs["shr"]["n"]
    x_loc = x[hs["gind"][:m]]
    recv_buf = x[hs["gind"][m:][hs["recv"]["J"]]]
    for k in range(ns):
        raw.recv.val[k] = recv_buf[k]
    dA = DeviceShard()
    assert lib.device_SHARD_CSC_create(raw, C.byref(dA)) == 0
    y0 = np.random.default_rng(10).standard_normal(m)
    dX, dXs, dY = dvec(lib, x_loc), dvec(lib, np.zeros(max(ns, 1))), dvec(lib, y0)
    assert lib.device_SHARD_CSC_spmxv(dA, Vector(), dX, dXs, dY, None, 2.0, 0.5, None, None, None) == 0
    ys = orc.shard_spmv(hs, x_loc, recv_buf)
    assert np.array_equal(down(lib, dY), 2.0 * ys + 0.5 * y0)
    assert lib.device_SHARD_CSC_destroy(C.byref(dA)) == 0
    hbs.shard_free(hs)


def test_device_vector_zero_copy_and_async_spmv():
    lib = _L2()
    rp, c, v, xt, b = convdiff64()
    n = rp.size - 1
    csr = hbs.HostCSR(rp.ctypes.data_as(C.POINTER(C.c_int)), c.ctypes.data_as(C.POINTER(C.c_int)),
                      v.ctypes.data_as(_dp), c.size, n, n)
    dA = DeviceCSR()
    assert lib.device_csr_create(csr, C.byref(dA)) == 0
    y0 = np.random.default_rng(3).standard_normal(n)
    dx, dy, dz = dvec(lib, xt), dvec(lib, y0), dvec(lib, np.ones(n))
    assert lib.device_vector_GPUtoGPU(dy, dz) == 0
    assert np.array_equal(down(lib, dz), y0)
    # asynchronous: queued on the handle's stream, the caller synchronises
    assert lib.device_csr_spmv_async(None, dA, dx, dy, -1.5, 0.25, None) == 0
    assert hbs.lib().hbs_ctx_sync(lib.hhp_default_handle()) == 0
    ax = orc.spmv(orc.CSR(rp, c, v), xt)
    assert np.array_equal(down(lib, dy), -1.5 * ax + 0.25 * y0)
    assert lib.device_vector_zero(dz) == 0
    assert np.all(down(lib, dz) == 0.0)
    for vv in (dx, dy, dz):
        assert lib.device_vector_destroy(C.byref(vv)) == 0
    assert lib.device_csr_destroy(C.byref(dA)) == 0


def test_dp_update_omega_and_add():
    lib = L()
    st = C.c_void_p(hbs.lib().hbs_ctx_stream(lib.hhp_default_handle()))
    s = np.array([0.0, 0.0, 3.7, -1.1, 0.0, 2.5, 0.375])
    d = dvec(lib, s)
    q = lambda k: C.cast(C.c_void_p(C.cast(d.vals, C.c_void_p).value + 8 * k), _dp)
    lib.hhp_dp_update_omega(q(0), q(1), q(2), q(3), st)   # omega = ts / tt ; neg_omega = -omega
    lib.hhp_dp_add(q(4), q(5), q(6), st)                    # total = a + b
    out = down(lib, d)
    assert out[0] == 3.7 / -1.1 and out[1] == -(3.7 / -1.1)
    assert out[4] == 2.5 + 0.375
    assert lib.device_vector_destroy(C.byref(d)) == 0


@pytest.mark.parametrize("name,x", [("sample", "X_mul.txt"), ("sample320", None)])
def test_plain_c_caller_runs_reference_harness_sequence(tmp_path, name, x):
    """The plain-C caller (tests/c/spmv_caller.c), compiled against the
    headers, follows spmv_gpu.c:90-136 and checks A*x = [1..m] (sample:
    x = X_mul.txt; sample320: x = X_solution.txt, since A*[1..320] = B)."""
    import subprocess
    from c_caller import build_caller
    exe = build_caller(tmp_path)
    m, n, rp, c, v = hbs.mm_read(golden(name, "A.mtx"))
    if x is None:
        # sample320's B is A*[1..320], so feed A x = [1..m] by solving for x:
        # take x = A^-1 [1..m] from a dense solve, written with full digits
        A = np.zeros((m, n))
        for i in range(m):
            for j in range(rp[i], rp[i + 1]):
                A[i, c[j]] += v[j]
        xv = np.linalg.solve(A, np.arange(1, m + 1, dtype=float))
        xp = tmp_path / "x.txt"
        xp.write_text("".join(f"{float(t)!r}\n" for t in xv))
        xpath = str(xp)
    else:
        xpath = golden(name, x)
    p = subprocess.run([exe, golden(name, "A.mtx"), xpath, "10"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(p.stdout, flush=True)
    assert p.returncode == 0
    kv = dict(l.split(" : ", 1) for l in p.stdout.splitlines() if " : " in l)
    assert int(kv["n_iters"].strip()) == 10
    # the SpMV is the oracle's CSR_spmxv_seq bit for bit, so the printed
    # error is the oracle's own (printed with 7 significant digits)
    xin = hbs.vector_read(xpath, n)
    err_orc = float(np.max(np.abs(orc.spmv(orc.CSR(rp, c, v), xin) - np.arange(1, m + 1))))
    assert float(kv["max_abs_err_vs_1ton"]) == float(f"{err_orc:.6e}")
    assert err_orc < (1e-13 if x else 1e-9)
    assert float(kv["async_err"]) == 0.0
    assert float(kv["blas_err"]) <= 1e-12 * m
    assert float(kv["dot_rel_err"]) <= 1e-14
    assert float(kv["dp_add"]) == 0.75
