"""Adjoint of the FP64 solve on the MI355X (hbs_adjoint_f64) and its
torch.autograd binding.  No reference counterpart: for L(x) with
x = A^-1 b, lambda = A^-T dL/dx, dL/db = lambda, dL/dA_ij = -lambda_i x_j.
Pinned by the oracle's BiCGStab on A^T (the oracle's CSR->CSC conversion,
hhp_matrix.c:1476-1535, gives A^T as CSR), by central finite differences
(h = 1e-6) of L through a direct sparse solve, and by
torch.autograd.gradcheck on a small system.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import pyoracle as orc
from hethyppar_amd import hbs
from hethyppar_amd.autograd import SparseSystem

pytestmark = pytest.mark.gpu


def _fem(nbr=120, seed=6):
    rp, c, v = hbs.gen_fem(nbr, 12, 1.1, seed)
    return rp, c, v


def _problem(rp, c, v, seed=1):
    n = rp.size - 1
    xt = hbs.gen_uniform(n, -1.0, 1.0, seed)
    b = hbs.gen_uniform(n, -1.0, 1.0, seed + 100)
    return xt, b


def _adjoint_gpu(ctx, rp, c, v, b, xt, tol=1e-12):
    n = rp.size - 1
    A = hbs.DeviceCSR(ctx, rp, c, v, flags=hbs.HBS_CSR_TRANSPOSE)
    S = hbs.Solver(A)
    db, dx = ctx.to_device(b), ctx.empty(n)
    dx.zero()
    S.bicgstab(db, dx, tol=tol, maxit=5000)
    x = dx.download()
    dg = ctx.to_device(x - xt)
    dl, dgv = ctx.empty(n), ctx.empty(int(c.size))
    info = S.adjoint(dg, dx, dl, dgv, tol=tol, maxit=5000)
    return x, dl.download(), dgv.download(), info


def test_adjoint_matches_oracle_on_transpose(ctx):
    rp, c, v = _fem()
    xt, b = _problem(rp, c, v)
    x, lam, gv, info = _adjoint_gpu(ctx, rp, c, v, b, xt, tol=1e-10)
    A = orc.CSR(rp, c, v)
    tp, tc, tv = orc.csr_to_csc(A)         # CSC of A == CSR of A^T
    g = x - xt
    lo, ito, _, _ = orc.bicgstab(orc.CSR(tp, tc, tv), g, tol=1e-10, maxit=5000)
    print(f"adjoint gpu iters {info.iters} relres {info.relres:.3e} | oracle iters {ito}", flush=True)
    assert info.status == 0 and abs(info.iters - ito) <= 2 and info.relres <= 1.01e-10
    assert np.linalg.norm(lam - lo) / np.linalg.norm(lo) <= 1e-6
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert np.array_equal(gv, -(lam[rows] * x[c]))          # one rounded product per entry


def test_adjoint_central_differences(ctx):
    rp, c, v = _fem(60, 9)
    n = rp.size - 1
    xt, b = _problem(rp, c, v, 3)
    x, lam, gv, _ = _adjoint_gpu(ctx, rp, c, v, b, xt)
    rows = np.repeat(np.arange(n), np.diff(rp))

    def loss(vals, rhs):
        M = sp.csr_matrix((vals, c, rp), shape=(n, n))
        y = spla.spsolve(M.tocsc(), rhs)
        return 0.5 * np.dot(y - xt, y - xt)

    h = 1e-6
    rng = np.random.default_rng(6)
    for j in rng.choice(c.size, 32, replace=False):
        vp, vm = v.copy(), v.copy()
        vp[j] += h
        vm[j] -= h
        fd = (loss(vp, b) - loss(vm, b)) / (2 * h)
        assert abs(fd - gv[j]) <= 1e-6 * max(1.0, abs(fd)), (j, rows[j], c[j], fd, gv[j])
    for k in rng.choice(n, 8, replace=False):
        bp, bm = b.copy(), b.copy()
        bp[k] += h
        bm[k] -= h
        fd = (loss(v, bp) - loss(v, bm)) / (2 * h)
        assert abs(fd - lam[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, lam[k])


def test_autograd_function_matches_adjoint(ctx):
    rp, c, v = _fem(40, 4)
    n = rp.size - 1
    xt, b = _problem(rp, c, v, 5)
    x_ref, lam_ref, gv_ref, _ = _adjoint_gpu(ctx, rp, c, v, b, xt)
    system = SparseSystem(rp, c, v, tol=1e-12)
    dev = system.device
    vals = torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device=dev, requires_grad=True)
    xtt = torch.tensor(xt, dtype=torch.float64, device=dev)
    x = system(vals, bt)
    loss = 0.5 * ((x - xtt) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    assert np.allclose(x.detach().cpu().numpy(), x_ref, rtol=0, atol=1e-12)
    assert np.allclose(bt.grad.cpu().numpy(), lam_ref, rtol=1e-9, atol=1e-12)
    assert np.allclose(vals.grad.cpu().numpy(), gv_ref, rtol=1e-9, atol=1e-12)
    system.close()


def test_autograd_gradcheck_small():
    rp, c, v = hbs.gen_convdiff2d(4, 3, 0.4)
    system = SparseSystem(rp, c, v, tol=1e-14, maxit=500)
    dev = system.device
    vals = torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(hbs.gen_uniform(12, -1.0, 1.0, 2), dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, y: system(a, y), (vals, b), eps=1e-6, atol=1e-6, rtol=1e-4)
    system.close()
