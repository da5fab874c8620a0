"""torch.autograd binding of the FP64 BiCGStab solve through the C-ABI.

    system = SparseSystem(rowptr, col, values0)        # pattern fixed, on cuda:0
    x = system(values, b)                              # values[nnz], b[n]: float64 cuda
    loss(x).backward()                                 # values.grad, b.grad

Forward: x = A(values)^-1 b with hbs_bicgstab_f64 (x0 = 0).  Backward: the
adjoint solve A^T lambda = dL/dx (hbs_adjoint_f64, BiCGStab on the stored
transpose), then dL/db = lambda and dL/dA_ij = -lambda_i x_j on A's pattern
in CSR order.  torch only supplies device memory: the library's context
has its own stream, so forward and backward first wait for torch's current
stream (the inputs, and the output buffers torch allocated and cleared,
are then complete), and every library call returns only once its results
are.  The HIP kernels are the product; with libhbs.so missing the import
fails.
"""
import torch

from . import hbs


def _hip_runtimes():
    paths = set()
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                paths.add(line.split()[-1])
    return paths


def _check_one_runtime():
    # torch brings its own HIP runtime; libhbs.so must bind to the same copy,
    # which holds when torch is loaded first (the loader then matches it by
    # soname).  Two copies would give two device contexts in one process.
    rt = _hip_runtimes()
    if len(rt) > 1:
        raise RuntimeError("two HIP runtimes are loaded (%s): import torch before the first libhbs call"
                           % ", ".join(sorted(rt)))


class SparseSystem:
    """A CSR pattern on one GPU whose values are an autograd input."""

    def __init__(self, rowptr, col, values, device=0, tol=1e-10, maxit=10000):
        import numpy as np
        torch.cuda.init()
        hbs.lib()
        _check_one_runtime()
        self.device = torch.device("cuda", device)
        self.ctx = hbs.Context(device)
        v = values.detach().cpu().numpy() if torch.is_tensor(values) else np.asarray(values)
        self.A = hbs.DeviceCSR(self.ctx, rowptr, col, v, flags=hbs.HBS_CSR_TRANSPOSE)
        self.S = hbs.Solver(self.A)
        self.n, self.nnz = self.A.m, self.A.nnz
        self.opts = hbs.default_opts(tol=tol, maxit=maxit)
        self.last_info = None

    def _check(self, t, size, name):
        if not (t.is_cuda and t.dtype == torch.float64 and t.numel() == size and t.device == self.device):
            raise ValueError(f"{name} must be a float64 tensor of {size} values on {self.device}")
        return t.contiguous()

    def __call__(self, values, b):
        return SparseSolve.apply(values, b, self)

    def wait_for_torch(self):
        torch.cuda.current_stream(self.device).synchronize()

    def close(self):
        self.ctx.close()


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, system):
        values = system._check(values, system.nnz, "values")
        b = system._check(b, system.n, "b")
        x = torch.zeros_like(b)
        system.wait_for_torch()
        system.A.set_values(hbs.DevPtr(values))
        system.last_info = system.S.bicgstab(hbs.DevPtr(b), hbs.DevPtr(x), system.opts)
        ctx.save_for_backward(values, x)
        ctx.system = system
        return x

    @staticmethod
    def backward(ctx, gx):
        values, x = ctx.saved_tensors
        s = ctx.system
        gx = s._check(gx, s.n, "grad of x")
        lam = torch.empty_like(x)
        gv = torch.empty_like(values) if ctx.needs_input_grad[0] else None
        s.wait_for_torch()
        s.A.set_values(hbs.DevPtr(values))     # the forward's values, whatever ran since
        s.S.adjoint(hbs.DevPtr(gx), hbs.DevPtr(x), hbs.DevPtr(lam), hbs.DevPtr(gv) if gv is not None else None,
                    s.opts)
        return gv, (lam if ctx.needs_input_grad[1] else None), None
