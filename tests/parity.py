"""The single-system BiCGStab bar, shared by the GPU tests.

Against the oracle's restatement of the reference CPU solver
(bicgstab_cpu.c:172-255 with the stop test of bicgstab_gpu.c:148-150,
231-236, 289-293): iterations within +-2, relative residual
sqrt(||b-Ax||^2/||b||^2) <= 1e-8 (1% slack for the stopping iterate),
solution within 1e-6 relative of the oracle's.

Order-sensitive systems.  Every parallel reduction re-associates the dot
sums, and for nearly symmetric systems that alone moves the oracle's own
stopping iteration (its blocked-dot runs).  For such a system the bar is the
iteration range those re-associated oracle runs span, +-2, and a forward
error (against the known solution) within 4x the worst of them.  check()
returns whether that bar was used, so a test can cap how many systems or
columns may need it.
"""
import numpy as np

import pyoracle as orc


def check(x, info, xo, ito, what, A=None, b=None, xt=None, maxit=2000, tol=1e-8):
    """Assert the bar; returns True when the order-sensitive bar was applied."""
    print(f"{what}: gpu iters {info.iters} stop {info.stop} relres {info.relres:.3e} | oracle iters {ito}",
          flush=True)
    assert info.status == 0
    assert info.relres <= 1.01 * tol
    dist = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    if abs(info.iters - ito) <= 2 and dist <= 1e-6:
        return False
    assert A is not None, f"{what}: iterations {info.iters} vs {ito}, distance {dist:.2e}"
    _, lo, hi, sols = orc.iteration_spread(A, b, tol=tol, maxit=maxit)
    fe = lambda y: np.linalg.norm(y - xt) / np.linalg.norm(xt)
    worst = max(fe(y) for y in sols)
    print(f"{what}: ORDER-SENSITIVE BAR: re-associated oracle runs span {lo}..{hi} iterations, "
          f"forward errors up to {worst:.2e}; gpu {fe(x):.2e} ({dist:.2e} from the oracle)", flush=True)
    assert lo - 2 <= info.iters <= hi + 2
    assert fe(x) <= 4 * worst
    return True
