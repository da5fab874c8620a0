"""hbs_dot_f64_noalloc: hbs_dot_f64 without any device allocation (its plan,
partials and result live in a file-scope device array of the library).  It
runs hbs_dot_f64's kernels over the same vector-workgroup plan, so the two
agree bit for bit; both are held to the summation bound against numpy."""
import ctypes as C

import numpy as np
import pytest

from hethyppar_amd import hbs

pytestmark = pytest.mark.gpu


def _dot(fn, ctx, n, x, y):
    out = C.c_double()
    hbs.check(fn(ctx.h, n, x.ptr, y.ptr, C.byref(out)))
    return out.value


def _check(ctx, n, seed):
    L = hbs.lib()
    rng = np.random.default_rng(seed)
    xh, yh = rng.standard_normal(n), rng.standard_normal(n)
    dx, dy = ctx.to_device(xh), ctx.to_device(yh)
    a = _dot(L.hbs_dot_f64_noalloc, ctx, n, dx, dy)
    b = _dot(L.hbs_dot_f64, ctx, n, dx, dy)
    print(f"n={n}: noalloc {a!r} dot {b!r}", flush=True)
    assert a == b
    u = np.finfo(np.float64).eps / 2
    assert abs(a - float(np.dot(xh, yh))) <= 2 * n * u * float(np.abs(xh) @ np.abs(yh))
    dx.free()
    dy.free()


def test_noalloc_dot_matches_dot_then_smaller_input(ctx):
    # two calls in a row, the second on a smaller input, then smaller still
    for n, seed in ((3_000_001, 1), (1_234_567, 2), (4097, 3), (1, 4)):
        _check(ctx, n, seed)


def test_noalloc_dot_follows_plan_changes_and_other_streams(ctx):
    """A new vector-workgroup size rewrites the plan the array holds; a second
    context (its own stream) on the same GPU uses the same array after the
    first's call has returned."""
    try:
        hbs.set_vec_blocks(64)
        _check(ctx, 1_000_003, 5)
    finally:
        hbs.set_vec_blocks(512)
    _check(ctx, 1_000_003, 6)
    ctx2 = hbs.Context(0)
    try:
        _check(ctx2, 777_777, 7)
        _check(ctx, 777_777, 8)
    finally:
        ctx2.close()


def test_noalloc_dot_rejects_sizes_beyond_its_maximum(ctx):
    L = hbs.lib()
    d = ctx.empty(16)
    out = C.c_double()
    rc = L.hbs_dot_f64_noalloc(ctx.h, (1 << 26) + 1, d.ptr, d.ptr, C.byref(out))
    assert rc == -1
    assert b"maximum" in L.hbs_last_error()
    assert L.hbs_dot_f64_noalloc(ctx.h, 0, d.ptr, d.ptr, C.byref(out)) == 0 and out.value == 0.0
