"""Test inputs: the reference's fixtures and small synthetic matrices."""
import numpy as np

from conftest import golden, read_ints
from hethyppar_amd import hbs


def sample():
    m, n, rp, c, v = hbs.mm_read(golden("sample", "A.mtx"))
    return rp, c, v


def sample320():
    m, n, rp, c, v = hbs.mm_read(golden("sample320", "A.mtx"))
    return rp, c, v


def convdiff64():
    """64x64 5-point convection-diffusion (conv 0.5), x_target ~ U[-1,1] seed 1,
    b = A x_target, x0 = 0."""
    rp, c, v = hbs.gen_convdiff2d(64, 64, 0.5)
    xt = hbs.gen_uniform(rp.size - 1, -1.0, 1.0, 1)
    b = hbs.host_spmv(rp, c, v, xt)
    return rp, c, v, xt, b


def markov(n, seed=2):
    rp, c, v = hbs.gen_markov(n, 12, 22, max(1, int(round(0.02 * n))), 1.05, seed)
    return rp, c, v


def random_csr(m, n, density, seed, empty_rows=True):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(m):
        k = rng.binomial(n, density)
        if not empty_rows:
            k = max(k, 1)
        cols = np.sort(rng.choice(n, size=k, replace=False)) if k else np.zeros(0, np.int64)
        rows.append(cols)
    rp = np.zeros(m + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate(rows).astype(np.int32) if rp[-1] else np.zeros(0, np.int32)
    val = rng.standard_normal(col.size)
    return rp, col, val


def dense(rp, c, v, n=None):
    m = rp.size - 1
    n = m if n is None else n
    A = np.zeros((m, n))
    for i in range(m):
        for j in range(rp[i], rp[i + 1]):
            A[i, c[j]] += v[j]
    return A


def seq_bound(rp, c, v, x):
    """Per-row bound of any summation order vs another: 2*gamma_k*sum|a||x|."""
    m = rp.size - 1
    u = np.finfo(np.float64).eps / 2
    k = np.diff(rp).astype(np.float64)
    g = 2 * k * u / (1 - k * u)
    rows = np.repeat(np.arange(m), np.diff(rp))
    s = np.bincount(rows, weights=np.abs(v * x[c]), minlength=m)
    return 2 * g * s


PARTITIONS = [
    ("sample", "part2.txt", 2),
    ("sample", "part2-simple.txt", 2),
    ("sample", "part2-imbal.txt", 2),
    ("sample320", "row_part_2", 2),
    ("sample320", "row_part_16", 16),
]


def partition(name, fname):
    return read_ints(golden(name, fname))
