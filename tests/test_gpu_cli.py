"""The command-line programs on the reference's own fixtures, with the
reference's stdout format (key : value lines, src/entry/bicgstab_gpu.c:
579-606, spmv_gpu.c:139-152) and output vector format ("%lf" per line)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import golden
from hethyppar_amd import hbs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hethyppar_amd", "bin")


def _metrics(text):
    out = {}
    for line in text.splitlines():
        m = re.match(r"^\s*([A-Za-z_][A-Za-z0-9_]*)\s*:\s*(\S+)", line)
        if m:
            out[m.group(1)] = m.group(2)
    return out


def test_bicgstab_cli_sample320_fixed_iterations(tmp_path):
    out = tmp_path / "X.txt"
    p = subprocess.run([os.path.join(BIN, "hbs_bicgstab_gpu"), "-m", golden("sample320", "A.mtx"),
                        "-x", golden("sample320", "X_init.txt"), "-y", golden("sample320", "B.txt"),
                        "-o", str(out), "-n", "20"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    m = _metrics(p.stdout)
    for key in ("n_iters", "spmv", "file_read", "relative_residual", "everything_total"):
        assert key in m
    assert int(m["n_iters"]) == 20
    x = hbs.vector_read(str(out), 320)
    gold = hbs.vector_read(golden("sample320", "out_X_seq.txt"), 320)
    assert np.max(np.abs(x - gold)) <= 1e-6 + 1e-9


def test_bicgstab_cli_tolerance_and_profile(tmp_path):
    out = tmp_path / "X.txt"
    env = dict(os.environ, HHP_PROFILE="1")
    p = subprocess.run([os.path.join(BIN, "hbs_bicgstab_gpu"), "-m", golden("sample320", "A.mtx"),
                        "-x", golden("sample320", "X_init.txt"), "-y", golden("sample320", "B.txt"),
                        "-o", str(out), "-n", "1000", "-t", "1e-8"], capture_output=True, text=True,
                       timeout=120, env=env)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    m = _metrics(p.stdout)
    assert int(m["n_iters"]) < 1000 and float(m["relative_residual"]) <= 1.01e-8
    assert m["PROFILE_enabled"] == "1" and "PROFILE_spmv_pct" in m


def test_spmv_cli_sample_one_to_n(tmp_path):
    out = tmp_path / "Y.txt"
    p = subprocess.run([os.path.join(BIN, "hbs_spmv_gpu"), "-m", golden("sample", "A.mtx"),
                        "-x", golden("sample", "X_mul.txt"), "-o", str(out), "-n", "5"],
                       capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    m = _metrics(p.stdout)
    assert int(m["n_iters"]) == 5 and float(m["max_abs_err_vs_1ton"]) < 1e-12
    y = hbs.vector_read(str(out), 16)
    assert np.max(np.abs(y - np.arange(1, 17))) < 1e-6
