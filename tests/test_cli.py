"""The command-line programs (hethyppar_amd/bin) keep the reference mains'
argument checks (src/entry/bicgstab_gpu.c:350-357, spmv_gpu.c:66-70); these
checks run before any device call, so they are tested on the CPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hethyppar_amd", "bin")


@pytest.mark.parametrize("name,msg", [
    ("hbs_bicgstab_gpu", "Error: All input and output files must be specified."),
    ("hbs_spmv_gpu", "Error: -m, -x, and -o must all be specified."),
])
def test_cli_requires_files(name, msg):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), "build the package first (hethyppar_amd.build)"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and msg in p.stderr


def test_cli_help_lists_reference_options():
    p = subprocess.run([os.path.join(BIN, "hbs_bicgstab_gpu"), "--help"], capture_output=True, text=True,
                       timeout=60)
    for opt in ("--input-matrix", "--input-x", "--input-y", "--output-x", "--n-iters"):
        assert opt in p.stdout


def test_cli_missing_input_file(tmp_path):
    p = subprocess.run([os.path.join(BIN, "hbs_spmv_gpu"), "-m", str(tmp_path / "none.mtx"), "-x",
                        str(tmp_path / "none.txt"), "-o", str(tmp_path / "y.txt")],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "does not exist" in p.stderr
