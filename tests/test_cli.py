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


def test_patpart_writes_the_partition_file(tmp_path):
    """hbs_patpart takes the reference's patpart arguments (matrix, parts,
    weight file, imbalance %, seed, output, log) and writes one part id per
    row, as the .part files the multi-rank drivers read."""
    import numpy as np
    from conftest import golden
    from hethyppar_amd import hbs
    w = tmp_path / "w.txt"
    w.write_text("0.75 0.25\n")
    out, log = tmp_path / "a.part", tmp_path / "a.log"
    mtx = golden("sample320", "A.mtx")
    p = subprocess.run([os.path.join(BIN, "hbs_patpart"), mtx, "2", str(w), "3", "42", str(out), str(log)],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    pv = np.array([int(t) for t in out.read_text().split()], dtype=np.int32)
    _, _, rp, c, v = hbs.mm_read(mtx)
    assert pv.size == rp.size - 1 and set(pv.tolist()) == {0, 1}
    ref, cut = hbs.partition(rp, c, 2, target=[0.75, 0.25], final_imbal=0.03, seed=42)
    assert np.array_equal(pv, ref)
    assert f"cut : {cut}" in p.stdout and f"cut  : {cut}" in log.read_text()
    w0 = np.diff(rp)[pv == 0].sum()
    assert f"[0] {w0} " in p.stdout


def test_patpart_rejects_short_weight_file(tmp_path):
    w = tmp_path / "w.txt"
    w.write_text("1.0\n")
    p = subprocess.run([os.path.join(BIN, "hbs_patpart"), os.path.join(ROOT, "tests", "golden", "sample", "A.mtx"),
                        "2", str(w), "3", "1", str(tmp_path / "o.part")], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "weights" in p.stderr
