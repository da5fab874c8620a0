"""Build the plain-C caller of the reference-shaped device API
(tests/c/spmv_caller.c) against include/ and libhbs.so."""
import os
import subprocess

from conftest import ROOT


def build_caller(out_dir):
    exe = os.path.join(str(out_dir), "spmv_caller")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    lib_dir = os.path.join(ROOT, "hethyppar_amd")
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{rocm}/include",
           os.path.join(ROOT, "tests", "c", "spmv_caller.c"), "-o", exe,
           f"-L{lib_dir}", "-lhbs", "-lm", f"-Wl,-rpath,{lib_dir}", f"-Wl,-rpath,{rocm}/lib"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout
    return exe
