"""Build libhbs.so (HIP kernels for gfx950 + C++ host layer) in-tree.

    python -m hethyppar_amd.build        # or hethyppar_amd.build.build()

Kernels: hipcc --offload-arch=gfx950, no packed-FP32 instructions, no FP
contraction (FP64 element-wise updates keep the reference's two roundings).
Host files: g++ against the HIP runtime headers.  Link: g++ -shared.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhbs.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("HBS_ARCH", "gfx950")

KERNELS = ["hbs_kernels.hip", "hbs_aux.hip", "hbs_block.hip", "hbs_noalloc.hip"]
HOST = ["hbs_api.cpp", "hbs_solver.cpp", "hbs_block.cpp", "hbs_shard.cpp", "hbs_plan.cpp", "hbs_layout.cpp", "hbs_gen.cpp",
        "hhp_hip.cpp", "hbs_part.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-Wall",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
    jobs = []
    objs = []
    for f in KERNELS:
        o = os.path.join(OUT, f + ".o")
        objs.append(o)
        jobs.append([HIPCC] + HIP_FLAGS + inc + ["-c", os.path.join(CSRC, f), "-o", o])
    for f in HOST:
        o = os.path.join(OUT, f + ".o")
        objs.append(o)
        jobs.append(["g++"] + HOST_FLAGS + inc + ["-c", os.path.join(CSRC, f), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs +
         [f"-L{ROCM}/lib", "-lamdhip64", "-fopenmp", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    build_cli()
    return LIB


CLI = ["hbs_bicgstab_gpu", "hbs_spmv_gpu", "hbs_patpart"]
BIN = os.path.join(HERE, "bin")


def build_cli():
    """Command-line programs with the reference's options and output
    (hethyppar_amd/bin/), linked against libhbs.so."""
    os.makedirs(BIN, exist_ok=True)
    out = []
    for name in CLI:
        exe = os.path.join(BIN, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
              os.path.join(HERE, "cli", name + ".cpp"), "-o", exe,
              f"-L{HERE}", "-lhbs", "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM}/lib"])
        out.append(exe)
    return out


def build_oracle():
    """CPU oracle (test infrastructure): oracle/liboracle.so."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
