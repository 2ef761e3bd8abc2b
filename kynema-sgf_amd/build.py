"""Build libksgf.so (HIP kernels + C ABI) for gfx950, in-tree.

Usage: python kynema-sgf_amd/build.py [--jobs N]

Every HIP translation unit is compiled with
  -O3 --offload-arch=gfx950 -ffp-contract=off
  -Xclang -target-feature -Xclang -packed-fp32-ops   (no packed FP32 ops)
and linked with hipcc -shared into kynema-sgf_amd/libksgf.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libksgf.so")
HOSTLIB = os.path.join(HERE, "libksgf_host.so")
STEPEXE = os.path.join(HERE, "ksgf_step")
LEVCHK = os.path.join(HERE, "ksgf_levels_check")
HOSTDIR = os.path.join(HERE, "host")

SOURCES = ["ksgf_layout.hip", "ksgf_godunov.hip", "ksgf_mlmg.hip", "ksgf_ops.hip", "ksgf_adjoint.hip",
           "ksgf_io.hip", "ksgf_nodal.hip", "ksgf_levels.hip", "ksgf_mlmg_vb.hip", "ksgf_tensor.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    "-Wno-unknown-warning-option",
    "-Wno-pass-failed",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ host layer (Hydro API + ApplyPredictor driver) over the C ABI
    hobj = os.path.join(BUILD, "ksgf_host.o")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(HOSTDIR, "ksgf_host.cpp"), "-o", hobj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host layer compile failed:\n{r.stderr}")
    tmp = HOSTLIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", hobj, "-L" + HERE, "-l:libksgf.so", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host layer link failed:\n{r.stderr}")
    os.replace(tmp, HOSTLIB)
    tmp = STEPEXE + ".tmp"
    cmd = [HIPCC, "-O3", "-std=c++17", "-Wall", os.path.join(HOSTDIR, "ksgf_step_main.cpp"), "-L" + HERE,
           "-l:libksgf.so", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ksgf_step build failed:\n{r.stderr}")
    os.replace(tmp, STEPEXE)
    tmp = LEVCHK + ".tmp"
    cmd = [HIPCC, "-O3", "-std=c++17", "-Wall", os.path.join(HOSTDIR, "ksgf_levels_check.cpp"), "-L" + HERE,
           "-l:libksgf.so", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ksgf_levels_check build failed:\n{r.stderr}")
    os.replace(tmp, LEVCHK)
    if verbose:
        print("built", LIB, HOSTLIB, STEPEXE, LEVCHK)
    return LIB


if __name__ == "__main__":
    jobs = 4
    if "--jobs" in sys.argv:
        jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(jobs=jobs, verbose=True)
