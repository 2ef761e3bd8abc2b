"""CPU check of the packed red-black multigrid kernels: the per-cell bodies
the HIP kernels run (kynema-sgf_amd/csrc/ksgf_mg_grid.h) are compiled as host
C++ with AddressSanitizer and compared bit for bit with natural-layout
half sweeps and restriction, over periodic / Neumann / Dirichlet sides,
semicoarsened ratios and every sweep mode (tests/mg_sweep_emulation.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_packed_mg_kernels_match_natural_layout(tmp_path):
    exe = tmp_path / "mg_emu"
    src = os.path.join(HERE, "mg_sweep_emulation.cpp")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", src, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().startswith("ok"), r.stdout + r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_ppm_trace_pair_matches_single_traces(tmp_path):
    """The Godunov stage-A kernel's shared-slope trace pair against two
    single-cell traces (and the dual-number evaluation's values), bit for
    bit, on the host build of ksgf_ppm.h (tests/ppm_pair_check.cpp)."""
    exe = tmp_path / "ppm_pair"
    src = os.path.join(HERE, "ppm_pair_check.cpp")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", src, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().startswith("ok"), r.stdout + r.stderr
