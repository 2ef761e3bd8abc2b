"""The C++ host layer's two-level Hydro::MacProjector (MacProjector over
Geom(0..1), initProjector / setDomainBC / setUMAC / project / getPhiLevels),
run by the ksgf_levels_check program: after the projection the composite
divergence is zero to the solver tolerance."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kynema-sgf_amd", "ksgf_levels_check")


def test_host_two_level_macprojector(gpu):
    print("host-layer two-level MacProjector", flush=True)
    assert os.path.exists(EXE), "ksgf_levels_check is not built"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "levels check ok" in r.stdout
