"""The ragged box layout of the batching workload: 1,500 boxes with edges
drawn from {8,16,24,32,48,64} per axis (seed 4) over a 320x320x480
periodic domain, Taylor-Green plus noise.  Every kernel of the step works
per cell or per face from values that do not depend on how the domain is
cut into boxes (ghosts are copies of neighbour cells, the multigrid runs
domain-wide), so the step on 1,500 ragged boxes must equal the step on a
uniform 64^3 decomposition bit for bit; a smaller ragged layout is checked
against the oracle."""
import importlib

import numpy as np
import pytest

from helpers import rel_l2, to_np

pytestmark = pytest.mark.gpu
CELL = 0


def test_ragged_layout_tracks_oracle(ksgf, oracle, workloads, gpu):
    st_mod = importlib.import_module("kynema-sgf_amd.stepper")
    case = workloads.ragged((64, 64, 80), nbox=None, edges=(8, 16, 24, 32))
    assert case.ba.nbox >= 24, case.ba.nbox
    G, vel, rho, gp = oracle.make_state(case)
    S = st_mod.AdvectProjectStep(case)
    S.load()
    clock = {}
    for n in range(2):
        vel, diag = oracle.advect_project_step(case, G, vel, rho, gp, clock=clock)
        S.step()
        assert S.last_info["iters"] == diag["mlmg"]["iters"], f"step {n}"
        a = case.ba.gather(to_np(S.vel), CELL, 3, 3)
        b = case.ba.gather(vel, CELL, 3, 3)
        assert rel_l2(a, b) <= 1e-12, f"step {n}: {rel_l2(a, b)}"


def test_ragged_1500_boxes_equal_uniform_layout(ksgf, workloads, gpu):
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    lay = importlib.import_module("kynema-sgf_amd.layout")
    rag = workloads.ragged()
    assert rag.ba.nbox == 1500
    e = rag.ba.boxes[:, 3:] - rag.ba.boxes[:, :3] + 1
    assert set(np.unique(e)) == {8, 16, 24, 32, 48, 64}
    out = []
    for ba in (rag.ba, lay.BoxArray(rag.geom, lay.chop_uniform(rag.geom.n_cell, 64))):
        case = rag
        case.ba = ba
        print(f"{ba.nbox} boxes over {rag.geom.n_cell}", flush=True)
        S = hs.HostStepper(case)
        S.load()
        dts = [S.step(check_divergence=True) for _ in range(2)]
        assert S.last_divmax <= 1e-10, S.last_divmax
        out.append((dts, S.last_info, S.velocity()))
        S.close()
    (dt_a, info_a, v_a), (dt_b, info_b, v_b) = out
    assert dt_a == dt_b
    assert info_a["iters"] == info_b["iters"] and info_a["rfinal"] == info_b["rfinal"]
    assert np.array_equal(v_a, v_b)
