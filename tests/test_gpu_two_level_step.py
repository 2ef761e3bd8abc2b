"""Two-level advect + MAC-project steps on the GPU (levels.TwoLevelStep:
per-level Godunov, composite MAC projection, face and cell fill-patch, fine
fluxes averaged onto the coarse faces, average_down) against the oracle's
two_level_step: Taylor-Green on a 16^3 level 0 (periodic, or with no-slip
walls in y that level-1 boxes touch) with three level-1 boxes at ratio 2,
two steps.  Everything but the composite solve is
bit-exact; the solve agrees to its tolerance, so the velocities agree to
1e-9 of their scale.  The GPU step conserves the level-0 momentum too."""
import importlib

import numpy as np
import pytest

from helpers import to_np

pytestmark = pytest.mark.gpu

CELL = 0


@pytest.mark.parametrize("walls", [False, True])
def test_two_level_steps_match_oracle(ksgf, oracle, gpu, walls):
    import test_multilevel_oracle as T

    print(f"two-level steps, tgv16 + 3 fine boxes, y walls {walls}", flush=True)
    lv = importlib.import_module("kynema-sgf_amd.levels")
    case, fba, vel_c, vel_f, gp_c, gp_f = T.tgv_two_level_walls() if walls else T.tgv_two_level()
    S = lv.TwoLevelStep(case, fba)
    S.load(vel_c, vel_f, gp_c, gp_f)
    ref_c, ref_f = vel_c.copy(), vel_f.copy()
    tot0 = case.ba.gather(vel_c, CELL, 3, 3).sum(axis=(1, 2, 3))
    clock = {}
    for s in range(2):
        rc, rf, rinfo = oracle.two_level_step(case, fba, ref_c, ref_f, gp_c, gp_f, clock=clock)
        ref_c, ref_f = rc, rf
        dt = S.step()
        print(f"  step {s}: dt {dt:.6e} (oracle {rinfo['dt']:.6e}), composite iterations {S.last_info['iters']} "
              f"(oracle {rinfo['iters']})", flush=True)
        assert abs(dt - rinfo["dt"]) <= 1e-13 * rinfo["dt"]
        assert abs(S.last_info["iters"] - rinfo["iters"]) <= 1
    gc, gf = S.velocity()
    gc, gf = to_np(gc), to_np(gf)
    scale = np.abs(ref_c).max()
    vc = case.ba.gather(gc, CELL, 3, 3)
    assert np.abs(vc - case.ba.gather(ref_c, CELL, 3, 3)).max() <= 1e-9 * scale
    fm = np.zeros(tuple(fba.geom.n_cell[::-1]), dtype=bool)
    for b in fba.boxes:
        fm[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1] = True
    vf = fba.gather(gf, CELL, 3, 3)
    assert np.abs(vf[:, fm] - fba.gather(ref_f, CELL, 3, 3)[:, fm]).max() <= 1e-9 * scale
    tot2 = vc.sum(axis=(1, 2, 3))
    assert np.abs(tot2 - tot0).max() <= 1e-13 * np.abs(vc).sum()
    S.close()


@pytest.mark.parametrize("walls", [False, True])
def test_cpp_two_level_driver_equals_python(ksgf, oracle, gpu, walls):
    """ksgf::amr::TwoLevelPredictor (the C++ host layer, through
    libksgf_host.so) makes the same calls as levels.TwoLevelStep in the same
    order: two steps give the same dt, iteration counts and velocities on
    both levels (ghosts included) bit for bit"""
    import test_multilevel_oracle as T

    print(f"C++ two-level driver, y walls {walls}", flush=True)
    lv = importlib.import_module("kynema-sgf_amd.levels")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    case, fba, vel_c, vel_f, gp_c, gp_f = T.tgv_two_level_walls() if walls else T.tgv_two_level()
    P = lv.TwoLevelStep(case, fba)
    P.load(vel_c, vel_f, gp_c, gp_f)
    H = hs.HostTwoLevel(case, fba)
    H.load(vel_c, vel_f, gp_c, gp_f)
    for s in range(2):
        dp = P.step()
        dh = H.step()
        print(f"  step {s}: dt {dh:.6e}, composite iterations {H.last_info['iters']}", flush=True)
        assert dh == dp and H.last_info["iters"] == P.last_info["iters"]
        pc, pf = P.velocity()
        hc, hf = H.velocity()
        assert np.array_equal(hc, to_np(pc)) and np.array_equal(hf, to_np(pf)), s
    H.close()
    P.close()
