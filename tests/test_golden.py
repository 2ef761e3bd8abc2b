"""The oracle against the golden fixtures (tests/golden, written by
tests/golden/make_golden.py): a change to the oracle's arithmetic shows up as
a bit-level difference here."""
import importlib

import numpy as np
import pytest

import golden_cases

CELL, XFACE = 0, 1


@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_oracle_reproduces_golden(oracle, name):
    case, z = golden_cases.load(name)
    ba = case.ba
    G, vel, rho, gp = oracle.make_state(case)
    vel_new, diag = oracle.advect_project_step(case, G, vel, rho, gp)
    assert diag["dt"] == float(z["dt"])
    assert diag["mlmg"]["iters"] == int(z["mlmg_iters"])
    src = np.zeros(ba.mf_size(CELL, 3, 1))
    mac0 = oracle.extrap_vel_to_faces(G, vel, src, case.bcrec, diag["dt"], 1, 0, 1, 0)
    for d in range(3):
        assert np.array_equal(ba.gather(mac0[d], XFACE + d, 1, 1), z[f"umac_pred{d}"])
        assert np.array_equal(ba.gather(diag["mac"][d], XFACE + d, 1, 1), z[f"umac{d}"])
        assert np.array_equal(ba.gather(diag["flux"][d], XFACE + d, 3, 0), z[f"flux{d}"])
    assert np.array_equal(ba.gather(diag["conv"], CELL, 3, 0), z["conv"])
    assert np.array_equal(ba.gather(vel_new, CELL, 3, 3), z["vel_new"])
