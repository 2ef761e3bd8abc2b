"""Order of accuracy of the HIP step (the C++ host layer's GodunovPredictor:
advect + MAC project + predictor + nodal projection) on the reference's
convecting Taylor vortex verification case (ConvectingTaylorVortex.cpp:216-337,
docs/sphinx/developer/verification.rst:89-107; the deck's [0, 2]^3 periodic
cube at n^3, u0 = v0 = 1, pressure gradient starting exact, t = 0.2,
CFL 0.4, inviscid).

The GPU step tracks the oracle's step to 1e-10 (the multigrid solves differ
by reduction order only) and converges at order > 1.8 for u and v over
32 -> 64 -> 128, the second order the reference's verification expects.  The
oracle side of the same check fails with planted transverse-term mistakes
(tests/test_ctv_order.py), so this pins the HIP transverse terms too."""
import math

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu

CELL = 0
T_END = 0.2
RTOL_NODAL, ATOL_NODAL = 1e-11, 1e-14


def _gpu_ctv(hs, workloads, n):
    case = workloads.convecting_taylor_vortex(n, n)
    S = hs.HostStepper(case)
    S.load()
    S.enable_nodal_projection(RTOL_NODAL, ATOL_NODAL)
    nsteps = n // 2
    dt = T_END / nsteps
    for _ in range(nsteps):
        S.step(dt)
        assert S.last_info["converged"] and S.nodal_info()["converged"]
    v = S.velocity()
    S.close()
    want, _ = case.extra["exact"](T_END)
    return case, v, [math.sqrt(np.mean((v[c] - want[c]) ** 2)) for c in range(2)]


def test_ctv_gpu_second_order_and_oracle(ksgf, oracle, workloads, gpu):
    import importlib

    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    errs = []
    for n in (32, 64, 128):
        print(f"CTV n={n}", flush=True)
        case, v, e = _gpu_ctv(hs, workloads, n)
        errs.append(e)
        print(f"  L2 u, v: {e}", flush=True)
        if n == 32:
            G, vel, rho, gp = oracle.make_state(case)
            dt = T_END / (n // 2)
            for _ in range(n // 2):
                vel, _ = oracle.advect_project_step(case, G, vel, rho, gp, dt=dt)
                oracle.apply_projection(case, G, vel, rho, gp, dt, RTOL_NODAL, ATOL_NODAL)
            r = rel_l2(v, case.ba.gather(vel, CELL, 3, 3))
            print(f"  rel-L2 vs oracle: {r:.2e}", flush=True)
            assert r <= 1e-10, r
    for a, b in zip(errs[:-1], errs[1:]):
        order = [math.log2(x / y) for x, y in zip(a, b)]
        print(f"  order {order}", flush=True)
        assert min(order) > 1.8, (errs, order)
