"""Order of accuracy of the oracle's full momentum step (advect + MAC
project + predictor + nodal projection) on the reference's convecting
Taylor vortex verification case.

The reference verifies the Godunov step by the L2 error of u and v against
the exact convecting Taylor vortex under refinement (ConvectingTaylorVortex.cpp:216-337,
docs/sphinx/developer/verification.rst:89-107, deck test/test_files/ctv_godunov_plm:
[0, 2]^3 periodic, u0 = v0 = 1, t = 0.2), expecting second order.  Here the
case runs inviscid at CFL 0.4 on 32^2 and 64^2 slabs (4 cells in z, the field
does not depend on z) with the pressure gradient starting exact
(CTV.activate_pressure) and the nodal projection at the end of every step.

The velocity converges at order > 1.8.  The same check run on an oracle
with a planted mistake in the transverse (corner-transport) terms of
ExtrapVelToFaces -- dropped, doubled or of the wrong sign -- or with the
transverse terms of the flux computation dropped, fails: the order falls to
1.0-1.45.  That is what pins the transverse terms, which flows without
transverse gradients (the uniform-flow test) leave unexercised."""
import math

import numpy as np
import pytest

CELL = 0
T_END = 0.2
RTOL_NODAL, ATOL_NODAL = 1e-11, 1e-14


def ctv_errors(oracle, workloads, n, nz=4):
    """L2 errors (ConvectingTaylorVortex.cpp:216-305 norm: volume-weighted,
    divided by the volume) of u and v at t = 0.2 after n/2 steps"""
    case = workloads.convecting_taylor_vortex(n, nz)
    G, vel, rho, gp = oracle.make_state(case)
    nsteps = n // 2
    dt = T_END / nsteps
    clock = {}
    for _ in range(nsteps):
        vel, diag = oracle.advect_project_step(case, G, vel, rho, gp, dt=dt, clock=clock)
        info, _ = oracle.apply_projection(case, G, vel, rho, gp, dt, RTOL_NODAL, ATOL_NODAL)
        assert diag["mlmg"]["converged"] and info["converged"]
    want, _ = case.extra["exact"](T_END)
    got = case.ba.gather(vel, CELL, 3, 3)
    return [math.sqrt(np.mean((got[c] - want[c]) ** 2)) for c in range(2)]


def ctv_order(oracle, workloads):
    e32 = ctv_errors(oracle, workloads, 32)
    e64 = ctv_errors(oracle, workloads, 64)
    return [math.log2(a / b) for a, b in zip(e32, e64)], e64


def test_ctv_second_order(oracle, workloads):
    oracle.set_plant(0)
    order, e64 = ctv_order(oracle, workloads)
    print(f"CTV order u, v: {order}, L2 at 64^2: {e64}")
    assert min(order) > 1.8, order
    # the vortex is symmetric under x <-> y with u <-> v
    assert abs(e64[0] - e64[1]) <= 1e-12 * e64[0]


@pytest.mark.parametrize("plant", [1, 2, 4, 8], ids=["extrap-transverse-dropped", "extrap-transverse-doubled",
                                                      "extrap-transverse-sign", "flux-transverse-dropped"])
def test_ctv_order_catches_planted_mistakes(oracle, workloads, plant):
    oracle.set_plant(plant)
    try:
        order, _ = ctv_order(oracle, workloads)
    finally:
        oracle.set_plant(0)
    print(f"plant {plant}: order {order}")
    assert max(order) < 1.8, order
