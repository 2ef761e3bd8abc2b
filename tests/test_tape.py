"""The CPU-taped reference of the step (oracle/tape.py, the checker of the
GPU adjoint) against the oracle's C step: same forward map up to the MLMG
tolerance (the tape projects exactly), and its taped gradient against
central differences of itself on a smooth direction."""
import numpy as np
import pytest
import torch

from helpers import rel_l2

CELL = 0


def _tape():
    import tape

    return tape


def _fields(workloads, n):
    case = workloads.differentiation(n)
    return case, torch.from_numpy(case.vel.copy()), torch.from_numpy(case.extra["scalar"][0].copy())


def test_tape_forward_matches_oracle_step(oracle, workloads):
    T = _tape()
    case, u, s = _fields(workloads, 16)
    dt = case.extra["dt"]
    dx = case.geom.dx
    G, vel, rho, gp = oracle.make_state(case)
    sbc = np.zeros((1, 6), dtype=np.int32)
    sf = case.ba.scatter(case.extra["scalar"], CELL, 3)
    oracle.fill_ghosts(G, sf, CELL, 1, 3, sbc, np.zeros((6, 1)))
    ut, st = u, s
    for step in range(3):
        vel, diag = oracle.advect_project_step(case, G, vel, rho, gp, dt=dt, scal=sf, scal_bc=sbc,
                                               scal_bcval=np.zeros((6, 1)))
        sf = diag["scal"]
        ut, st = T.step(ut, st, dt, dx)
        a = case.ba.gather(vel, CELL, 3, 3)
        b = case.ba.gather(sf, CELL, 1, 3)[0]
        assert rel_l2(ut.numpy(), a) <= 1e-10, f"step {step}: velocity {rel_l2(ut.numpy(), a)}"
        assert rel_l2(st.numpy(), b) <= 1e-10, f"step {step}: scalar {rel_l2(st.numpy(), b)}"


@pytest.mark.parametrize("kind", ["scalar", "velocity"])
def test_tape_gradient_central_differences(workloads, kind):
    """J = ||s_N||: the taped directional derivative against central
    differences of the tape with h = 1e-7 (for the scalar, a direction
    scaled by the blob, so that no limiter switch is crossed in its tails)."""
    T = _tape()
    case, u, s = _fields(workloads, 12)
    dt, dx = case.extra["dt"], case.geom.dx
    J, gu, gs = T.gradient(u, s, 2, dt, dx)
    g = torch.Generator().manual_seed(4)
    if kind == "scalar":
        v = torch.randn(s.shape, generator=g, dtype=torch.float64) * s
        v /= v.norm()
        f = lambda h: float(T.run(u, s + h * v, 2, dt, dx)[1].square().sum().sqrt())  # noqa: E731
        ad = float((gs * v).sum())
    else:
        v = torch.randn(u.shape, generator=g, dtype=torch.float64)
        v /= v.norm()
        f = lambda h: float(T.run(u + h * v, s, 2, dt, dx)[1].square().sum().sqrt())  # noqa: E731
        ad = float((gu * v).sum())
    h = 1e-7
    fd = (f(h) - f(-h)) / (2 * h)
    assert abs(fd - ad) <= 1e-4 * abs(ad) + 1e-9, (fd, ad)
