"""The tensor viscous operator on the GPU (ksgf_tensorop_*: amrex::MLTensorOp,
the default ICNS diffusion of icns_diffusion.H:18-50) against the oracle's
restatement (tests/test_tensor_oracle.py pins it by the exact Fourier symbol,
self-adjointness, a dense solve and manufactured solutions).

* apply: bit for bit on ragged boxes, for periodic, ABL (no-slip bottom,
  slip top) and inflow / outflow BCs, variable eta and a, both scalar pairs
  the reference uses ((0, -1) for divtau, (1, dt) for the solve).
* solve: the converged answer equals the oracle's BiCGStab solve to 1e-9 of
  its scale, and the GPU's own apply of it returns the rhs.
* The GPU solve converges at second order on a manufactured solution.
* average_velocity_eta_to_faces: bit for bit.
"""
import importlib

import numpy as np
import pytest
import sympy as sp

from helpers import to_np

import test_tensor_oracle as TT

pytestmark = pytest.mark.gpu

lay = importlib.import_module("kynema-sgf_amd.layout")
P_, N_, D_ = 0, 1, 2
CELL, XFACE = 0, 1

PER = [[P_, P_, P_]] * 3
CASES = [
    ("periodic", (16, 12, 20), (1.0, 0.75, 1.25), (1, 1, 1), PER, PER),
    ("abl", (24, 16, 16), (480.0, 320.0, 64.0), (1, 1, 0), TT.ABL_LO, TT.ABL_HI),
    ("inflow_outflow", (20, 12, 16), (1.0, 0.6, 0.8), (0, 1, 0), TT.INOUT_LO, TT.INOUT_HI),
]


def _say(msg):
    print(msg, flush=True)


def _setup(ksgf, n, prob_hi, periodic, seed):
    rng = np.random.default_rng(seed)
    geom = lay.Geometry(n, prob_hi=prob_hi, periodic=periodic)
    ba = lay.BoxArray(geom, lay.chop_random(n, [4, 8], rng))
    dba = ksgf.DeviceBoxArray(ba)
    eta_cc = 1.0 + 0.5 * rng.uniform(size=(n[2], n[1], n[0]))
    eta3 = TT.oracle.velocity_eta_to_faces(eta_cc, periodic)
    a = 1.0 + 0.5 * rng.uniform(size=(n[2], n[1], n[0]))
    u = rng.standard_normal((3, n[2], n[1], n[0]))
    return rng, geom, ba, dba, eta3, a, u


def _upload_faces(ba, dba, eta3):
    out = []
    for d in range(3):
        out.append(dba.upload(ba.scatter(eta3[d][None], XFACE + d, 0)))
    return out


@pytest.mark.parametrize("name,n,prob_hi,periodic,lo,hi", CASES)
def test_apply_bitexact(ksgf, oracle, gpu, name, n, prob_hi, periodic, lo, hi):
    _say(f"tensorop apply {name}")
    rng, geom, ba, dba, eta3, a, u = _setup(ksgf, n, prob_hi, periodic, 3)
    op = ksgf.TensorOp(dba, lo, hi)
    op.set_shear_viscosity(_upload_faces(ba, dba, eta3), 0)
    op.set_acoeffs(dba.upload(ba.scatter(a[None], CELL, 1)), 1)
    uin = dba.upload(ba.scatter(u, CELL, 2))
    for alpha, beta in ((0.0, -1.0), (1.0, 0.37)):
        op.set_scalars(alpha, beta)
        out = dba.alloc(CELL, 3, 0)
        op.apply(out, 0, uin, 2)
        got = ba.gather(to_np(out), CELL, 3, 0)
        want = oracle.tensor_apply(n, geom.dx, lo, hi, alpha, a.ravel(), beta, eta3, u.reshape(3, -1))
        assert np.array_equal(got.reshape(3, -1), want), (alpha, beta)
    op.close()


@pytest.mark.parametrize("name,n,prob_hi,periodic,lo,hi", CASES[1:])
def test_solve_matches_oracle(ksgf, oracle, gpu, name, n, prob_hi, periodic, lo, hi):
    _say(f"tensorop solve {name}")
    rng, geom, ba, dba, eta3, a, u = _setup(ksgf, n, prob_hi, periodic, 5)
    dt = 0.5 * min(geom.dx) ** 2  # dt eta / dx^2 ~ 1: the cross terms matter
    op = ksgf.TensorOp(dba, lo, hi)
    op.set_shear_viscosity(_upload_faces(ba, dba, eta3), 0)
    op.set_acoeffs(dba.upload(ba.scatter(a[None], CELL, 0)), 0)
    op.set_scalars(1.0, dt)
    rhs = u * a[None]
    sol = dba.alloc(CELL, 3, 1)
    info = op.solve(sol, 1, dba.upload(ba.scatter(rhs, CELL, 0)), 0, 1e-11, 0.0)
    _say(f"  GPU iterations {info['iters']} levels {info['nlevels']} r0 {info['r0']:.3e} rfinal {info['rfinal']:.3e}")
    assert info["converged"] == 1 and info["rfinal"] <= 1e-11 * info["r0"]
    got = ba.gather(to_np(sol), CELL, 3, 1)
    want, oinfo = oracle.tensor_solve(n, geom.dx, lo, hi, 1.0, a.ravel(), dt, eta3, rhs.reshape(3, -1), 1e-12, 0.0)
    assert oinfo["status"] == 0
    scale = np.abs(want).max()
    assert np.abs(got.reshape(3, -1) - want).max() <= 1e-9 * scale
    # the converged answer satisfies the oracle's operator
    back = oracle.tensor_apply(n, geom.dx, lo, hi, 1.0, a.ravel(), dt, eta3, got.reshape(3, -1))
    assert np.abs(back - rhs.reshape(3, -1)).max() <= 2e-11 * np.abs(rhs).max()
    op.close()


def test_solve_converges_second_order(ksgf, gpu):
    """the ABL walls (no-slip bottom, slip top) with a manufactured solution,
    variable eta (zero normal slope at the walls) and a, at 16^3, 32^3, 64^3"""
    _say("tensorop MMS")
    errs = _mms_errors(ksgf, (16, 32, 64), ksgf.default_options(do_semicoarsening=1))
    orders = [(np.log2(errs[i - 1][0] / errs[i][0]), np.log2(errs[i - 1][1] / errs[i][1])) for i in (1, 2)]
    _say(f"  errors {errs} orders {orders}")
    assert orders[-1][0] > 1.8 and orders[-1][1] > 1.7, orders


def _mms_errors(ksgf, sizes, options=None):
    X, Y, Z = sp.symbols("x y z")
    Lx, Ly, H = 2.0, 1.5, 1.0
    f = sp.sin(sp.pi * Z / (2 * H))
    g = sp.sin(sp.pi * Z / H)
    U = [sp.cos(2 * sp.pi * X / Lx) * sp.sin(2 * sp.pi * Y / Ly) * f + 0.5 * f,
         sp.sin(2 * sp.pi * X / Lx) * sp.cos(4 * sp.pi * Y / Ly) * f,
         sp.sin(2 * sp.pi * X / Lx) * sp.sin(2 * sp.pi * Y / Ly) * g]
    ETA = 1 + 0.3 * sp.cos(2 * sp.pi * X / Lx) * sp.cos(2 * sp.pi * Y / Ly) + 0.2 * sp.cos(sp.pi * Z / H)
    A_ = 1.0 + 0.5 * sp.cos(2 * sp.pi * (X / Lx + Y / Ly))
    alpha, beta = 1.0, 0.05
    grad = [[sp.diff(U[i], v) for v in (X, Y, Z)] for i in range(3)]
    div = sum(grad[i][i] for i in range(3))
    tau = [[ETA * (grad[i][j] + grad[j][i]) - (sp.Rational(2, 3) * ETA * div if i == j else 0) for j in range(3)]
           for i in range(3)]
    divtau = [sum(sp.diff(tau[j][i], v) for j, v in enumerate((X, Y, Z))) for i in range(3)]
    rhs_e = [sp.lambdify((X, Y, Z), alpha * A_ * U[i] - beta * divtau[i], "numpy") for i in range(3)]
    ue = [sp.lambdify((X, Y, Z), U[i], "numpy") for i in range(3)]
    eta_f = sp.lambdify((X, Y, Z), ETA, "numpy")
    a_f = sp.lambdify((X, Y, Z), A_, "numpy")
    errs = []
    for N in sizes:
        n = (N, N, N)
        geom = lay.Geometry(n, prob_hi=(Lx, Ly, H), periodic=(1, 1, 0))
        ba = lay.BoxArray(geom, lay.chop_random(n, [8, 16], np.random.default_rng(N)))
        dba = ksgf.DeviceBoxArray(ba)
        h = geom.dx
        zz, yy, xx = np.meshgrid(*[(np.arange(n[d]) + 0.5) * h[d] for d in (2, 1, 0)], indexing="ij")
        eta_cc = eta_f(xx, yy, zz) + 0 * xx
        # the viscosity's face values by the GPU's own average_velocity_eta_to_faces
        eta_dev = dba.upload(ba.scatter(eta_cc[None], CELL, 1))
        dba.fill_boundary(eta_dev, CELL, 1, 1)
        faces = dba.average_velocity_eta_to_faces(eta_dev, 1)
        op = ksgf.TensorOp(dba, TT.ABL_LO, TT.ABL_HI, options)
        op.set_shear_viscosity(faces, 0)
        op.set_acoeffs(dba.upload(ba.scatter((a_f(xx, yy, zz) + 0 * xx)[None], CELL, 0)), 0)
        op.set_scalars(alpha, beta)
        rhs = np.stack([rhs_e[i](xx, yy, zz) + 0 * xx for i in range(3)])
        sol = dba.alloc(CELL, 3, 0)
        info = op.solve(sol, 0, dba.upload(ba.scatter(rhs, CELL, 0)), 0, 1e-12, 0.0)
        assert info["converged"] == 1
        _say(f"  {N}^3: {info['iters']} iterations")
        got = ba.gather(to_np(sol), CELL, 3, 0)
        ex = np.stack([ue[i](xx, yy, zz) + 0 * xx for i in range(3)])
        errs.append((np.sqrt(np.mean((got - ex) ** 2)), np.abs(got - ex).max()))
        op.close()
    return errs


def test_eta_to_faces_bitexact(ksgf, oracle, gpu):
    n = (12, 10, 8)
    rng = np.random.default_rng(4)
    geom = lay.Geometry(n, prob_hi=(1.2, 1.0, 0.8), periodic=(1, 0, 0))
    ba = lay.BoxArray(geom, lay.chop_random(n, [4, 8], rng))
    dba = ksgf.DeviceBoxArray(ba)
    cc = 1.0 + rng.uniform(size=(n[2], n[1], n[0]))
    dev = dba.upload(ba.scatter(cc[None], CELL, 1))
    dba.fill_boundary(dev, CELL, 1, 1)
    faces = dba.average_velocity_eta_to_faces(dev, 1)
    want = oracle.velocity_eta_to_faces(cc, (1, 0, 0))
    for d in range(3):
        got = ba.gather(to_np(faces[d]), XFACE + d, 1, 0)[0]
        assert np.array_equal(got, want[d]), d


def test_errors(ksgf, gpu):
    n = (8, 8, 8)
    geom = lay.Geometry(n, prob_hi=(1.0, 1.0, 1.0), periodic=(1, 1, 1))
    ba = lay.BoxArray(geom, [[0, 0, 0, 7, 7, 7]])
    dba = ksgf.DeviceBoxArray(ba)
    op = ksgf.TensorOp(dba, PER, PER)
    u = dba.alloc(CELL, 3, 0)
    with pytest.raises(ksgf.KsgfError, match="set_scalars"):
        op.apply(dba.alloc(CELL, 3, 0), 0, u, 0)
    op.set_scalars(0.0, 1.0)
    with pytest.raises(ksgf.KsgfError, match="singular"):
        op.solve(u, 0, dba.alloc(CELL, 3, 0), 0, 1e-10, 0.0)
    op.close()
    with pytest.raises(ksgf.KsgfError, match="periodic"):
        ksgf.TensorOp(dba, TT.ABL_LO, TT.ABL_HI)
