"""Pins of the oracle's tensor viscous operator (amrex::MLTensorOp as
ICNSDiffTensorOp uses it, icns_diffusion.H:18-50 and DiffusionOps.H:54-71:
alpha a u - beta div(eta (grad u + grad u^T) - (2/3) eta (div u) I)).

* The exact discrete Fourier symbol on a periodic box, worked out here from
  the face differences (2 i sin(theta/2)/h), face averages (cos(theta/2)) and
  central differences (i sin(theta)/h) of a plane wave: it fixes the 4/3
  normal coefficient, the transpose term's index and the -(2/3) div u term;
  each of the oracle's three planted mistakes fails it.
* With constant eta on a periodic box the operator is self-adjoint (the
  symbol is real symmetric: S_nq = beta eta s_n s_q / 3 off the diagonal).
* The oracle's BiCGStab solve equals a dense solve of the matrix its apply
  assembles (walls, variable eta and a).
* Manufactured solutions with walls (no-slip bottom, slip top; an inflow
  wall and a zero-gradient outflow), variable eta and a: the discrete
  solution converges at second order, the criterion of verification.rst:67-107.
  (eta has zero normal slope at the walls: the reference takes the wall-face
  viscosity from the adjacent cell, a first-order value where it does not.)
* average_velocity_eta_to_faces: cell means on interior and periodic faces,
  the adjacent cell on a non-periodic domain face.
"""
import numpy as np
import pytest
import sympy as sp

import oracle

oracle.build()
P_, N_, D_ = 0, 1, 2


def _symbol(k, h, alpha, a, beta, eta):
    th = [k[d] * h[d] for d in range(3)]
    dl = [2j * np.sin(th[d] / 2) / h[d] for d in range(3)]  # difference across a face
    av = [np.cos(th[d] / 2) for d in range(3)]  # mean of the two cells beside a face
    cd = [1j * np.sin(th[d]) / h[d] for d in range(3)]  # central difference at a cell
    S = np.zeros((3, 3), complex)
    for n in range(3):
        S[n, n] += alpha * a
        for d in range(3):
            S[n, n] -= beta * eta * dl[d] * (4.0 / 3.0 if d == n else 1.0) * dl[d]
        for q in range(3):
            if q != n:
                S[n, q] -= beta * eta * dl[q] * av[q] * cd[n]  # face q: d u_q / d x_n
                S[n, q] -= beta * eta * dl[n] * (-2.0 / 3.0) * av[n] * cd[q]  # face n: -(2/3) d u_q / d x_q
    return S


def _const_faces(n, eta):
    return [np.full((n[2], n[1], n[0] + 1), eta), np.full((n[2], n[1] + 1, n[0]), eta),
            np.full((n[2] + 1, n[1], n[0]), eta)]


def _fourier_error(plant):
    n, h = (8, 6, 10), (0.3, 0.5, 0.2)
    L = [n[d] * h[d] for d in range(3)]
    alpha, a, beta, eta = 0.7, 1.3, 0.9, 1.7
    per = np.zeros((3, 3), int)
    z, y, x = np.meshgrid(*[(np.arange(n[d]) + 0.5) * h[d] for d in (2, 1, 0)], indexing="ij")
    rng = np.random.default_rng(0)
    worst = 0.0
    oracle.set_plant(plant)
    try:
        for m in [(1, 2, 3), (3, 1, 2), (2, 2, 4), (0, 1, 1), (4, 3, 5)]:
            k = [2 * np.pi * m[d] / L[d] for d in range(3)]
            ph = k[0] * x + k[1] * y + k[2] * z
            ah = rng.standard_normal(3) + 1j * rng.standard_normal(3)
            u = np.real(ah[:, None, None, None] * np.exp(1j * ph)[None])
            out = oracle.tensor_apply(n, h, per, per, alpha, np.full(int(np.prod(n)), a), beta, _const_faces(n, eta),
                                      u.reshape(3, -1)).reshape(u.shape)
            ref = np.real((_symbol(k, h, alpha, a, beta, eta) @ ah)[:, None, None, None] * np.exp(1j * ph)[None])
            worst = max(worst, np.abs(out - ref).max() / np.abs(ref).max())
    finally:
        oracle.set_plant(0)
    return worst


def test_fourier_symbol():
    assert _fourier_error(0) <= 1e-13


@pytest.mark.parametrize("plant", [16, 32, 64])
def test_fourier_symbol_fails_planted_mistakes(plant):
    """dropping the transpose term, dropping -(2/3) div u, or taking the
    face derivative from one cell: each is a large error"""
    assert _fourier_error(plant) > 0.05


def _dense(n, h, lo, hi, alpha, a, beta, eta3):
    N = 3 * int(np.prod(n))
    A = np.zeros((N, N))
    for c in range(N):
        e = np.zeros(N)
        e[c] = 1.0
        A[:, c] = oracle.tensor_apply(n, h, lo, hi, alpha, a, beta, eta3, e.reshape(3, -1)).ravel()
    return A


def test_self_adjoint_constant_eta_periodic():
    n, h = (4, 5, 3), (0.4, 0.3, 0.5)
    per = np.zeros((3, 3), int)
    A = _dense(n, h, per, per, 0.0, None, 1.0, _const_faces(n, 1.3))
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
    # the constant fields are its null space
    for q in range(3):
        e = np.zeros((3, int(np.prod(n))))
        e[q] = 1.0
        assert np.abs(A @ e.ravel()).max() <= 1e-13 * np.abs(A).max()


ABL_LO = [[P_, P_, D_]] * 3  # periodic x, y; no-slip bottom
ABL_HI = [[P_, P_, N_], [P_, P_, N_], [P_, P_, D_]]  # slip top: w Dirichlet, u and v Neumann
INOUT_LO = [[D_, P_, N_], [D_, P_, N_], [D_, P_, D_]]  # wall (inflow at rest) at x lo, slip z walls
INOUT_HI = [[N_, P_, N_], [N_, P_, N_], [N_, P_, D_]]  # zero-gradient outflow at x hi


def test_solve_equals_dense_solve():
    rng = np.random.default_rng(3)
    n, h = (4, 3, 4), (0.5, 0.4, 0.3)
    eta3 = [1 + rng.random((n[2], n[1], n[0] + 1)), 1 + rng.random((n[2], n[1] + 1, n[0])),
            1 + rng.random((n[2] + 1, n[1], n[0]))]
    a = 1 + rng.random(int(np.prod(n)))
    for lo, hi, alpha in ((ABL_LO, ABL_HI, 1.0), (INOUT_LO, INOUT_HI, 0.0)):
        A = _dense(n, h, lo, hi, alpha, a, 0.7, eta3)
        rhs = rng.standard_normal((3, int(np.prod(n))))
        x, info = oracle.tensor_solve(n, h, lo, hi, alpha, a, 0.7, eta3, rhs, 1e-13, 0.0)
        assert info["status"] == 0
        ref = np.linalg.solve(A, rhs.ravel())
        assert np.abs(x.ravel() - ref).max() <= 1e-10 * np.abs(ref).max()


def _mms(lo, hi, U, Lx, Ly, H, sizes):
    X, Y, Z = sp.symbols("x y z")
    U = [u(X, Y, Z) for u in U]
    # zero normal slope of eta at every side: the domain-face viscosity is the
    # adjacent cell's (fixup_eta_on_domain_faces), first-order accurate where
    # eta has a normal gradient at a wall
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
    per = tuple(int(lo[0][d] == P_) for d in range(3))
    errs = []
    for N in sizes:
        n, h = (N, N, N), (Lx / N, Ly / N, H / N)
        zz, yy, xx = np.meshgrid(*[(np.arange(n[d]) + 0.5) * h[d] for d in (2, 1, 0)], indexing="ij")
        eta3 = oracle.velocity_eta_to_faces(eta_f(xx, yy, zz) + 0 * xx, per)
        a = (a_f(xx, yy, zz) + 0 * xx).ravel()
        rhs = np.stack([(rhs_e[i](xx, yy, zz) + 0 * xx).ravel() for i in range(3)])
        sol, info = oracle.tensor_solve(n, h, lo, hi, alpha, a, beta, eta3, rhs, 1e-12, 0.0)
        assert info["status"] == 0
        ex = np.stack([(ue[i](xx, yy, zz) + 0 * xx).ravel() for i in range(3)])
        errs.append((np.sqrt(np.mean((sol - ex) ** 2)), np.abs(sol - ex).max()))
    return [(np.log2(errs[i - 1][0] / errs[i][0]), np.log2(errs[i - 1][1] / errs[i][1])) for i in range(1, len(errs))]


def test_mms_second_order_abl_walls():
    """no-slip bottom, slip top (u, v: zero normal derivative; w = 0)"""
    Lx, Ly, H = 2.0, 1.5, 1.0

    def f(z):
        return sp.sin(sp.pi * z / (2 * H))

    def g(z):
        return sp.sin(sp.pi * z / H)

    U = [lambda x, y, z: sp.cos(2 * sp.pi * x / Lx) * sp.sin(2 * sp.pi * y / Ly) * f(z) + 0.5 * f(z),
         lambda x, y, z: sp.sin(2 * sp.pi * x / Lx) * sp.cos(4 * sp.pi * y / Ly) * f(z),
         lambda x, y, z: sp.sin(2 * sp.pi * x / Lx) * sp.sin(2 * sp.pi * y / Ly) * g(z)]
    orders = _mms(ABL_LO, ABL_HI, U, Lx, Ly, H, (8, 16, 32))
    assert orders[-1][0] > 1.8 and orders[-1][1] > 1.7, orders


def test_mms_second_order_inflow_outflow():
    """x: a wall at rest (all Dirichlet) at the low side, zero gradient at
    the high side; y periodic; z slip walls on both sides"""
    Lx, Ly, H = 1.0, 1.5, 1.0

    def p(x):  # 0 at x = 0, zero slope at x = Lx
        return sp.sin(sp.pi * x / (2 * Lx))

    U = [lambda x, y, z: p(x) * sp.cos(sp.pi * z / H) * (1 + 0.5 * sp.sin(2 * sp.pi * y / Ly)),
         lambda x, y, z: p(x) * sp.cos(2 * sp.pi * y / Ly) * sp.cos(sp.pi * z / H),
         lambda x, y, z: p(x) * sp.sin(sp.pi * z / H) * sp.sin(2 * sp.pi * y / Ly)]
    orders = _mms(INOUT_LO, INOUT_HI, U, Lx, Ly, H, (8, 16, 32))
    assert orders[-1][0] > 1.8 and orders[-1][1] > 1.5, orders


def test_velocity_eta_to_faces():
    rng = np.random.default_rng(1)
    cc = rng.random((3, 4, 5))
    fx, fy, fz = oracle.velocity_eta_to_faces(cc, (1, 0, 0))
    assert fx.shape == (3, 4, 6) and fy.shape == (3, 5, 5) and fz.shape == (4, 4, 5)
    assert np.array_equal(fx[:, :, 1:-1], 0.5 * (cc[:, :, :-1] + cc[:, :, 1:]))
    assert np.array_equal(fx[:, :, 0], 0.5 * (cc[:, :, -1] + cc[:, :, 0])) and np.array_equal(fx[:, :, 0], fx[:, :, -1])
    assert np.array_equal(fy[:, 0, :], cc[:, 0, :]) and np.array_equal(fy[:, -1, :], cc[:, -1, :])
    assert np.array_equal(fz[0], cc[0]) and np.array_equal(fz[-1], cc[-1])
    const = oracle.velocity_eta_to_faces(np.full((3, 4, 5), 2.5), (0, 0, 0))
    assert all(np.all(f == 2.5) for f in const)
