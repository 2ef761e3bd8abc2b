"""Pins of the oracle's nodal projection (NodalProjector / MLNodeLaplacian with a
constant sigma; src/projection/incflo_apply_nodal_projection.cpp:149-465).

* The reference's own known answer (unit_tests/projection/test_pressure_offset.cpp):
  a 2x2x16 unit box with slip walls and a pressure_outflow top, uniform
  w = w0, dt = 1, rho = 1: the nodal pressure on the bottom plane averages to
  -w0 (9.81 for w0 = -9.81, 0 without it) to eps * 1e8.
* The 27-point stencil, the Neumann mirror rows and the Dirichlet nodes against
  a finite-element matrix assembled here independently, from Kronecker
  products of the 1-D element matrices, and solved densely.
* The nodal right-hand side against the finite-element divergence assembled
  the same way.
* Decomposition independence, and multigrid convergence.
"""
import importlib
import itertools

import numpy as np
import pytest

import oracle

layout = importlib.import_module("kynema-sgf_amd.layout")
oracle.build()

N_, D_, P_ = 1, 2, 0  # LinOp BCs: Neumann, Dirichlet, periodic
EPS = np.finfo(float).eps


def _ghosts_as_apply_projection(lay, G, vel, ng):
    """vel.setBndry(0) then FillBoundary(periodicity): physical-boundary ghosts
    zero, interior and periodic ghosts from the valid cells"""
    ba = G.ba
    for b in range(ba.nbox):
        v = ba.fab_view(vel, b, layout.CELL, 3, ng)
        inner = v[:, ng:-ng, ng:-ng, ng:-ng].copy()
        v[...] = 0.0
        v[:, ng:-ng, ng:-ng, ng:-ng] = inner
    oracle.fill_boundary(G, vel, layout.CELL, 3, ng)


def _fe_matrices(n, h, lobc, hibc, sigma):
    """(L, B, keep): the oracle's nodal system assembled from Q1 element
    integrals.  L = -D K / V on the kept (non-Dirichlet) nodes, with K the
    stiffness matrix and D = 2^(Neumann planes of the node); B maps cell
    velocities (3, ncell) to -D G u / V, the divergence right-hand side."""
    N = [n[d] if lobc[d] == P_ else n[d] + 1 for d in range(3)]
    nn = int(np.prod(N))
    K1 = [np.array([[1.0, -1.0], [-1.0, 1.0]]) / h[d] for d in range(3)]
    M1 = [np.array([[1.0 / 3, 1.0 / 6], [1.0 / 6, 1.0 / 3]]) * h[d] for d in range(3)]
    G1 = [np.array([-1.0, 1.0]) for d in range(3)]  # integral of dN/dx over the cell
    m1 = [np.array([0.5, 0.5]) * h[d] for d in range(3)]  # integral of N over the cell

    def kron3(a, b, c):  # node order (z, y, x) within the element
        return np.kron(c, np.kron(b, a))

    Ke = sigma * (kron3(K1[0], M1[1], M1[2]) + kron3(M1[0], K1[1], M1[2]) + kron3(M1[0], M1[1], K1[2]))
    Ge = [kron3(G1[0], m1[1], m1[2]), kron3(m1[0], G1[1], m1[2]), kron3(m1[0], m1[1], G1[2])]
    K = np.zeros((nn, nn))
    ncell = int(np.prod(n))
    Bg = np.zeros((nn, 3 * ncell))
    for k, j, i in itertools.product(range(n[2]), range(n[1]), range(n[0])):
        ids = []
        for c, b, a in itertools.product((0, 1), (0, 1), (0, 1)):
            x, y, z = i + a, j + b, k + c
            x %= N[0]
            y %= N[1]
            z %= N[2]
            ids.append(x + N[0] * (y + N[1] * z))
        ids = np.array(ids)
        K[np.ix_(ids, ids)] += Ke
        cell = i + n[0] * (j + n[1] * k)
        for d in range(3):
            Bg[ids, d * ncell + cell] += Ge[d]
    V = h[0] * h[1] * h[2]
    Dg = np.ones(nn)
    keep = np.ones(nn, dtype=bool)
    for k, j, i in itertools.product(range(N[2]), range(N[1]), range(N[0])):
        idx = i + N[0] * (j + N[1] * k)
        for d, q in enumerate((i, j, k)):
            for side, bc in ((0, lobc[d]), (N[d] - 1, hibc[d])):
                if bc == P_ or q != side:
                    continue
                if bc == N_:
                    Dg[idx] *= 2.0
                else:
                    keep[idx] = False
    L = -(Dg[:, None] * K) / V
    B = -(Dg[:, None] * Bg) / V
    return L, B, keep, Dg


CASES = [
    ((4, 3, 5), (0.5, 0.7, 0.3), (P_, N_, N_), (P_, N_, D_)),
    ((4, 4, 4), (1.0, 0.5, 0.25), (N_, N_, N_), (N_, N_, N_)),
    ((4, 2, 6), (0.3, 0.3, 0.2), (P_, P_, D_), (P_, P_, D_)),
]


@pytest.mark.parametrize("n,h,lobc,hibc", CASES)
def test_nodal_solve_equals_fe_dense_solve(n, h, lobc, hibc):
    sigma = 0.7
    L, _, keep, Dg = _fe_matrices(n, h, lobc, hibc, sigma)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(L.shape[0])
    b[~keep] = 0.0
    Lk = L[np.ix_(keep, keep)]
    singular = all(lobc[d] != D_ and hibc[d] != D_ for d in range(3))
    if singular:
        # make b solvable: weighted mean zero, weights 1/D
        w = 1.0 / Dg
        b[keep] -= np.sum(w[keep] * b[keep]) / np.sum(w[keep])
        x, *_ = np.linalg.lstsq(Lk, b[keep], rcond=None)
    else:
        x = np.linalg.solve(Lk, b[keep])
    dx = np.array(h) * 1.0
    phi, info = oracle.nodal_solve(n, dx, lobc, hibc, sigma, b, 1e-11, 0.0)
    assert info["converged"], info
    got = phi[keep]
    if singular:
        got = got - got.mean()
        x = x - x.mean()
    assert np.abs(got - x).max() <= 1e-9 * np.abs(x).max()
    assert np.all(phi[~keep] == 0.0)


def test_nodal_rhs_equals_fe_divergence():
    n, h = (4, 3, 5), (0.5, 0.7, 0.3)
    lobc, hibc = (P_, N_, N_), (P_, N_, D_)
    geom = layout.Geometry(n, prob_hi=tuple(n[d] * h[d] for d in range(3)), periodic=(1, 0, 0))
    ba = layout.BoxArray(geom, layout.chop_uniform(n, 2))
    G = oracle.Geom(ba)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((3, n[2], n[1], n[0]))
    ng = 1
    vel = ba.scatter(u, layout.CELL, ng)
    _ghosts_as_apply_projection(layout, G, vel, ng)
    rhs = oracle.nodal_rhs(G, vel, ng, lobc, hibc)
    _, B, keep, _ = _fe_matrices(n, h, lobc, hibc, 1.0)
    ref = B @ u.reshape(3, -1).ravel()
    assert np.abs(rhs[keep] - ref[keep]).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("w0", [0.0, -9.81])
def test_pressure_offset_known_answer(w0):
    """unit_tests/projection/test_pressure_offset.cpp: dynamic_only / full_pressure"""
    n = (2, 2, 16)
    geom = layout.Geometry(n, prob_hi=(1.0, 1.0, 1.0), periodic=(0, 0, 0))
    ba = layout.BoxArray(geom, layout.chop_uniform(n, 16))
    G = oracle.Geom(ba)
    ng = 3
    u = np.zeros((3, n[2], n[1], n[0]))
    u[2] = w0
    vel = ba.scatter(u, layout.CELL, ng)
    _ghosts_as_apply_projection(layout, G, vel, ng)
    lobc, hibc = (N_, N_, N_), (N_, N_, D_)
    rtol, atol = EPS * 1e5, EPS * 1e2  # MLMGOptions defaults (MLMGOptions.H:61-65)
    info, phi, gp = oracle.nodal_project(G, vel, ng, lobc, hibc, 1.0, rtol, atol)
    assert info["status"] == 0
    N = (3, 3, 17)
    p = phi.reshape(N[2], N[1], N[0])
    pbottom = p[0].sum() / ((n[0] + 1) * (n[1] + 1))
    assert abs(pbottom - (-w0)) <= EPS * 1e8
    g = ba.gather(gp, layout.CELL, 3, 0)
    assert np.abs(g[2] - w0).max() <= 1e-9
    v = ba.gather(vel, layout.CELL, 3, ng)
    assert np.abs(v).max() <= 1e-9


def _tgv(n, seed=1):
    x = [(np.arange(n[d]) + 0.5) / n[d] for d in range(3)]
    X, Y, Z = np.meshgrid(x[0], x[1], x[2], indexing="ij")
    tp = 2 * np.pi
    u = np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z)
    v = -np.cos(tp * X) * np.sin(tp * Y) * np.cos(tp * Z)
    w = np.zeros_like(u)
    rng = np.random.default_rng(seed)
    f = np.stack([a.transpose(2, 1, 0) for a in (u, v, w)])
    return f + rng.uniform(-1e-3, 1e-3, f.shape)


def test_nodal_projection_decomposition_independent_and_converges():
    n = (32, 32, 32)
    geom = layout.Geometry(n, periodic=(1, 1, 1))
    u = _tgv(n)
    # add a smooth gradient part the projection must remove:
    # grad of sin(2 pi x) sin(4 pi y) cos(2 pi z)
    x = [(np.arange(n[d]) + 0.5) / n[d] for d in range(3)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    tp = 2 * np.pi
    u = u + np.stack([tp * np.cos(tp * X) * np.sin(2 * tp * Y) * np.cos(tp * Z),
                      2 * tp * np.sin(tp * X) * np.cos(2 * tp * Y) * np.cos(tp * Z),
                      -tp * np.sin(tp * X) * np.sin(2 * tp * Y) * np.sin(tp * Z)])
    outs = []
    for mgs in (32, 16):
        ba = layout.BoxArray(geom, layout.chop_uniform(n, mgs))
        G = oracle.Geom(ba)
        vel = ba.scatter(u, layout.CELL, 1)
        _ghosts_as_apply_projection(layout, G, vel, 1)
        rhs0 = oracle.nodal_rhs(G, vel, 1, (P_,) * 3, (P_,) * 3)
        info, phi, gp = oracle.nodal_project(G, vel, 1, (P_,) * 3, (P_,) * 3, 0.5, 1e-11, 0.0)
        assert info["status"] == 0 and info["iters"] <= 30, info
        _ghosts_as_apply_projection(layout, G, vel, 1)
        rhs1 = oracle.nodal_rhs(G, vel, 1, (P_,) * 3, (P_,) * 3)
        # approximate projection: the nodal divergence left is O(h^2) of a smooth field's
        assert np.abs(rhs1).max() < 0.05 * np.abs(rhs0).max()
        outs.append((phi, ba.gather(vel, layout.CELL, 3, 1), info["iters"]))
    assert outs[0][2] == outs[1][2]
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


def test_nodal_walled_anisotropic_semicoarsening_converges():
    """ABL-like: periodic x/y, walls in z, dz = dx/5"""
    n = (32, 32, 16)
    h = (20.0, 20.0, 4.0)
    rng = np.random.default_rng(7)
    lobc, hibc = (P_, P_, N_), (P_, P_, N_)
    N = (32, 32, 17)
    b = rng.standard_normal(int(np.prod(N)))
    opts = oracle.default_opts(semicoarsening=1)
    phi, info = oracle.nodal_solve(n, h, lobc, hibc, 1.0, b, 1e-11, 0.0, opts)
    assert info["converged"] and info["iters"] <= 40, info
