"""Pins of the oracle's composite two-level MAC projection (MacProjector over
levels 0 and 1, icns_advection.cpp:88-95,331,353; AMReX's MLMG composite solve,
restated: coarse-fine ghosts quadratic through the tangentially interpolated
coarse value, reflux of the averaged fine fluxes on uncovered coarse cells).

* Second order: u* = grad(psi) of a smooth psi is projected to zero and phi
  recovers psi; the errors of both drop by at least 3x when h halves (a
  dropped tangential slope, cross term or reflux drops the ratio to 2 or less).
* Conservation: the volume sum of the composite operator over uncovered coarse
  and fine cells is zero for periodic / Neumann boundaries (interior fluxes
  telescope only if each coarse-fine face carries the fine average).
* After the projection the composite divergence is zero to solver tolerance,
  and every coarse face under or on the fine level holds the fine average.
* A fine level covering the whole domain gives the single-level projection on
  the fine grid.
"""
import importlib

import numpy as np
import pytest

import oracle

oracle.build()
lay = importlib.import_module("kynema-sgf_amd.layout")
P_, N_, D_ = 0, 1, 2
K = 2 * np.pi


def fine_boxes(n):
    """two fine boxes over the coarse cells [n/4, 3n/4)^3 (fine indices)"""
    a, b = n // 2, 3 * n // 2
    m = (a + b) // 2
    return [[a, a, a, m - 1, b - 1, b - 1], [m, a, a, b - 1, b - 1, b - 1]]


def _faces(m, h, d, fn):
    shp = list(m)
    shp[d] += 1
    ax = [(np.arange(shp[e]) + (0.0 if e == d else 0.5)) * h[e] for e in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return np.ascontiguousarray(fn(X, Y, Z)).ravel()


def _cells(m, h, fn):
    ax = [(np.arange(m[e]) + 0.5) * h[e] for e in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return fn(X, Y, Z)


def _fmask(n, fb):
    fm = np.zeros((2 * n[2], 2 * n[1], 2 * n[0]), dtype=bool)
    for b in np.asarray(fb).reshape(-1, 6):
        fm[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1] = True
    return fm


def _cmask(n, fb):
    return _fmask(n, fb).reshape(n[2], 2, n[1], 2, n[0], 2).all(axis=(1, 3, 5))


def _fine_face_max(umf, n, fb):
    """max |u| over the faces of fine cells"""
    fm = _fmask(n, fb)
    m = 0.0
    for d in range(3):
        shp = [2 * n[2], 2 * n[1], 2 * n[0]]
        shp[2 - d] += 1
        u = umf[d].reshape(shp)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[2 - d], hi[2 - d] = slice(0, -1), slice(1, None)
        m = max(m, np.abs(u[tuple(lo)][fm]).max(), np.abs(u[tuple(hi)][fm]).max())
    return m


PSI = {
    "neumann": (
        ((P_, N_, P_), (P_, N_, P_)),
        lambda X, Y, Z: np.cos(K * X) * np.cos(K * Y) * np.cos(K * Z),
        [lambda X, Y, Z: -K * np.sin(K * X) * np.cos(K * Y) * np.cos(K * Z),
         lambda X, Y, Z: -K * np.cos(K * X) * np.sin(K * Y) * np.cos(K * Z),
         lambda X, Y, Z: -K * np.cos(K * X) * np.cos(K * Y) * np.sin(K * Z)]),
    "dirichlet": (
        ((P_, D_, P_), (P_, D_, P_)),
        lambda X, Y, Z: np.cos(K * X) * np.sin(np.pi * Y) * np.cos(K * Z),
        [lambda X, Y, Z: -K * np.sin(K * X) * np.sin(np.pi * Y) * np.cos(K * Z),
         lambda X, Y, Z: np.pi * np.cos(K * X) * np.cos(np.pi * Y) * np.cos(K * Z),
         lambda X, Y, Z: -K * np.cos(K * X) * np.sin(np.pi * Y) * np.sin(K * Z)]),
}


def _gradient_case(kind, nn):
    (lobc, hibc), psi, grad = PSI[kind]
    n = (nn, nn, nn)
    h = [1.0 / nn] * 3
    hf = [v / 2 for v in h]
    nf = [2 * v for v in n]
    fb = fine_boxes(nn)
    umc = [_faces(n, h, d, grad[d]) for d in range(3)]
    umf = [_faces(nf, hf, d, grad[d]) for d in range(3)]
    info, pc, pf = oracle.mac_project_two_levels(n, h, lobc, hibc, fb, 1.0, umc, umf, 1e-11, 0.0)
    assert info["status"] == 0 and info["iters"] <= 30, info
    fm = _fmask(n, fb)
    pce, pfe = _cells(n, h, psi), _cells(nf, hf, psi)
    pcr, pfr = pc.reshape(pce.shape), pf.reshape(pfe.shape)
    off = np.mean(pcr - pce) if kind == "neumann" else 0.0
    cm = _cmask(n, fb)
    return dict(uf=_fine_face_max(umf, n, fb), uc=max(np.abs(a).max() for a in umc),
                pf=np.abs(pfr - pfe - off)[fm].max(), pc=np.abs(pcr - pce - off)[~cm].max())


@pytest.mark.parametrize("kind", ["neumann", "dirichlet"])
def test_composite_second_order(kind):
    e16, e32 = _gradient_case(kind, 16), _gradient_case(kind, 32)
    for key in ("uf", "uc", "pf", "pc"):
        assert e16[key] / e32[key] >= 3.0, (key, e16, e32)
    assert e32["uf"] < 1e-2 and e32["pf"] < 1e-3


@pytest.mark.parametrize("bcs", [((P_, P_, P_), (P_, P_, P_)), ((P_, N_, P_), (N_, N_, P_))])
def test_composite_operator_conserves(bcs):
    n = (12, 8, 16)
    h = (0.5, 0.75, 0.25)
    fb = [[4, 0, 8, 11, 7, 15], [12, 0, 8, 15, 3, 23], [20, 12, 0, 23, 15, 3]]  # reaches the y wall and z side
    rng = np.random.default_rng(3)
    pc = rng.standard_normal(int(np.prod(n)))
    pf = rng.standard_normal(8 * int(np.prod(n)))
    Lc, Lf = oracle.tl_apply(n, h, bcs[0], bcs[1], fb, 0.7, pc, pf)
    fm, cm = _fmask(n, fb).ravel(), _cmask(n, fb).ravel()
    tot = 8.0 * Lc[~cm].sum() + Lf[fm].sum()
    scale = 8.0 * np.abs(Lc[~cm]).sum() + np.abs(Lf[fm]).sum()
    assert abs(tot) <= 1e-13 * scale
    assert np.all(Lc[cm] == 0.0) and np.all(Lf[~fm] == 0.0)


def _div(u, m, h):
    ux = u[0].reshape(m[2], m[1], m[0] + 1)
    uy = u[1].reshape(m[2], m[1] + 1, m[0])
    uz = u[2].reshape(m[2] + 1, m[1], m[0])
    return (np.diff(ux, axis=2) / h[0] + np.diff(uy, axis=1) / h[1]) + np.diff(uz, axis=0) / h[2]


@pytest.mark.parametrize("bcs", [((P_, N_, P_), (P_, N_, P_)), ((P_, D_, P_), (P_, N_, P_))])
def test_composite_projection_divergence_free(bcs):
    n = (16, 12, 16)
    h = (1.0 / 16, 0.75 / 12, 1.0 / 16)
    hf = [v / 2 for v in h]
    nf = [2 * v for v in n]
    fb = [[0, 4, 8, 7, 11, 23], [24, 8, 8, 31, 19, 15], [8, 12, 24, 23, 23, 31]]  # periodic sides, y walls
    rng = np.random.default_rng(11)
    umc, umf = [], []
    for m, lst in ((n, umc), (nf, umf)):
        for d in range(3):
            shp = [m[2], m[1], m[0]]
            shp[2 - d] += 1
            a = rng.standard_normal(shp)
            if d == 1:  # no flow through the y walls
                a[:, 0, :] = 0.0
                a[:, -1, :] = 0.0
            if d != 1:  # periodic: the image faces agree
                sl = [slice(None)] * 3
                sl[2 - d] = -1
                sl0 = list(sl)
                sl0[2 - d] = 0
                a[tuple(sl)] = a[tuple(sl0)]
            lst.append(np.ascontiguousarray(a).ravel())
    info, pc, pf = oracle.mac_project_two_levels(n, h, bcs[0], bcs[1], fb, 1.3, umc, umf, 1e-11, 0.0)
    assert info["status"] == 0 and info["iters"] <= 30, info
    fm, cm = _fmask(n, fb), _cmask(n, fb)
    dc, df = _div(umc, n, h), _div(umf, nf, hf)
    # the residual target is 1e-11 of the initial residual (= |div u*|)
    assert np.abs(df[fm]).max() <= 1e-9 * info["r0"] and np.abs(dc[~cm]).max() <= 1e-9 * info["r0"]
    # covered coarse cells: the average of the fine divergences (fine faces averaged)
    avg = df.reshape(n[2], 2, n[1], 2, n[0], 2).mean(axis=(1, 3, 5))
    assert np.abs(dc[cm] - avg[cm]).max() <= 1e-9 * info["r0"]


def test_composite_full_cover_equals_single_level():
    n = (8, 8, 8)
    h = (0.125, 0.125, 0.125)
    nf = (16, 16, 16)
    hf = (0.0625,) * 3
    lobc, hibc = (P_, D_, P_), (P_, N_, P_)
    fb = [[0, 0, 0, 7, 15, 15], [8, 0, 0, 15, 15, 15]]
    rng = np.random.default_rng(2)
    umf = []
    for d in range(3):
        shp = [16, 16, 16]
        shp[2 - d] += 1
        a = rng.standard_normal(shp)
        if d == 1:
            a[:, -1, :] = 0.0
        else:
            sl = [slice(None)] * 3
            sl[2 - d] = -1
            sl0 = list(sl)
            sl0[2 - d] = 0
            a[tuple(sl)] = a[tuple(sl0)]
        umf.append(np.ascontiguousarray(a).ravel())
    umc = [np.zeros(((8 + (d == 2)) * (8 + (d == 1)) * (8 + (d == 0)),)) for d in range(3)]
    ref = [a.copy() for a in umf]
    info, pc, pf = oracle.mac_project_two_levels(n, h, lobc, hibc, fb, 1.0, umc, umf, 1e-12, 0.0)
    assert info["status"] == 0, info
    g = lay.Geometry(nf, prob_hi=(1.0, 1.0, 1.0), periodic=(1, 0, 1))
    ba = lay.BoxArray(g, [[0, 0, 0, 15, 15, 15]])
    shp = [[16 + (d == 2), 16 + (d == 1), 16 + (d == 0)] for d in range(3)]
    mac = [ba.scatter(ref[d].reshape(shp[d])[None], 1 + d, 1) for d in range(3)]  # face MultiFabs, 1 ghost
    sinfo, _ = oracle.mac_project(oracle.Geom(ba), mac, lobc, hibc, 1.0, 1e-12, 0.0)
    assert sinfo["status"] == 0
    for d in range(3):
        single = ba.gather(mac[d], 1 + d, 1, 1)[0]
        assert np.abs(umf[d].reshape(shp[d]) - single).max() <= 1e-9, d
