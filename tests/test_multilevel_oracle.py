"""Pins of the oracle's two-level averaging (amrex::average_down on cell data,
incflo_apply_nodal_projection.cpp:457-460; amrex::average_down_faces on the
advective fluxes, icns_advection.H:401-407) on a ragged two-level layout
(fine level at ratio 2 over part of the domain).

* Linear fields: the average of the children is the value at the parent's
  centre (face centre), exactly up to rounding.
* Flux conservation: the divergence of the averaged face fluxes over a covered
  coarse cell equals the average of the fine divergences of its 8 children
  (the interior fine faces telescope).
* Coarse cells and faces without fine data keep their values; a coarse face
  shared by two coarse boxes holds the same value in both FABs.
"""
import importlib

import numpy as np
import pytest

import oracle

lay = importlib.import_module("kynema-sgf_amd.layout")
oracle.build()
CELL, XFACE = 0, 1


def two_level(seed=4, n=(16, 12, 16), periodic=(1, 0, 1)):
    """coarse BoxArray (ragged) and a fine BoxArray at ratio 2 covering the
    coarse region [4, 12) x [2, 10) x [4, 12) (fine boxes of 4 or 8 cells)"""
    rng = np.random.default_rng(seed)
    cg = lay.Geometry(n, prob_hi=(1.0, 0.75, 1.0), periodic=periodic)
    cba = lay.BoxArray(cg, lay.chop_random(n, [4, 8], rng, granularity=4))
    nf = tuple(2 * v for v in n)
    fg = lay.Geometry(nf, prob_hi=cg.prob_hi, periodic=periodic)
    lo, hi = (8, 4, 8), (24, 20, 24)  # fine indices of the refined region
    boxes = []
    for z0 in range(lo[2], hi[2], 8):
        for y0 in range(lo[1], hi[1], 8):
            x0 = lo[0]
            while x0 < hi[0]:
                w = int(rng.choice([4, 8]))
                w = min(w, hi[0] - x0)
                boxes.append([x0, y0, z0, x0 + w - 1, min(y0 + 7, hi[1] - 1), min(z0 + 7, hi[2] - 1)])
                x0 += w
    fba = FineBoxArray(fg, boxes)
    return cba, fba


class FineBoxArray(lay.BoxArray):
    """a BoxArray that covers part of its domain (the fine level)"""

    def __init__(self, geom, boxes):
        self.geom = geom
        self.boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self.nbox = self.boxes.shape[0]


def _fill_cells(ba, ncomp, ng, fn):
    out = np.zeros(ba.mf_size(CELL, ncomp, ng))
    h = ba.geom.dx
    for b in range(ba.nbox):
        v = ba.fab_view(out, b, CELL, ncomp, ng)
        lo = ba.boxes[b, :3] - ng
        nz, ny, nx = v.shape[1:]
        z, y, x = np.meshgrid(*[(lo[d] + np.arange(m) + 0.5) * h[d] for d, m in ((2, nz), (1, ny), (0, nx))],
                              indexing="ij")
        for c in range(ncomp):
            v[c] = fn(c, x, y, z)
    return out


def test_average_down_linear_exact_and_untouched():
    cba, fba = two_level()
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    lin = lambda c, x, y, z: 1.0 + (c + 1) * x - 2.0 * y + 0.5 * z * (c - 1)  # noqa: E731
    fine = _fill_cells(fba, 2, 1, lin)
    crse = np.full(cba.mf_size(CELL, 2, 1), -7.0)
    oracle.average_down(Gf, fine, 1, Gc, crse, 1, 2, 2)
    exact = _fill_cells(cba, 2, 1, lin)
    cov = np.zeros(tuple(cba.geom.n_cell[::-1]), dtype=bool)
    cov[4:12, 2:10, 4:12] = True
    got = cba.gather(crse, CELL, 2, 1)
    ref = cba.gather(exact, CELL, 2, 1)
    assert np.abs(got[:, cov] - ref[:, cov]).max() <= 1e-14 * np.abs(ref).max()
    assert np.all(got[:, ~cov] == -7.0)


def test_average_down_faces_conserves_flux_divergence():
    cba, fba = two_level(seed=7)
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    rng = np.random.default_rng(3)
    nf = fba.geom.n_cell
    fine = []
    for d in range(3):  # one domain-wide field per direction: shared faces agree
        shp = [1, nf[2], nf[1], nf[0]]
        shp[3 - d] += 1
        fine.append(fba.scatter(rng.standard_normal(shp), XFACE + d, 0))
    crse = [np.zeros(cba.mf_size(XFACE + d, 1, 0)) for d in range(3)]
    oracle.average_down_faces(Gf, fine, 0, Gc, crse, 0, 1, 2)
    # fine divergences over each fine box, averaged to the parents
    hf, hc = fba.geom.dx, cba.geom.dx
    nfc = tuple(cba.geom.n_cell[::-1])
    avg_div = np.zeros(nfc)
    for b in range(fba.nbox):
        fx = fba.fab_view(fine[0], b, XFACE, 1, 0)[0]
        fy = fba.fab_view(fine[1], b, XFACE + 1, 1, 0)[0]
        fz = fba.fab_view(fine[2], b, XFACE + 2, 1, 0)[0]
        div = (fx[:, :, 1:] - fx[:, :, :-1]) / hf[0] + (fy[:, 1:, :] - fy[:, :-1, :]) / hf[1] + \
              (fz[1:, :, :] - fz[:-1, :, :]) / hf[2]
        lo = fba.boxes[b, :3] // 2
        nz, ny, nx = div.shape
        c = div.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).mean(axis=(1, 3, 5))
        avg_div[lo[2]:lo[2] + nz // 2, lo[1]:lo[1] + ny // 2, lo[0]:lo[0] + nx // 2] = c
    cx = cba.gather(crse[0], XFACE, 1, 0)[0]
    cy = cba.gather(crse[1], XFACE + 1, 1, 0)[0]
    cz = cba.gather(crse[2], XFACE + 2, 1, 0)[0]
    cdiv = (cx[:, :, 1:] - cx[:, :, :-1]) / hc[0] + (cy[:, 1:, :] - cy[:, :-1, :]) / hc[1] + \
           (cz[1:, :, :] - cz[:-1, :, :]) / hc[2]
    cov = np.zeros(nfc, dtype=bool)
    cov[4:12, 2:10, 4:12] = True
    assert np.abs(cdiv[cov] - avg_div[cov]).max() <= 1e-12 * np.abs(avg_div[cov]).max()


def test_average_down_faces_linear_and_shared_faces():
    cba, fba = two_level(seed=9)
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    hf = fba.geom.dx
    fine, crse = [], []
    for d in range(3):
        f = np.zeros(fba.mf_size(XFACE + d, 1, 0))
        for b in range(fba.nbox):
            v = fba.fab_view(f, b, XFACE + d, 1, 0)[0]
            lo = fba.boxes[b, :3]
            nz, ny, nx = v.shape
            idx = np.meshgrid(np.arange(nz) + lo[2], np.arange(ny) + lo[1], np.arange(nx) + lo[0], indexing="ij")
            # position of the face centre: nodal along d, cell-centred across
            pos = [(idx[2 - e] + (0.0 if e == d else 0.5)) * hf[e] for e in range(3)]
            v[...] = 2.0 * pos[0] - pos[1] + 3.0 * pos[2]
        fine.append(f)
        crse.append(np.full(cba.mf_size(XFACE + d, 1, 0), np.nan))
    oracle.average_down_faces(Gf, fine, 0, Gc, crse, 0, 1, 2)
    hc = cba.geom.dx
    for d in range(3):
        for b in range(cba.nbox):
            v = cba.fab_view(crse[d], b, XFACE + d, 1, 0)[0]
            lo = cba.boxes[b, :3]
            nz, ny, nx = v.shape
            idx = np.meshgrid(np.arange(nz) + lo[2], np.arange(ny) + lo[1], np.arange(nx) + lo[0], indexing="ij")
            pos = [(idx[2 - e] + (0.0 if e == d else 0.5)) * hc[e] for e in range(3)]
            ref = 2.0 * pos[0] - pos[1] + 3.0 * pos[2]
            m = ~np.isnan(v)
            if m.any():
                assert np.abs(v[m] - ref[m]).max() <= 1e-13
        # every coarse FAB holding a covered face was written (shared faces in both)
        clo, chi = np.array([4, 2, 4]), np.array([11, 9, 11])
        chi[d] += 1
        expect = 0
        for b in range(cba.nbox):
            blo, bhi = cba.boxes[b, :3].copy(), cba.boxes[b, 3:].copy()
            bhi[d] += 1
            ext = np.minimum(bhi, chi) - np.maximum(blo, clo) + 1
            expect += int(np.prod(np.maximum(ext, 0)))
        assert int(np.isfinite(crse[d]).sum()) == expect
