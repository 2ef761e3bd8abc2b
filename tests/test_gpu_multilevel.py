"""Two-level averaging on the GPU (ksgf_average_down, ksgf_average_down_faces)
against the CPU oracle: bit for bit (same child summation order), on a ragged
coarse level with a fine level at ratio 2 over part of the domain, periodic in
x and z (so the image faces are exercised where the fine level reaches the
periodic sides)."""
import importlib

import numpy as np
import pytest

from helpers import to_np

pytestmark = pytest.mark.gpu

lay = importlib.import_module("kynema-sgf_amd.layout")
CELL, XFACE = 0, 1


def _say(msg):
    print(msg, flush=True)


def _levels(reach_sides):
    import test_multilevel_oracle as T

    rng = np.random.default_rng(21)
    n = (16, 12, 16)
    cg = lay.Geometry(n, prob_hi=(1.0, 0.75, 1.0), periodic=(1, 0, 1))
    cba = lay.BoxArray(cg, lay.chop_random(n, [4, 8], rng, granularity=4))
    fg = lay.Geometry(tuple(2 * v for v in n), prob_hi=cg.prob_hi, periodic=(1, 0, 1))
    if reach_sides:  # fine boxes touching the periodic low/high sides in x and the y walls
        boxes = [[0, 0, 8, 7, 7, 15], [24, 16, 8, 31, 23, 15], [8, 8, 0, 15, 15, 7], [16, 8, 24, 23, 15, 31]]
    else:
        boxes = [[8, 4, 8, 15, 11, 15], [16, 4, 8, 23, 11, 15], [8, 12, 8, 23, 19, 15], [8, 4, 16, 23, 19, 23]]
    return cba, T.FineBoxArray(fg, boxes)


@pytest.mark.parametrize("reach_sides", [False, True])
def test_average_down_bitexact(ksgf, oracle, gpu, reach_sides):
    _say(f"average_down reach_sides={reach_sides}")
    cba, fba = _levels(reach_sides)
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    rng = np.random.default_rng(5)
    fine = rng.standard_normal(fba.mf_size(CELL, 3, 1))
    crse = rng.standard_normal(cba.mf_size(CELL, 3, 2))
    ref = crse.copy()
    oracle.average_down(Gf, fine, 1, Gc, ref, 2, 3, 2)
    d = dc.upload(crse.copy())
    df.average_down_to(df.upload(fine), 1, dc, d, 2, 3, 2)
    assert np.array_equal(to_np(d), ref)


@pytest.mark.parametrize("reach_sides", [False, True])
def test_average_down_faces_bitexact(ksgf, oracle, gpu, reach_sides):
    _say(f"average_down_faces reach_sides={reach_sides}")
    cba, fba = _levels(reach_sides)
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    rng = np.random.default_rng(6)
    nf = fba.geom.n_cell
    fine, crse = [], []
    for d in range(3):
        shp = [2, nf[2], nf[1], nf[0]]
        shp[3 - d] += 1
        fine.append(fba.scatter(rng.standard_normal(shp), XFACE + d, 0))
        crse.append(rng.standard_normal(cba.mf_size(XFACE + d, 2, 1)))
    ref = [c.copy() for c in crse]
    oracle.average_down_faces(Gf, fine, 0, Gc, ref, 1, 2, 2)
    dcr = [dc.upload(c.copy()) for c in crse]
    df.average_down_faces_to([df.upload(f) for f in fine], 0, dc, dcr, 1, 2, 2)
    for d in range(3):
        assert np.array_equal(to_np(dcr[d]), ref[d]), f"dir {d}"


@pytest.mark.parametrize("reach_sides", [False, True])
def test_fillpatch_two_levels_bitexact(ksgf, oracle, gpu, reach_sides):
    """ghost cells of the fine level from the coarse level (conservative
    linear interpolation), then the fine level's own data, then the physical
    BCs (y walls: Dirichlet in x/z velocity, extrapolation in y)"""
    _say(f"fillpatch_two_levels reach_sides={reach_sides}")
    cba, fba = _levels(reach_sides)
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    rng = np.random.default_rng(7)
    bc = np.array([[0, 3, 0, 0, 3, 0], [0, 2, 0, 0, 2, 0], [0, 3, 0, 0, 3, 0]], dtype=np.int32)
    bcval = rng.standard_normal((6, 3))
    nc, nf = cba.geom.n_cell, fba.geom.n_cell
    crse = cba.scatter(rng.standard_normal((3, nc[2], nc[1], nc[0])), CELL, 1)
    oracle.fill_ghosts(Gc, crse, CELL, 3, 1, bc, bcval)
    for ng in (1, 3):
        fine = fba.scatter(rng.standard_normal((3, nf[2], nf[1], nf[0])), CELL, ng)
        ref = fine.copy()
        oracle.fillpatch_two_levels(Gf, ref, ng, Gc, crse, 1, 3, 2, bc, bcval)
        d = df.upload(fine)
        df.fillpatch_two_levels(d, ng, dc, dc.upload(crse), 1, 3, bc, bcval)
        got = to_np(d)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(np.nan_to_num(got), np.nan_to_num(ref)), f"ng {ng}"


@pytest.mark.parametrize("reach_sides", [False, True])
def test_fillpatch_faces_two_levels_bitexact(ksgf, oracle, gpu, reach_sides):
    """ghost faces of the fine level from the coarse faces (face-linear), then
    the fine level's own faces, then the physical BCs, every direction"""
    _say(f"fillpatch_faces_two_levels reach_sides={reach_sides}")
    cba, fba = _levels(reach_sides)
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    rng = np.random.default_rng(8)
    nc, nf = cba.geom.n_cell, fba.geom.n_cell
    for d in range(3):
        bc = np.array([[0, 3, 0, 0, 3, 0], [0, 2, 0, 0, 2, 0], [0, 3, 0, 0, 3, 0]], dtype=np.int32)[d:d + 1]
        bcval = rng.standard_normal((6, 1))
        shp_c = [2, nc[2], nc[1], nc[0]]
        shp_c[3 - d] += 1
        crse = cba.scatter(rng.standard_normal(shp_c), XFACE + d, 1)
        shp_f = [2, nf[2], nf[1], nf[0]]
        shp_f[3 - d] += 1
        fine = fba.scatter(rng.standard_normal(shp_f), XFACE + d, 2)
        ref = fine.copy()
        oracle.fillpatch_faces_two_levels(Gf, ref, XFACE + d, 2, Gc, crse, 1, 2, np.repeat(bc, 2, axis=0),
                                          np.repeat(bcval, 2, axis=1))
        t = df.upload(fine)
        df.fillpatch_faces_two_levels(t, XFACE + d, 2, dc, dc.upload(crse), 1, 2, np.repeat(bc, 2, axis=0),
                                      np.repeat(bcval, 2, axis=1))
        assert np.array_equal(to_np(t), ref), f"dir {d}"
