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


# --------------------------------------------------------------------------
# FillPatchTwoLevels (Field::fillpatch on level > 0, FieldFillPatchOps.H:272-336):
# conservative linear interpolation of the ghost cells from the coarse level
# --------------------------------------------------------------------------
def _one_box_levels(seed=11):
    """ragged coarse level 16x12x16 and one fine box over coarse [4,7]^3, so
    every ghost cell of the fine box (ng 2) lies inside the domain and its
    parent's face neighbours lie in the coarse valid region"""
    rng = np.random.default_rng(seed)
    n = (16, 12, 16)
    cg = lay.Geometry(n, prob_hi=(1.0, 0.75, 1.0), periodic=(1, 0, 1))
    cba = lay.BoxArray(cg, lay.chop_random(n, [4, 8], rng, granularity=4))
    fg = lay.Geometry(tuple(2 * v for v in n), prob_hi=cg.prob_hi, periodic=(1, 0, 1))
    return cba, FineBoxArray(fg, [[8, 8, 8, 15, 15, 15]])


_BC = np.array([[0, 3, 0, 0, 3, 0]], dtype=np.int32)


def _fillpatch(cba, fba, crse_dom, fine_dom, ng=2):
    """run the oracle's fill-patch; crse_dom/fine_dom: domain-wide (nz, ny, nx)"""
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    crse = cba.scatter(crse_dom[None], CELL, 1)
    oracle.fill_ghosts(Gc, crse, CELL, 1, 1, _BC, np.zeros((6, 1)))
    fine = np.full(fba.mf_size(CELL, 1, ng), np.nan)
    v = fba.fab_view(fine, 0, CELL, 1, ng)[0]
    lo, hi = fba.boxes[0, :3], fba.boxes[0, 3:]
    v[ng:-ng, ng:-ng, ng:-ng] = fine_dom[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
    oracle.fillpatch_two_levels(Gf, fine, ng, Gc, crse, 1, 1, 2, _BC, np.zeros((6, 1)))
    return fba.fab_view(fine, 0, CELL, 1, ng)[0], lo - ng


def test_fillpatch_linear_field_exact():
    """MC slopes of a linear field are exact and the bounding factor stays 1,
    so the ghosts take the linear field at the fine cell centres"""
    cba, fba = _one_box_levels()
    lin = lambda x, y, z: 0.3 + 2.0 * x - 1.5 * y + 0.75 * z  # noqa: E731

    def dom(g):
        n = g.n_cell
        z, y, x = np.meshgrid(*[(np.arange(n[d]) + 0.5) * g.dx[d] for d in (2, 1, 0)], indexing="ij")
        return lin(x, y, z)

    v, lo = _fillpatch(cba, fba, dom(cba.geom), dom(fba.geom))
    h = fba.geom.dx
    z, y, x = np.meshgrid(*[(lo[d] + np.arange(v.shape[2 - d]) + 0.5) * h[d] for d in (2, 1, 0)], indexing="ij")
    assert np.abs(v - lin(x, y, z)).max() <= 1e-14


def test_fillpatch_conservative_and_bounded():
    """random coarse data: the 8 children of every coarse cell whose children
    are all ghosts average to the parent, and no child leaves the range of its
    parent and the parent's 6 face neighbours"""
    cba, fba = _one_box_levels(seed=12)
    rng = np.random.default_rng(8)
    cd = rng.standard_normal(tuple(cba.geom.n_cell[::-1]))
    v, lo = _fillpatch(cba, fba, cd, rng.standard_normal(tuple(fba.geom.n_cell[::-1])))
    nchecked = 0
    for K in range(3, 9):
        for J in range(3, 9):
            for I in range(3, 9):
                if 4 <= I <= 7 and 4 <= J <= 7 and 4 <= K <= 7:
                    continue
                ch = v[2 * K - lo[2]:2 * K - lo[2] + 2, 2 * J - lo[1]:2 * J - lo[1] + 2,
                       2 * I - lo[0]:2 * I - lo[0] + 2]
                nb = [cd[K, J, I], cd[K, J, I - 1], cd[K, J, I + 1], cd[K, J - 1, I], cd[K, J + 1, I],
                      cd[K - 1, J, I], cd[K + 1, J, I]]
                assert ch.min() >= min(nb) - 1e-15 and ch.max() <= max(nb) + 1e-15
                assert abs(ch.mean() - cd[K, J, I]) <= 1e-15 * 8
                nchecked += 1
    assert nchecked == 6 ** 3 - 4 ** 3
    assert not np.isnan(v).any()


def test_fillpatch_known_answers():
    """hand-worked parents on the fine box's low-x ghost slab (children
    i = 6, 7 at offsets -1/4, +1/4).
    A = coarse (3,5,5): x neighbours 0 | 1 | 4, y flat, z 1 | 1 | 0.9: MC slope
    in x min(2, 2, 6) = 2 (minmod would give 1), none in y or z, bound factor
    1: children 0.5 and 1.5.
    B = coarse (3,5,7): neighbours 0.9 | 1 | 1.3 in every direction: slopes
    0.2 each, deviation 3 * 0.2 / 4 = 0.15 > 1 - 0.9, so alpha = 2/3 and the
    children are 1 + (+-1 +-1 +-1) / 30 (extremes exactly 0.9 and 1.1)."""
    cba, fba = _one_box_levels(seed=13)
    cd = np.ones(tuple(cba.geom.n_cell[::-1]))  # cd[K, J, I]
    cd[5, 5, 2], cd[5, 5, 4] = 0.0, 4.0
    cd[7, 5, 2], cd[7, 5, 4] = 0.9, 1.3
    cd[7, 4, 3], cd[7, 6, 3] = 0.9, 1.3
    cd[6, 5, 3], cd[8, 5, 3] = 0.9, 1.3
    v, lo = _fillpatch(cba, fba, cd, np.zeros(tuple(fba.geom.n_cell[::-1])))

    def kids(I, J, K):
        return v[2 * K - lo[2]:2 * K - lo[2] + 2, 2 * J - lo[1]:2 * J - lo[1] + 2, 2 * I - lo[0]:2 * I - lo[0] + 2]

    a = kids(3, 5, 5)
    assert np.abs(a[:, :, 0] - 0.5).max() <= 1e-15 and np.abs(a[:, :, 1] - 1.5).max() <= 1e-15
    s = np.array([-1.0, 1.0])
    ref = 1.0 + (s[:, None, None] + s[None, :, None] + s[None, None, :]) / 30.0
    assert np.abs(kids(3, 5, 7) - ref).max() <= 1e-14


# --------------------------------------------------------------------------
# face fill-patch (fillpatch_sibling_fields on the MAC velocities,
# icns_advection.H:231-237): face-linear interpolation at ratio 2
# --------------------------------------------------------------------------
@pytest.mark.parametrize("d", [0, 1, 2])
def test_fillpatch_faces_linear_normal_exact_and_fine_data_kept(d):
    """a face field linear in its normal coordinate (constant across) is
    reproduced exactly on every fine ghost face inside the domain (coarse-face
    copies and midpoint means are exact for it), valid faces are untouched, and
    a ghost face another fine box holds carries that box's value"""
    cba, fba = two_level(seed=31)
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    fn = lambda x: 0.25 + 3.0 * x  # noqa: E731

    def faces(ba, ng):
        out = np.zeros(ba.mf_size(XFACE + d, 1, ng))
        h = ba.geom.dx[d]
        for b in range(ba.nbox):
            v = ba.fab_view(out, b, XFACE + d, 1, ng)[0]
            lo = ba.boxes[b, d] - ng
            shape = [1, 1, 1]
            shape[2 - d] = v.shape[2 - d]
            x = ((lo + np.arange(v.shape[2 - d])) * h).reshape(shape)
            v[...] = fn(x)
        return out

    ng = 2
    crse = faces(cba, 0)
    ref = faces(fba, ng)
    fine = np.full(fba.mf_size(XFACE + d, 1, ng), np.nan)
    for b in range(fba.nbox):  # valid faces only
        vf = fba.fab_view(fine, b, XFACE + d, 1, ng)[0]
        vr = fba.fab_view(ref, b, XFACE + d, 1, ng)[0]
        sl = tuple(slice(ng, -ng) for _ in range(3))
        vf[sl] = vr[sl]
    bc = np.zeros((1, 6), dtype=np.int32)
    bc[0, 1] = bc[0, 4] = 2  # y walls: first-order extrapolation
    oracle.fillpatch_faces_two_levels(Gf, fine, XFACE + d, ng, Gc, crse, 0, 1, bc, np.zeros((6, 1)))
    nf = fba.geom.n_cell
    for b in range(fba.nbox):
        vf = fba.fab_view(fine, b, XFACE + d, 1, ng)[0]
        vr = fba.fab_view(ref, b, XFACE + d, 1, ng)[0]
        lo = fba.boxes[b, :3] - ng
        nz, ny, nx = vf.shape
        idx = np.meshgrid(np.arange(nz) + lo[2], np.arange(ny) + lo[1], np.arange(nx) + lo[0], indexing="ij")
        inside = np.ones(vf.shape, dtype=bool)
        if d != 1:  # y is the only wall direction: faces beyond it take the physical BC
            inside &= (idx[1] >= 0) & (idx[1] <= nf[1] - 1)
        else:
            inside &= (idx[1] >= 0) & (idx[1] <= nf[1])
        assert np.abs(vf[inside] - vr[inside]).max() <= 1e-14, b


# --------------------------------------------------------------------------
# one two-level advect + MAC-project step (every operator over both levels)
# --------------------------------------------------------------------------
def tgv_two_level(n=16, seed=2):
    """Taylor-Green on a periodic n^3 level 0 (boxes of n/2) and a level 1 at
    ratio 2 over the coarse cells [n/4, 3n/4)^3 in three boxes; the covered
    coarse cells hold the average of the fine velocity"""
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    case = wl.taylor_green(n, max_grid_size=n // 2, noise=0.0)
    fg = lay.Geometry((2 * n,) * 3, prob_hi=(1.0, 1.0, 1.0), periodic=(1, 1, 1))
    a, b, m = n // 2, 3 * n // 2, n
    fba = FineBoxArray(fg, [[a, a, a, m - 1, b - 1, b - 1], [m, a, a, b - 1, m - 1, b - 1],
                            [m, m, a, b - 1, b - 1, b - 1]])
    x = fg.cell_centers(0)[None, None, :]
    y = fg.cell_centers(1)[None, :, None]
    z = fg.cell_centers(2)[:, None, None]
    tp = 2.0 * np.pi
    rng = np.random.default_rng(seed)
    u = np.sin(tp * x) * np.cos(tp * y) * np.cos(tp * z)
    v = -np.cos(tp * x) * np.sin(tp * y) * np.cos(tp * z)
    vf = np.stack(np.broadcast_arrays(u, v, 0.1 * u * v)).astype(np.float64) + 1e-3 * rng.standard_normal((3,) + u.shape)
    vel_f = fba.scatter(vf, CELL, 3)
    vel_c = case.ba.scatter(case.vel, CELL, 3)
    oracle.average_down(oracle.Geom(fba), vel_f, 3, oracle.Geom(case.ba), vel_c, 3, 3, 2)
    gp_c = np.zeros(case.ba.mf_size(CELL, 3, 0))
    gp_f = np.zeros(fba.mf_size(CELL, 3, 0))
    return case, fba, vel_c, vel_f, gp_c, gp_f


def tgv_two_level_walls(n=16, seed=3):
    """as tgv_two_level, with no-slip walls in y (velocity ext_dir 0, MAC
    Neumann) and level-1 boxes that touch the low wall"""
    import dataclasses

    wl = importlib.import_module("kynema-sgf_amd.workloads")
    g = lay.Geometry((n,) * 3, prob_hi=(1.0, 1.0, 1.0), periodic=(1, 0, 1))
    ba = lay.BoxArray(g, lay.chop_uniform(g.n_cell, n // 2))
    base = wl.taylor_green(n, max_grid_size=n // 2, noise=0.0)
    bcrec = np.array([[0, 3, 0, 0, 3, 0]] * 3, dtype=np.int32)
    case = dataclasses.replace(base, name=f"tgv{n}_ywalls", geom=g, ba=ba, bcrec=bcrec, bcval=np.zeros((6, 3)),
                               lobc=(0, 1, 0), hibc=(0, 1, 0))
    fg = lay.Geometry((2 * n,) * 3, prob_hi=(1.0, 1.0, 1.0), periodic=(1, 0, 1))
    a, b, m = n // 2, 3 * n // 2, n
    fba = FineBoxArray(fg, [[a, 0, a, m - 1, m // 2 - 1, b - 1], [m, 0, a, b - 1, m - 1, b - 1],
                            [a, m // 2, a, m - 1, m - 1, b - 1]])
    x = fg.cell_centers(0)[None, None, :]
    y = fg.cell_centers(1)[None, :, None]
    z = fg.cell_centers(2)[:, None, None]
    tp = 2.0 * np.pi
    rng = np.random.default_rng(seed)
    u = np.sin(tp * x) * np.sin(np.pi * y) * np.cos(tp * z)
    w = np.cos(tp * x) * np.sin(np.pi * y) * np.sin(tp * z)
    vf = np.stack(np.broadcast_arrays(u, 0.0 * u, w)).astype(np.float64) + 1e-3 * rng.standard_normal((3,) + u.shape)
    xc = g.cell_centers(0)[None, None, :]
    yc = g.cell_centers(1)[None, :, None]
    zc = g.cell_centers(2)[:, None, None]
    uc = np.sin(tp * xc) * np.sin(np.pi * yc) * np.cos(tp * zc)
    wc = np.cos(tp * xc) * np.sin(np.pi * yc) * np.sin(tp * zc)
    vel_c = ba.scatter(np.stack(np.broadcast_arrays(uc, 0.0 * uc, wc)).astype(np.float64), CELL, 3)
    vel_f = fba.scatter(vf, CELL, 3)
    oracle.average_down(oracle.Geom(fba), vel_f, 3, oracle.Geom(ba), vel_c, 3, 3, 2)
    return case, fba, vel_c, vel_f, np.zeros(ba.mf_size(CELL, 3, 0)), np.zeros(fba.mf_size(CELL, 3, 0))


@pytest.mark.parametrize("walls", [False, True])
def test_two_level_step_conserves_momentum(walls):
    """no pressure gradient, periodic or no-slip walls (no advective flux
    through a wall): the level-0 total momentum (covered cells averaged down)
    is conserved to rounding over a two-level step, which needs the fine
    fluxes on the coarse faces of the coarse-fine boundary
    (average_down_faces) and the final average_down"""
    case, fba, vel_c, vel_f, gp_c, gp_f = tgv_two_level_walls() if walls else tgv_two_level()
    tot0 = case.ba.gather(vel_c, CELL, 3, 3).sum(axis=(1, 2, 3))
    scale = np.abs(case.ba.gather(vel_c, CELL, 3, 3)).sum()
    vc1, vf1, info = oracle.two_level_step(case, fba, vel_c, vel_f, gp_c, gp_f)
    assert info["status"] == 0, info
    tot1 = case.ba.gather(vc1, CELL, 3, 3).sum(axis=(1, 2, 3))
    assert np.abs(tot1 - tot0).max() <= 1e-14 * scale, (tot0, tot1)
    # the step moved the field
    assert np.abs(case.ba.gather(vc1, CELL, 3, 3) - case.ba.gather(vel_c, CELL, 3, 3)).max() > 1e-3
    # covered coarse cells are the average of the new fine velocity
    chk = vc1.copy()
    oracle.average_down(oracle.Geom(fba), vf1, 3, oracle.Geom(case.ba), chk, 3, 3, 2)
    assert np.array_equal(chk, vc1)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_fillpatch_faces_shared_faces_keep_fine_values(d):
    """a ghost face that is a valid face of another fine box -- including that
    box's high face on the coarse-fine boundary, which no fine cell beyond it
    owns -- takes that box's value, not the coarse interpolation"""
    case, fba, *_ = tgv_two_level()
    cba = case.ba
    Gc, Gf = oracle.Geom(cba), oracle.Geom(fba)
    rng = np.random.default_rng(12 + d)
    nc, nf = cba.geom.n_cell, fba.geom.n_cell
    shp_c = [1, nc[2], nc[1], nc[0]]
    shp_c[3 - d] += 1
    crse = cba.scatter(rng.standard_normal(shp_c), XFACE + d, 0)
    shp_f = [1, nf[2], nf[1], nf[0]]
    shp_f[3 - d] += 1
    dom = rng.standard_normal(shp_f)
    fine = fba.scatter(dom, XFACE + d, 1)
    oracle.fillpatch_faces_two_levels(Gf, fine, XFACE + d, 1, Gc, crse, 0, 1, np.zeros((1, 6), np.int32),
                                      np.zeros((6, 1)))
    nshared = 0
    for b in range(fba.nbox):
        v = fba.fab_view(fine, b, XFACE + d, 1, 1)[0]
        lo = fba.boxes[b, :3] - 1
        for b2 in range(fba.nbox):
            if b2 == b:
                continue
            flo = fba.boxes[b2, :3].copy()
            fhi = fba.boxes[b2, 3:].copy()
            fhi[d] += 1
            # faces of b2 inside b's FAB storage but outside b's valid faces
            glo = np.maximum(flo, lo)
            ghi = np.minimum(fhi, lo + np.array(v.shape[::-1]) - 1)
            if np.any(glo > ghi):
                continue
            for k in range(glo[2], ghi[2] + 1):
                for j in range(glo[1], ghi[1] + 1):
                    for i in range(glo[0], ghi[0] + 1):
                        p = np.array([i, j, k])
                        vb = fba.boxes[b]
                        vhi = vb[3:].copy()
                        vhi[d] += 1
                        if np.all(p >= vb[:3]) and np.all(p <= vhi):
                            continue
                        assert v[k - lo[2], j - lo[1], i - lo[0]] == dom[0, k, j, i]
                        nshared += 1
    assert nshared > 0
