"""Composite two-level MAC projection on the GPU (ksgf_macproj_levels_*)
against the CPU oracle's ko_mac_project_two_levels: a ragged coarse level,
fine boxes at ratio 2 that reach the periodic sides and the y walls, random
face velocities.  Both run the same iteration (fine Jacobi, one coarse
V-cycle per composite iteration); the coarse V-cycles agree to rounding, so
the iteration counts match and the corrected velocities agree to 1e-9 of the
input scale."""
import importlib

import numpy as np
import pytest

from helpers import to_np

pytestmark = pytest.mark.gpu

lay = importlib.import_module("kynema-sgf_amd.layout")
P_, N_, D_ = 0, 1, 2


def _say(msg):
    print(msg, flush=True)


def _random_faces(rng, m, walls_y):
    out = []
    for d in range(3):
        shp = [m[2], m[1], m[0]]
        shp[2 - d] += 1
        a = rng.standard_normal(shp)
        if d == 1 and walls_y:
            a[:, 0, :] = 0.0
            a[:, -1, :] = 0.0
        if d != 1:  # periodic image faces agree
            sl = [slice(None)] * 3
            sl[2 - d] = -1
            sl0 = list(sl)
            sl0[2 - d] = 0
            a[tuple(sl)] = a[tuple(sl0)]
        out.append(np.ascontiguousarray(a))
    return out


CASES = [
    ("neumann_y", (P_, N_, P_), (P_, N_, P_), 1.3),
    ("dirichlet_ylo", (P_, D_, P_), (P_, N_, P_), 0.7),
    ("dirichlet_y", (P_, D_, P_), (P_, D_, P_), 1.0),  # a level-1 box on the high Dirichlet wall
]


@pytest.mark.parametrize("name,lobc,hibc,beta", CASES)
def test_macproj_levels_matches_oracle(ksgf, oracle, gpu, name, lobc, hibc, beta):
    import test_multilevel_oracle as T

    _say(f"macproj_levels {name}")
    rng = np.random.default_rng(17)
    n = (16, 12, 16)
    cg = lay.Geometry(n, prob_hi=(1.0, 0.75, 1.0), periodic=(1, 0, 1))
    cba = lay.BoxArray(cg, lay.chop_random(n, [4, 8], rng, granularity=4))
    nf = tuple(2 * v for v in n)
    fg = lay.Geometry(nf, prob_hi=cg.prob_hi, periodic=(1, 0, 1))
    fboxes = [[0, 4, 8, 7, 11, 23], [24, 8, 8, 31, 19, 15], [8, 12, 24, 23, 23, 31], [8, 4, 8, 15, 11, 15]]
    fba = T.FineBoxArray(fg, fboxes)
    ucd = _random_faces(rng, n, True)
    ufd = _random_faces(rng, nf, True)
    # oracle on domain-wide arrays
    umc = [a.ravel().copy() for a in ucd]
    umf = [a.ravel().copy() for a in ufd]
    rinfo, rpc, rpf = oracle.mac_project_two_levels(n, cg.dx, lobc, hibc, fboxes, beta, umc, umf, 1e-11, 0.0)
    assert rinfo["status"] == 0, rinfo
    # GPU on MultiFabs (ngrow 1)
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    mc = [dc.upload(cba.scatter(ucd[d][None], 1 + d, 1)) for d in range(3)]
    mf = [df.upload(fba.scatter(ufd[d][None], 1 + d, 1)) for d in range(3)]
    mp = ksgf.MacProjectorLevels(dc, df, lobc, hibc, beta)
    info = mp.project(mc, mf, 1, 1e-11, 0.0)
    _say(f"  oracle iters {rinfo['iters']} r0 {rinfo['r0']:.3e}; gpu iters {info['iters']} r0 {info['r0']:.3e} "
         f"rfinal {info['rfinal']:.3e}")
    assert info["converged"] == 1
    assert abs(info["r0"] - rinfo["r0"]) <= 1e-13 * rinfo["r0"]
    assert abs(info["iters"] - rinfo["iters"]) <= 1
    scale = max(np.abs(a).max() for a in ucd + ufd)
    fm = np.zeros(nf[::-1], dtype=bool)
    for b in fboxes:
        fm[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1] = True
    for d in range(3):
        shp_c = list(n[::-1])
        shp_c[2 - d] += 1
        gc = cba.gather(to_np(mc[d]), 1 + d, 1, 1)[0]
        assert np.abs(gc - umc[d].reshape(shp_c)).max() <= 1e-9 * scale, f"coarse dir {d}"
        shp_f = list(nf[::-1])
        shp_f[2 - d] += 1
        gf = fba.gather(to_np(mf[d]), 1 + d, 1, 1)[0]
        ref = umf[d].reshape(shp_f)
        for lohi in (slice(0, -1), slice(1, None)):  # the low and high faces of every fine cell
            sl = [slice(None)] * 3
            sl[2 - d] = lohi
            assert np.abs(gf[tuple(sl)][fm] - ref[tuple(sl)][fm]).max() <= 1e-9 * scale, f"fine dir {d}"
    pc, pf = mp.phi()
    pc = to_np(pc)
    gpf = fba.gather(to_np(pf), 0, 1, 1)[0]
    sp = np.abs(rpc).max()
    assert np.abs(pc - rpc).max() <= 1e-9 * sp
    assert np.abs(gpf[fm] - rpf.reshape(nf[::-1])[fm]).max() <= 1e-9 * sp


def test_macproj_levels_rejects_bad_levels(ksgf, gpu):
    """boxes not aligned to the ratio, and a fine domain that is not the coarse
    one refined by 2, are errors at creation (before any solve)"""
    import test_multilevel_oracle as T

    _say("macproj_levels bad levels")
    n = (16, 12, 16)
    cg = lay.Geometry(n, prob_hi=(1.0, 0.75, 1.0), periodic=(1, 0, 1))
    cba = lay.BoxArray(cg, [[0, 0, 0, 15, 11, 15]])
    dc = ksgf.DeviceBoxArray(cba)
    fg = lay.Geometry((32, 24, 32), prob_hi=cg.prob_hi, periodic=(1, 0, 1))
    bad_align = ksgf.DeviceBoxArray(T.FineBoxArray(fg, [[1, 4, 8, 8, 11, 15]]), covers_domain=False)
    with pytest.raises(ksgf.KsgfError, match="aligned"):
        ksgf.MacProjectorLevels(dc, bad_align, (P_, N_, P_), (P_, N_, P_))
    fg3 = lay.Geometry((48, 36, 48), prob_hi=cg.prob_hi, periodic=(1, 0, 1))
    bad_ratio = ksgf.DeviceBoxArray(T.FineBoxArray(fg3, [[0, 0, 0, 11, 11, 11]]), covers_domain=False)
    with pytest.raises(ksgf.KsgfError, match="refined by 2"):
        ksgf.MacProjectorLevels(dc, bad_ratio, (P_, N_, P_), (P_, N_, P_))


def test_macproj_levels_maxiter_reports_not_converged(ksgf, oracle, gpu):
    """maxiter 2: the projection corrects with the iterate it has and reports
    KSGF_ERR_NOT_CONVERGED (a status, not an abort), as MLMG does at maxiter"""
    import test_multilevel_oracle as T

    _say("macproj_levels maxiter 2")
    rng = np.random.default_rng(5)
    n = (16, 12, 16)
    cg = lay.Geometry(n, prob_hi=(1.0, 0.75, 1.0), periodic=(1, 0, 1))
    cba = lay.BoxArray(cg, lay.chop_random(n, [4, 8], rng, granularity=4))
    nf = tuple(2 * v for v in n)
    fg = lay.Geometry(nf, prob_hi=cg.prob_hi, periodic=(1, 0, 1))
    fboxes = [[8, 4, 8, 23, 19, 23]]
    fba = T.FineBoxArray(fg, fboxes)
    ucd = _random_faces(rng, n, True)
    ufd = _random_faces(rng, nf, True)
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    mc = [dc.upload(cba.scatter(ucd[d][None], 1 + d, 1)) for d in range(3)]
    mf = [df.upload(fba.scatter(ufd[d][None], 1 + d, 1)) for d in range(3)]
    opts = ksgf.default_options(maxiter=2)
    mp = ksgf.MacProjectorLevels(dc, df, (P_, N_, P_), (P_, N_, P_), 1.0, opts)
    info = mp.project(mc, mf, 1, 1e-11, 0.0)
    assert info["iters"] == 2 and info["converged"] == 0
    assert 0.0 < info["rfinal"] < info["r0"]
    o = oracle.default_opts(maxiter=2)
    umc = [a.ravel().copy() for a in ucd]
    umf = [a.ravel().copy() for a in ufd]
    rinfo, _, _ = oracle.mac_project_two_levels(n, cg.dx, (P_, N_, P_), (P_, N_, P_), fboxes, 1.0, umc, umf, 1e-11,
                                                0.0, o)
    assert rinfo["status"] == 1 and rinfo["iters"] == 2
    assert abs(info["rfinal"] - rinfo["rfinal"]) <= 1e-8 * rinfo["r0"]


def test_macproj_levels_bitwise_reproducible(ksgf, gpu):
    """the composite mean that makes the singular (no Dirichlet side) solve
    unique is summed in a fixed block order: two projections of the same
    input give the same bits (and the same phi)"""
    import test_multilevel_oracle as T

    _say("macproj_levels reproducible")
    rng = np.random.default_rng(23)
    n = (16, 12, 16)
    cg = lay.Geometry(n, prob_hi=(1.0, 0.75, 1.0), periodic=(1, 0, 1))
    cba = lay.BoxArray(cg, lay.chop_random(n, [4, 8], rng, granularity=4))
    nf = tuple(2 * v for v in n)
    fg = lay.Geometry(nf, prob_hi=cg.prob_hi, periodic=(1, 0, 1))
    fba = T.FineBoxArray(fg, [[0, 4, 8, 7, 11, 23], [24, 8, 8, 31, 19, 15], [8, 12, 24, 23, 23, 31]])
    ucd = _random_faces(rng, n, True)
    ufd = _random_faces(rng, nf, True)
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    mp = ksgf.MacProjectorLevels(dc, df, (P_, N_, P_), (P_, N_, P_), 1.3)
    outs = []
    for _ in range(2):
        mc = [dc.upload(cba.scatter(ucd[d][None], 1 + d, 1)) for d in range(3)]
        mf = [df.upload(fba.scatter(ufd[d][None], 1 + d, 1)) for d in range(3)]
        info = mp.project(mc, mf, 1, 1e-11, 0.0)
        assert info["converged"] == 1
        pc, pf = mp.phi()
        outs.append([to_np(a).copy() for a in mc + mf] + [to_np(pc).copy(), to_np(pf).copy()])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
