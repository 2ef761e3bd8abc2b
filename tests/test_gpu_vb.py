"""MAC projection with face coefficients on the GPU (ksgf_macproj_set_beta_faces:
initProjector(LPInfo, Array<const MultiFab*,3> beta) / updateBeta(faces),
icns_advection.cpp:143-154, 234-291) against the oracle's restatement
(tests/test_vb_oracle.py pins it with a dense solve).

The sweeps evaluate the oracle's expressions in its order; the only
difference is the order of the rhs-mean sum of singular problems, so the
V-cycle counts are equal and the corrected faces agree to 1e-12 relative."""
import importlib

import numpy as np
import pytest

from helpers import rel_l2, to_np

import test_vb_oracle as TV

pytestmark = pytest.mark.gpu

lay = importlib.import_module("kynema-sgf_amd.layout")
P_, N_, D_ = 0, 1, 2
XFACE = 1


def _setup(n, prob_hi, periodic, lobc, hibc, seed, sizes):
    rng = np.random.default_rng(seed)
    geom = lay.Geometry(n, prob_hi=prob_hi, periodic=periodic)
    ba = lay.BoxArray(geom, lay.chop_random(n, sizes, rng))
    b3 = TV.periodic_faces(TV.faces(n, rng), lobc)
    mac = []
    for d in range(3):
        shp = [n[2], n[1], n[0]]
        shp[2 - d] += 1
        a = rng.standard_normal(shp)
        ax = 2 - d
        if periodic[d]:
            s0, s1 = [slice(None)] * 3, [slice(None)] * 3
            s0[ax], s1[ax] = 0, -1
            a[tuple(s1)] = a[tuple(s0)]
        else:  # no flow through the walls
            s0, s1 = [slice(None)] * 3, [slice(None)] * 3
            s0[ax], s1[ax] = 0, -1
            a[tuple(s0)] = 0.0
            a[tuple(s1)] = 0.0
        mac.append(a)
    return geom, ba, b3, mac


CASES = [
    ("periodic", (32, 24, 16), (1.0, 0.75, 0.5), (1, 1, 1), (P_, P_, P_), (P_, P_, P_), 0),
    ("abl_walls", (32, 32, 16), (640.0, 640.0, 64.0), (1, 1, 0), (P_, P_, N_), (P_, P_, N_), 1),
    ("dirichlet", (16, 24, 20), (1.0, 1.5, 1.25), (0, 1, 0), (N_, P_, N_), (N_, P_, D_), 0),
]


@pytest.mark.parametrize("name,n,prob_hi,periodic,lobc,hibc,semi", CASES)
def test_face_beta_projection_matches_oracle(ksgf, oracle, gpu, name, n, prob_hi, periodic, lobc, hibc, semi):
    print(f"face-beta projection {name}", flush=True)
    geom, ba, b3, mac = _setup(n, prob_hi, periodic, lobc, hibc, 7, [8, 16])
    opts_o = oracle.default_opts(semicoarsening=semi)
    ref, phi_r, info_r = oracle.mac_project_vb(ba, mac, lobc, hibc, b3, 1e-11, 0.0, opts_o)
    dba = ksgf.DeviceBoxArray(ba)
    o = ksgf.default_options(do_semicoarsening=semi)
    mp = ksgf.MacProjector(dba, lobc, hibc, 1.0, o)
    db = [dba.upload(ba.scatter(b3[d][None], XFACE + d, 0)) for d in range(3)]
    mp.update_beta_faces(db, 0)
    dm = [dba.upload(ba.scatter(mac[d][None], XFACE + d, 1)) for d in range(3)]
    info = mp.project(dm, 1e-11, 0.0)
    print(f"  V-cycles gpu {info['iters']} oracle {info_r['iters']}", flush=True)
    assert info["converged"] and info_r["converged"]
    assert info["iters"] == info_r["iters"]
    for d in range(3):
        got = ba.gather(to_np(dm[d]), XFACE + d, 1, 1)[0]
        assert rel_l2(got, ref[d]) <= 1e-12, (d, rel_l2(got, ref[d]))
    phi = to_np(mp.phi()).reshape(phi_r.shape)
    assert rel_l2(phi - phi.mean(), phi_r - phi_r.mean()) <= 1e-11
    mp.close()


def test_constant_face_beta_equals_scalar_beta(ksgf, gpu):
    """beta given as constant faces projects as updateBeta(Real) does (two
    different hierarchies, both converged to rtol 1e-12), and set_beta(Real)
    switches the projector back to the constant-coefficient path"""
    n, lobc, hibc = (32, 32, 16), (P_, P_, N_), (P_, P_, D_)
    geom, ba, _, mac = _setup(n, (1.0, 1.0, 0.5), (1, 1, 0), lobc, hibc, 3, [16])
    dba = ksgf.DeviceBoxArray(ba)
    mp = ksgf.MacProjector(dba, lobc, hibc, 0.8)
    out = {}
    for mode in ("scalar", "faces", "scalar-again"):
        dm = [dba.upload(ba.scatter(mac[d][None], XFACE + d, 1)) for d in range(3)]
        if mode == "faces":
            db = [dba.alloc(XFACE + d, 1, 0).fill_(0.8) for d in range(3)]
            mp.update_beta_faces(db, 0)
        elif mode == "scalar-again":
            mp.update_beta(0.8)
        info = mp.project(dm, 1e-12, 0.0)
        assert info["converged"], (mode, info)
        out[mode] = [ba.gather(to_np(m), XFACE + d, 1, 1)[0] for d, m in enumerate(dm)]
    for d in range(3):
        assert rel_l2(out["faces"][d], out["scalar"][d]) <= 1e-10
        assert np.array_equal(out["scalar-again"][d], out["scalar"][d])
    mp.close()


def test_average_cellcenter_to_face_bitexact(ksgf, oracle, gpu):
    n = (12, 10, 8)
    rng = np.random.default_rng(9)
    geom = lay.Geometry(n, prob_hi=(1.2, 1.0, 0.8), periodic=(1, 1, 0))
    ba = lay.BoxArray(geom, lay.chop_random(n, [4, 8], rng))
    rho = 1.0 + 0.5 * rng.uniform(size=(1, n[2], n[1], n[0]))
    flat = ba.scatter(rho, 0, 1)
    G = oracle.Geom(ba)
    bc = np.zeros((1, 6), dtype=np.int32)
    bc[0, 2] = bc[0, 5] = 2
    oracle.fill_ghosts(G, flat, 0, 1, 1, bc, np.zeros((6, 1)))
    dba = ksgf.DeviceBoxArray(ba)
    faces = dba.average_cellcenter_to_face(dba.upload(flat), 1)
    # the cell values of every face's two cells, ghosts included
    ext = np.pad(rho[0], ((1, 1), (0, 0), (0, 0)), mode="edge")
    ext = np.pad(ext, ((0, 0), (1, 1), (1, 1)), mode="wrap")
    want = [0.5 * (ext[1:-1, 1:-1, :-1] + ext[1:-1, 1:-1, 1:]), 0.5 * (ext[1:-1, :-1, 1:-1] + ext[1:-1, 1:, 1:-1]),
            0.5 * (ext[:-1, 1:-1, 1:-1] + ext[1:, 1:-1, 1:-1])]
    for d in range(3):
        got = ba.gather(to_np(faces[d]), XFACE + d, 1, 0)[0]
        assert np.array_equal(got, want[d]), d


def test_host_layer_variable_density_projection(ksgf, oracle, gpu):
    """MacProjOp's variable-density branch on the C++ Hydro layer
    (average_cellcenter_to_face, invert(factor), initProjector(faces),
    project) against the oracle's face-coefficient projection with
    beta = factor / (0.5 (rho_lo + rho_hi)) formed in numpy"""
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    n, lobc, hibc = (32, 24, 16), (P_, P_, N_), (P_, P_, N_)
    geom, ba, _, mac = _setup(n, (640.0, 480.0, 64.0), (1, 1, 0), lobc, hibc, 11, [8, 16])
    rng = np.random.default_rng(12)
    rho = 1.0 + 0.5 * rng.uniform(size=(n[2], n[1], n[0]))
    factor = 0.75
    ext = np.pad(rho, ((1, 1), (0, 0), (0, 0)), mode="edge")
    ext = np.pad(ext, ((0, 0), (1, 1), (1, 1)), mode="wrap")
    rf = [0.5 * (ext[1:-1, 1:-1, :-1] + ext[1:-1, 1:-1, 1:]), 0.5 * (ext[1:-1, :-1, 1:-1] + ext[1:-1, 1:, 1:-1]),
          0.5 * (ext[:-1, 1:-1, 1:-1] + ext[1:, 1:-1, 1:-1])]
    b3 = [factor / r for r in rf]
    o = ksgf.default_options(do_semicoarsening=1)
    got, phi, info = hs.mac_project_density(ba, lobc, hibc, rho, mac, factor, 1e-11, 0.0, o)
    ref, phi_r, info_r = oracle.mac_project_vb(ba, mac, lobc, hibc, b3, 1e-11, 0.0,
                                               oracle.default_opts(semicoarsening=1))
    print(f"host-layer variable-density projection: V-cycles {info['iters']} (oracle {info_r['iters']})", flush=True)
    assert info["converged"] and info["iters"] == info_r["iters"]
    for d in range(3):
        assert rel_l2(got[d], ref[d]) <= 1e-12, d
