"""PPM::PredictStateOn{X,Y,Z}Face through the C ABI
(ksgf_ppm_predict_state_on_face) on the GPU.

The reference's own known answers (unit_tests/multiphase/test_mflux_schemes.cpp:195-410:
upwind and minmod on constant, |i - c| and quadratic fields, minmod at hoextrap
boundaries, tolerance eps * 1e4) are evaluated through the HIP entry point,
and every limiter the entry point provides agrees bit for bit with the
oracle's restatement on seeded fields with walls."""
import numpy as np
import pytest

import test_oracle as TO
from helpers import to_np

pytestmark = pytest.mark.gpu

lay = TO.lay
EPS = TO.EPS
ADV, DT = TO.ADV, TO.DT


def _gpu_predict(ksgf, ba, d, s, vel_edge, bc, lim):
    dba = ksgf.DeviceBoxArray(ba)
    Im, Ip = dba.ppm_predict_state(d, dba.upload(s), 1, 2, 0, dba.upload(vel_edge), 1, bc, DT, lim)
    return to_np(Im), to_np(Ip)


def test_ppm_upwind_known_answers_gpu(ksgf, oracle, gpu):
    print("ppm upwind known answers", flush=True)
    ba, G, bc = TO._mflux_setup(oracle)
    c = 5 // 2
    s = TO._fill(ba, lambda i, j, k: 1.5 + 0 * i)
    Im, Ip = _gpu_predict(ksgf, ba, 0, s, TO._mac(ba, 0, ADV), bc, lay.PPM_UPWIND)
    assert abs(TO._at(ba, Im, c, c, c) - 1.5) < EPS * 1e4 and abs(TO._at(ba, Ip, c, c, c) - 1.5) < EPS * 1e4
    for d, vel in ((0, ADV), (1, ADV), (2, -ADV)):
        s = TO._fill(ba, lambda i, j, k: np.abs((i, j, k)[d] - c).astype(float))
        Im, Ip = _gpu_predict(ksgf, ba, d, s, TO._mac(ba, d, vel), bc, lay.PPM_UPWIND)
        assert abs(TO._at(ba, Im, c, c, c)) < EPS * 1e4 and abs(TO._at(ba, Ip, c, c, c)) < EPS * 1e4


def test_ppm_minmod_known_answers_gpu(ksgf, oracle, gpu):
    print("ppm minmod known answers", flush=True)
    ba, G, bc = TO._mflux_setup(oracle)
    c = 5 // 2
    dx = 0.4
    s = TO._fill(ba, lambda i, j, k: 1.5 + 0 * i)
    Im, Ip = _gpu_predict(ksgf, ba, 0, s, TO._mac(ba, 0, ADV), bc, lay.PPM_MINMOD)
    assert abs(TO._at(ba, Im, c, c, c) - 1.5) < EPS * 1e4
    ir = float(c)
    slp = (ir ** 2 - (ir - 1.0) ** 2) / dx
    val_p = ir ** 2 - slp * 0.5 * (DT * ADV - dx)
    val_n = ir ** 2 + slp * 0.5 * (DT * ADV - dx)
    for d, vel in ((0, ADV), (1, ADV)):
        s = TO._fill(ba, lambda i, j, k: ((i, j, k)[d] ** 2).astype(float))
        Im, Ip = _gpu_predict(ksgf, ba, d, s, TO._mac(ba, d, vel), bc, lay.PPM_MINMOD)
        assert abs(TO._at(ba, Im, c, c, c) - c * c) < EPS * 1e4
        assert abs(TO._at(ba, Ip, c, c, c) - val_p) < EPS * 1e4
    s = TO._fill(ba, lambda i, j, k: (k ** 2).astype(float))
    Im, Ip = _gpu_predict(ksgf, ba, 2, s, TO._mac(ba, 2, -ADV), bc, lay.PPM_MINMOD)
    assert abs(TO._at(ba, Im, c, c, c) - val_n) < EPS * 1e4
    assert abs(TO._at(ba, Ip, c, c, c) - c * c) < EPS * 1e4
    s = TO._fill(ba, lambda i, j, k: np.abs(i - c).astype(float))
    Im, Ip = _gpu_predict(ksgf, ba, 0, s, TO._mac(ba, 0, ADV), bc, lay.PPM_MINMOD)
    assert abs(TO._at(ba, Im, c, c, c)) < EPS * 1e4 and abs(TO._at(ba, Ip, c, c, c)) < EPS * 1e4


def test_ppm_minmod_boundary_known_answers_gpu(ksgf, oracle, gpu):
    print("ppm minmod hoextrap boundary known answers", flush=True)
    bc = np.full((1, 6), lay.BC_FOEXTRAP, dtype=np.int32)
    bc[0, 0] = lay.BC_HOEXTRAP
    bc[0, 4] = lay.BC_HOEXTRAP
    ba, G, _ = TO._mflux_setup(oracle)
    dx = 0.4
    i, j, k = 0, 2, 2
    s = TO._fill(ba, lambda a, b, c: (a ** 2).astype(float))
    Im, Ip = _gpu_predict(ksgf, ba, 0, s, TO._mac(ba, 0, -ADV), bc, lay.PPM_MINMOD)
    ir = float(i)
    slp = (ir ** 2 - (ir - 1.0) ** 2) / dx
    val_n = ir ** 2 + slp * 0.5 * (DT * ADV - dx)
    assert abs(TO._at(ba, Im, i, j, k) - val_n) < EPS * 1e4
    assert abs(TO._at(ba, Ip, i, j, k) - ir * ir) < EPS * 1e4
    i, j, k = 2, 4, 2
    s = TO._fill(ba, lambda a, b, c: (b ** 2).astype(float))
    Im, Ip = _gpu_predict(ksgf, ba, 1, s, TO._mac(ba, 1, ADV), bc, lay.PPM_MINMOD)
    ir = float(j)
    slp = ((ir + 1.0) ** 2 - ir ** 2) / dx
    val_p = ir ** 2 - slp * 0.5 * (DT * ADV - dx)
    assert abs(TO._at(ba, Im, i, j, k) - ir * ir) < EPS * 1e4
    assert abs(TO._at(ba, Ip, i, j, k) - val_p) < EPS * 1e4


@pytest.mark.parametrize("lim", [lay.PPM_VANLEER, lay.PPM_NOLIMITER, lay.PPM_UPWIND, lay.PPM_MINMOD])
def test_ppm_predict_bitexact_vs_oracle(ksgf, oracle, gpu, lim):
    """every direction, two boxes, walls of each kind on the non-periodic
    sides (ext_dir low, hoextrap high in y; hoextrap low, foextrap high in
    z), a velocity of both signs: Im and Ip equal the oracle's bit for bit"""
    print(f"ppm predict bit-exact lim={lim}", flush=True)
    n = (12, 10, 8)
    geom = lay.Geometry(n, (0.0, 0.0, 0.0), (1.2, 1.0, 0.6), (1, 0, 0))
    ba = lay.BoxArray(geom, [[0, 0, 0, 5, 9, 7], [6, 0, 0, 11, 9, 7]])
    G = oracle.Geom(ba)
    bc = np.zeros((1, 6), dtype=np.int32)
    bc[0, 1], bc[0, 4] = lay.BC_EXT_DIR, lay.BC_HOEXTRAP
    bc[0, 2], bc[0, 5] = lay.BC_HOEXTRAP, lay.BC_FOEXTRAP
    rng = np.random.default_rng(31)
    s = rng.standard_normal(ba.mf_size(TO.CELL, 2, 2))
    for d in range(3):
        vel = rng.uniform(-1.0, 1.0, ba.mf_size(TO.XFACE + d, 1, 1))
        Im_r, Ip_r = oracle.ppm_predict_state(G, d, s, 2, 2, 1, vel, 1, bc, DT, lim)
        dba = ksgf.DeviceBoxArray(ba)
        Im, Ip = dba.ppm_predict_state(d, dba.upload(s), 2, 2, 1, dba.upload(vel), 1, bc, DT, lim)
        assert np.array_equal(to_np(Im), Im_r), f"dir {d} Im"
        assert np.array_equal(to_np(Ip), Ip_r), f"dir {d} Ip"
