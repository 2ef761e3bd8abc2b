"""Device-controlled MAC projection (ksgf_macproj_project_begin / finish /
correct): the convergence test of every MLMG iteration on the device, so a
step can be captured into a hipGraph and read back once.

* begin(K) + finish + correct equal ksgf_macproj_project bit for bit (faces,
  phi, iteration count), FP64 and mixed, for K below the iteration count
  (finish continues on the host), equal to it, and above it (the extra
  iterations return at once).
* begin captured into a graph (torch.cuda.CUDAGraph on the capture stream)
  and replayed on three different inputs, with eager work and
  synchronisations between the replays: every replay equals the eager
  projection of its input bit for bit.
"""
import importlib

import numpy as np
import pytest

from helpers import to_np

pytestmark = pytest.mark.gpu

CELL, XFACE = 0, 1


def _say(msg):
    print(msg, flush=True)


def _faces(ksgf, oracle, workloads, seed, n=(64, 48, 32)):
    case = workloads.abl(n, extent=(20.0 * n[0], 20.0 * n[1], 4.0 * n[2]), sizes=(16, 32), nsteps=1, seed=seed)
    G, vel, rho, gp = oracle.make_state(case)
    dt = case.cfl / oracle.cfl_max(G, vel, 3)
    mac = oracle.extrap_vel_to_faces(G, vel, np.zeros(case.ba.mf_size(CELL, 3, 1)), case.bcrec, dt)
    return case, mac


@pytest.mark.parametrize("prec", [0, 1])
def test_split_projection_equals_project(ksgf, oracle, workloads, gpu, prec):
    case, mac = _faces(ksgf, oracle, workloads, 3)
    dba = ksgf.DeviceBoxArray(case.ba)
    opts = ksgf.default_options(**case.mlmg)
    opts.precision = prec
    mp = ksgf.MacProjector(dba, case.lobc, case.hibc, 1.0, opts)
    ref = [dba.upload(m.copy()) for m in mac]
    info = mp.project(ref, 1e-11, 1e-14)
    phi_ref = to_np(mp.phi()).copy()
    n_it = info["iters"]
    _say(f"split projection precision={prec}: eager {n_it} iterations")
    for K in (2, n_it, n_it + 4):
        g = [dba.upload(m.copy()) for m in mac]
        mp.project_begin(g, 1e-11, 1e-14, K)
        inf = mp.project_finish()
        mp.project_correct(g)
        _say(f"  K={K}: {inf['iters']} iterations, rfinal {inf['rfinal']:.3e} (eager {info['rfinal']:.3e})")
        assert inf["iters"] == n_it and inf["converged"] == 1
        assert inf["r0"] == info["r0"] and inf["rfinal"] == info["rfinal"]
        assert inf["bottom_iters"] == info["bottom_iters"]
        for d in range(3):
            assert np.array_equal(to_np(g[d]), to_np(ref[d])), (K, d)
        assert np.array_equal(to_np(mp.phi()), phi_ref), K
    mp.close()


def test_graph_replay_equals_eager(ksgf, oracle, workloads, gpu):
    import torch

    inputs = [_faces(ksgf, oracle, workloads, s)[1] for s in (1, 2, 5)]
    case, _ = _faces(ksgf, oracle, workloads, 1)
    dba = ksgf.DeviceBoxArray(case.ba)
    opts = ksgf.default_options(**case.mlmg)
    opts.precision = 1
    mp = ksgf.MacProjector(dba, case.lobc, case.hibc, 1.0, opts)
    # eager references (this also builds every launch table before the capture)
    refs = []
    for mac in inputs:
        g = [dba.upload(m.copy()) for m in mac]
        info = mp.project(g, 1e-11, 1e-14)
        refs.append(([to_np(a).copy() for a in g], to_np(mp.phi()).copy(), info["iters"]))
    K = max(r[2] for r in refs) + 2
    buf = [dba.upload(m.copy()) for m in inputs[0]]
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        mp.project_begin(buf, 1e-11, 1e-14, K)
    _say(f"graph captured with {K} iterations")
    for rep, mac in enumerate(inputs):
        for d in range(3):
            buf[d].copy_(torch.from_numpy(mac[d]).to(buf[d].device))
        torch.cuda.synchronize()
        graph.replay()
        inf = mp.project_finish()
        mp.project_correct(buf)
        torch.cuda.synchronize()
        _say(f"  replay {rep}: {inf['iters']} iterations (eager {refs[rep][2]})")
        assert inf["iters"] == refs[rep][2] and inf["converged"] == 1
        for d in range(3):
            assert np.array_equal(to_np(buf[d]), refs[rep][0][d]), (rep, d)
        assert np.array_equal(to_np(mp.phi()), refs[rep][1]), rep
    del graph
    mp.close()


@pytest.mark.parametrize("which", ["tgv64", "abl_scalar"])
def test_host_step_graph_equals_eager(ksgf, oracle, workloads, gpu, which):
    """GodunovPredictor's fixed-dt steps replayed from hipGraphs (one host
    read per step) equal the eager steps bit for bit: velocity, the scalar,
    the MAC solve's iteration count, over steps that capture both parities"""
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    if which == "tgv64":
        case = workloads.taylor_green(64, 32)
    else:
        case = workloads.abl((64, 48, 32), extent=(1280.0, 960.0, 128.0), sizes=(16, 32), nsteps=1)
    _say(f"host step graph {which}")
    n = case.geom.n_cell
    z, y, x = np.meshgrid(*[(np.arange(int(m)) + 0.5) / int(m) for m in n[::-1]], indexing="ij")
    blob = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) / (2 * 0.1 ** 2))[None]
    sbc = np.zeros((1, 6), dtype=np.int32)
    if which == "abl_scalar":
        sbc[0, 2] = sbc[0, 5] = 2  # foextrap at the walls
    steppers = []
    for graph in (False, True):
        H = hs.HostStepper(case)
        H.load()
        if which == "abl_scalar":
            H.enable_scalar(sbc)
            H.load_scalar(blob)
        if graph:
            H.enable_graph(True)
        steppers.append(H)
    E, Gs = steppers
    E.step()  # one adaptive step gives the fixed dt
    dt = E.last_dt
    Gs.step()
    assert np.array_equal(E.velocity(), Gs.velocity())
    for k in range(5):
        E.step(dt)
        Gs.step(dt)
        _say(f"  step {k}: MAC iterations eager {E.last_info['iters']} graph {Gs.last_info['iters']}")
        assert Gs.last_info["iters"] == E.last_info["iters"] and Gs.last_dt == dt
        assert np.array_equal(E.velocity(), Gs.velocity()), k
        if which == "abl_scalar":
            assert np.array_equal(E.scalar(), Gs.scalar()), k
    for H in steppers:
        H.close()


def test_host_graph_rejects_host_reads(ksgf, workloads, gpu):
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    case = workloads.taylor_green(16, 16)
    H = hs.HostStepper(case)
    H.load()
    H.enable_diffusion(0.01, 2, 1e-10, 1e-14)
    with pytest.raises(RuntimeError, match="enable_graph"):
        H.enable_graph(True)
    H.close()
