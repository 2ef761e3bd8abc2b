#!/usr/bin/env python3
"""Benchmark: cell-updates/s of one Godunov advect + MAC-project step on the
256^3 neutral-ABL box (uneven boxes from {32,48,64}) on one MI355X, with the
MLMG in FP64 and in mixed precision (FP32 V-cycles inside FP64 defect
correction); the headline value is the mixed path, the FP64 path is
reported beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n NX NY NZ]
                  [--precision both|mixed|fp64] [--no-cpu] [--workload abl|poisson-batch|nodal|levels|levels-step]

A step is one pass of the hot path over the whole grid, driven by the C++
host layer (GodunovPredictor): convective CFL, source terms,
ExtrapVelToFaces, MAC projection (MLMG to rtol 1e-11), face ghost fill,
momentum fluxes, convective term, predictor update -- the Godunov branch of
incflo::ApplyPredictor without diffusion and nodal projection.  Inputs are
resident in HBM before the timed region.  Prints one JSON line.
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, nargs=3, default=[256, 256, 256])
    p.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-n", type=int, nargs=3, default=[128, 128, 128])
    p.add_argument("--workload", choices=["abl", "poisson-batch", "nodal", "levels", "levels-step"], default="abl",
                   help="poisson-batch: 64 independent periodic Poisson problems of 32^3..128^3 (zero-mean "
                        "N(0,1) right-hand sides, seed 5) solved as one batch (hipGraph replay and eager) and one "
                        "after another; nodal: the full incflo step (advect + MAC projection + ApplyProjection's "
                        "nodal projection) on the same ABL box next to the step without it; each prints its own "
                        "JSON line")
    p.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                   help="summarise four rocprofv3 --pmc output directories (FETCH_SIZE and WRITE_SIZE of a "
                        "1-step mixed run with KSGF_SWEEP_PLANES fixed, and of --copy-probe) into HBM bytes per "
                        "launch of the level-0 half sweep; prints text for profiles/")
    p.add_argument("--copy-probe", action="store_true",
                   help="only run a plain 1 GiB -> 1 GiB elementwise copy (x*1.0) 4 times: the known-size "
                        "kernel that scales FETCH_SIZE / WRITE_SIZE counters to bytes")
    return p.parse_args()


def copy_probe():
    import torch

    n = 1 << 27
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(4):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 4}))


def run_gpu(case, precision, steps, warmup, profile):
    import torch

    ksgf = importlib.import_module("kynema-sgf_amd")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    opts = ksgf.default_options(**case.mlmg)
    opts.precision = 1 if precision == "mixed" else 0
    S = hs.HostStepper(case, options=opts)
    S.load()
    torch.cuda.synchronize()
    for _ in range(warmup):
        S.step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    infos = []
    if profile:
        ksgf.prof_reset()
    ev0.record(stream)
    for _ in range(steps):
        S.step()
        infos.append(dict(S.last_info))
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    prof = ksgf.prof_read() if profile else None
    # one more (untimed) step with the post-projection divergence check
    S.step(check_divergence=True)
    divmax = S.last_divmax
    vel = S.velocity()
    S.close()
    return ms, infos, prof, divmax, vel


def accuracy(case, nsteps):
    """The mixed path against the FP64 path in lockstep (the FP64 step's dt
    drives both), after every step: velocity rel-L2, pressure (the MAC
    projection's phi, means removed: the problem is singular) rel-L2 and
    max |div u_mac| after each projection."""
    ksgf = importlib.import_module("kynema-sgf_amd")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    S = {}
    for p in ("fp64", "mixed"):
        o = ksgf.default_options(**case.mlmg)
        o.precision = 1 if p == "mixed" else 0
        S[p] = hs.HostStepper(case, options=o)
        S[p].load()
    out = {"steps": nsteps, "velocity_rel_l2": [], "pressure_rel_l2": [], "max_div_fp64": [], "max_div_mixed": []}
    for _ in range(nsteps):
        dt = S["fp64"].step(check_divergence=True)
        S["mixed"].step(dt=dt, check_divergence=True)
        v64, v32 = S["fp64"].velocity(), S["mixed"].velocity()
        out["velocity_rel_l2"].append(float(np.linalg.norm(v32 - v64) / np.linalg.norm(v64)))
        p64, p32 = S["fp64"].phi(), S["mixed"].phi()
        p64 -= p64.mean()
        p32 -= p32.mean()
        out["pressure_rel_l2"].append(float(np.linalg.norm(p32 - p64) / np.linalg.norm(p64)))
        out["max_div_fp64"].append(S["fp64"].last_divmax)
        out["max_div_mixed"].append(S["mixed"].last_divmax)
    for x in S.values():
        x.close()
    return out


def cpu_baseline(n_cell):
    """The oracle step on a bounded sample (a smaller ABL box with the same
    cell sizes and physics), all host threads; cell-updates/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    ext = (20.0 * n_cell[0], 20.0 * n_cell[1], 4.0 * n_cell[2])
    case = wl.abl(tuple(n_cell), extent=ext, nsteps=1)
    G, vel, rho, gp = O.make_state(case)
    t0 = time.perf_counter()
    reps = 0
    clock = {}
    while True:
        vel, d = O.advect_project_step(case, G, vel, rho, gp, clock=clock)
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 10:
            break
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"value": case.ncell * reps / dt, "unit": "cell-updates/s", "cores": cores, "kind": "port",
           "sample": f"{reps} oracle steps of the {n_cell[0]}x{n_cell[1]}x{n_cell[2]} ABL box (same dx/dz and "
                     f"physics, {d['mlmg']['iters']} V-cycles in the last), {dt:.2f} s"}
    # the same oracle on one thread (the OpenMP runtime the oracle links)
    try:
        import ctypes

        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        G1, v1, r1, g1 = O.make_state(case)
        t1 = time.perf_counter()
        n1 = 0
        clock1 = {}
        while True:
            v1, _ = O.advect_project_step(case, G1, v1, r1, g1, clock=clock1)
            n1 += 1
            if time.perf_counter() - t1 > 5.0 or n1 >= 2:
                break
        dt1 = time.perf_counter() - t1
        gomp.omp_set_num_threads(cores)
        out["one_thread"] = {"value": case.ncell * n1 / dt1, "unit": "cell-updates/s",
                             "sample": f"{n1} oracle steps of the same box on 1 thread, {dt1:.2f} s"}
    except OSError:
        out["one_thread"] = None
    n = case.geom.n_cell
    return out, (case, reps, case.ba.gather(vel, 0, 3, 3), d["phi"].reshape(n[2], n[1], n[0]))


def vs_oracle(case, reps, vel_ref, phi_ref):
    """The GPU step (mixed precision, as the headline) against the FP64
    oracle after the same steps of the same case: velocity rel-L2, the
    rel-L2 of the last MAC projection's grad phi (cell differences along
    each axis) and the post-projection max |div u_mac| of the last step"""
    ksgf = importlib.import_module("kynema-sgf_amd")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    o = ksgf.default_options(**case.mlmg)
    o.precision = 1
    S = hs.HostStepper(case, options=o)
    S.load()
    for r in range(reps):
        S.step(check_divergence=(r == reps - 1))
    v = S.velocity()
    phi = S.phi()

    def grad(p):
        return np.concatenate([np.diff(p, axis=a).ravel() for a in (2, 1, 0)])

    gr = grad(phi_ref)
    out = {"workload": f"ABL {case.geom.n_cell[0]}x{case.geom.n_cell[1]}x{case.geom.n_cell[2]}", "steps": reps,
           "precision": "mixed", "velocity_rel_l2": float(np.linalg.norm(v - vel_ref) / np.linalg.norm(vel_ref)),
           "grad_phi_rel_l2": float(np.linalg.norm(grad(phi) - gr) / np.linalg.norm(gr)),
           "post_projection_max_div": S.last_divmax}
    S.close()
    return out


PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_pk_half_abl256_mixed.txt")


def _pmc_rows(d):
    import csv
    import glob
    import re

    out = []
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(path)):
            out.append((int(r["Dispatch_Id"]), re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", ""),
                        int(r["Grid_Size"]), float(r["Counter_Value"])))
    out.sort()
    return out


def pmc_summary(fdir, wdir, cfdir, cwdir, n=(256, 256, 256), kz=4):
    """HBM bytes per launch of the level-0 k_pk_half<float> launches of a
    one-step run: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) matched launch
    by launch, each scaled by what the counter reads for the 1 GiB copy."""
    import collections

    true_kib = (1 << 27) * 8 / 1024.0
    scale = []
    for d in (cfdir, cwdir):
        v = [x[3] for x in _pmc_rows(d) if "MulFunctor" in x[1]]
        scale.append(true_kib / (sum(v) / len(v)))
    # level-0 launch grid with kz planes per workgroup: (n0/2/64) x (n1/4) x (n2/kz) tiles of 256
    ntx = -(-(n[0] // 2) // 64)
    grid0 = ntx * (-(-n[1] // 4)) * (-(-n[2] // kz)) * 256
    F = [x for x in _pmc_rows(fdir) if "k_pk_half<float" in x[1] and x[2] == grid0]
    W = [x for x in _pmc_rows(wdir) if "k_pk_half<float" in x[1] and x[2] == grid0]
    if len(F) != len(W) or not F:
        raise SystemExit(f"unmatched level-0 launches: {len(F)} fetch vs {len(W)} write")
    per = collections.defaultdict(list)
    for f, w in zip(F, W):
        if f[1] != w[1]:
            raise SystemExit("launch order differs between the two counter runs")
        per[f[1]].append((f[3] * scale[0] * 1024.0, w[3] * scale[1] * 1024.0))
    lines = [f"copy probe scale: FETCH_SIZE x{scale[0]:.4f}, WRITE_SIZE x{scale[1]:.4f} "
             f"(1 GiB read + 1 GiB written by torch.mul(x, 1.0))",
             f"level-0 k_pk_half<float> launches (grid {grid0}, {len(F)} launches):"]
    allb = []
    for name, v in sorted(per.items()):
        fb = sum(x[0] for x in v) / len(v)
        wb = sum(x[1] for x in v) / len(v)
        allb += [x[0] + x[1] for x in v]
        lines.append(f"  {name:40s} n={len(v):4d} fetch {fb / 1e6:8.2f} MB  write {wb / 1e6:8.2f} MB")
    lines.append(f"traffic_bytes_per_launch: {sum(allb) / len(allb):.0f}")
    return "\n".join(lines)


def pmc_traffic():
    """HBM bytes per level-0 half-sweep launch from the committed summary, or None."""
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def poisson_batch(steps, warmup):
    """Time the batched solve of the 64-problem set against the same
    problems solved one at a time through ksgf_macproj_solve."""
    import torch

    ksgf = importlib.import_module("kynema-sgf_amd")
    lay = importlib.import_module("kynema-sgf_amd.layout")
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    probs, rhs = wl.poisson_problems(64, seed=5)
    cells = sum(int(np.prod(p["n"])) for p in probs)
    rt = [torch.from_numpy(r).to("cuda") for r in rhs]
    phis = [torch.zeros_like(r) for r in rt]
    pb = ksgf.PoissonBatch(probs)
    singles = []
    for p in probs:
        n = p["n"]
        geom = lay.Geometry(tuple(n), (0.0, 0.0, 0.0), tuple(p["dx"][d] * n[d] for d in range(3)), (1, 1, 1))
        ba = lay.BoxArray(geom, [0, 0, 0, n[0] - 1, n[1] - 1, n[2] - 1])
        singles.append(ksgf.MacProjector(ksgf.DeviceBoxArray(ba), p["lobc"], p["hibc"], p["beta"]))
    infos = {}

    def timed(name, fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            infos[name] = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    t_graph = timed("graph", lambda: pb.solve(rt, phis, 1e-11, 1e-14, use_graph=True))
    t_eager = timed("eager", lambda: pb.solve(rt, phis, 1e-11, 1e-14, use_graph=False))
    t_seq = timed("single", lambda: [mp.solve(rt[q], phis[q], 1e-11, 1e-14) for q, mp in enumerate(singles)])
    its = [i["iters"] for i in infos["graph"]]
    assert its == [i["iters"] for i in infos["single"]]
    return {
        "metric": "Poisson cells solved/s: 64 independent periodic problems 32^3..128^3, MLMG FP64 to rtol 1e-11",
        "value": cells / t_graph,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": t_graph * 1e3,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"64 periodic Poisson problems, edges {sorted(set(p['n'][0] for p in probs))}, "
                               f"{cells} cells, rhs N(0,1) minus mean (seed 5)",
                   "mlmg_iters_max": max(its), "mlmg_iters_mean": sum(its) / len(its)},
        "ms_batch_graph": t_graph * 1e3,
        "ms_batch_eager": t_eager * 1e3,
        "ms_one_at_a_time": t_seq * 1e3,
        "speedup_graph_vs_one_at_a_time": t_seq / t_graph,
    }


def nodal_step(n, steps, warmup, no_cpu, cpu_n):
    """The ABL step followed by incflo::ApplyProjection (nodal projection,
    nodal_proj rtol/atol defaults, semicoarsening as the MAC solve), timed
    beside the same step without it; the oracle's nodal projection on a
    bounded sample is the CPU baseline."""
    import torch

    ksgf = importlib.import_module("kynema-sgf_amd")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    ext = (20.0 * n[0], 20.0 * n[1], 4.0 * n[2])
    case = wl.abl(tuple(n), extent=ext, nsteps=steps)
    eps = np.finfo(float).eps
    res = {}
    for nodal in (False, True):
        opts = ksgf.default_options(**case.mlmg)
        opts.precision = 1
        S = hs.HostStepper(case, options=opts)
        S.load()
        if nodal:
            S.enable_nodal_projection(eps * 1e5, eps * 1e2, ksgf.default_options(**case.mlmg))
        for _ in range(warmup):
            S.step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its = []
        for _ in range(steps):
            S.step()
            its.append(S.nodal_info()["iters"] if nodal else S.last_info["iters"])
        torch.cuda.synchronize()
        res[nodal] = ((time.perf_counter() - t) / steps, its)
        S.close()
    ms_full, nits = res[True][0] * 1e3, res[True][1]
    ms_pred = res[False][0] * 1e3
    out = {
        "metric": "cell-updates/s per full incflo step (advect + MAC projection + nodal projection)",
        "value": case.ncell / (ms_full * 1e-3),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_full,
        "higher_is_better": True,
        "dtype": "f64+f32mg",
        "data": "synthetic",
        "config": {"workload": f"neutral ABL {n[0]}x{n[1]}x{n[2]}, dx=dy=20 m dz=4 m, uneven boxes {{32,48,64}}, "
                               f"CFL 0.9, MAC mixed precision rtol 1e-11, nodal FP64 rtol eps*1e5",
                   "nodal_vcycles": nits},
        "ms_step_without_nodal": ms_pred,
        "ms_nodal_projection": ms_full - ms_pred,
    }
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.build()
        ext = (20.0 * cpu_n[0], 20.0 * cpu_n[1], 4.0 * cpu_n[2])
        c2 = wl.abl(tuple(cpu_n), extent=ext, nsteps=1)
        G, vel, rho, gp = O.make_state(c2)
        t0 = time.perf_counter()
        reps = 0
        while True:
            v = vel.copy()
            g = gp.copy()
            info, _ = O.apply_projection(c2, G, v, rho, g, 1.0, eps * 1e5, eps * 1e2)
            reps += 1
            if time.perf_counter() - t0 > 10.0 or reps >= 5:
                break
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": c2.ncell * reps / dt, "unit": "cell-updates/s (nodal projection only)",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
                               "sample": f"{reps} oracle nodal projections of the {cpu_n[0]}x{cpu_n[1]}x{cpu_n[2]} "
                                         f"ABL box ({info['iters']} V-cycles), {dt:.2f} s"}
    return out


def _faces_on_device(ba, d, ng, fn):
    """face MultiFab (1 comp, ngrow ng) of direction d on the device, each FAB
    (ghost faces included) filled with fn(X, Y, Z) at its face centres"""
    import torch

    out = torch.empty(ba.mf_size(1 + d, 1, ng), dtype=torch.float64, device="cuda")
    lo, n = ba.fab_lo_n(1 + d, ng)
    off = ba.fab_offsets(1 + d, 1, ng)
    h = ba.geom.dx
    for b in range(ba.nbox):
        ax = [(int(lo[b, e]) + torch.arange(int(n[b, e]), dtype=torch.float64, device="cuda") +
               (0.0 if e == d else 0.5)) * h[e] for e in range(3)]
        Z, Y, X = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        out[int(off[b]):int(off[b + 1])] = fn(X, Y, Z).reshape(-1)
    return out


def levels_project(steps, warmup, no_cpu):
    """Composite two-level MAC projection (Hydro::MacProjector over levels 0
    and 1) on the 1,500-box ragged 320x320x480 level 0 with a level 1 at
    ratio 2 over 40% of the domain; the oracle's composite projection on a
    64x64x96 copy of the layout is the CPU baseline."""
    import torch

    ksgf = importlib.import_module("kynema-sgf_amd")
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    case, fba = wl.ragged_two_level()
    cba = case.ba
    dc = ksgf.DeviceBoxArray(cba)
    df = ksgf.DeviceBoxArray(fba, covers_domain=False)
    u0c = [_faces_on_device(cba, d, 1, lambda X, Y, Z, d=d: wl.two_level_faces(cba.geom, torch, X, Y, Z, d))
           for d in range(3)]
    u0f = [_faces_on_device(fba, d, 1, lambda X, Y, Z, d=d: wl.two_level_faces(fba.geom, torch, X, Y, Z, d))
           for d in range(3)]
    uc = [a.clone() for a in u0c]
    uf = [a.clone() for a in u0f]
    rtol, atol = 1e-11, 0.0
    mp = ksgf.MacProjectorLevels(dc, df, case.lobc, case.hibc, 1.0)
    st = torch.cuda.current_stream()
    ms, its = [], []
    for s in range(warmup + steps):
        for a, b in zip(uc + uf, u0c + u0f):
            a.copy_(b)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        info = mp.project(uc, uf, 1, rtol, atol)
        e1.record(st)
        torch.cuda.synchronize()
        if s >= warmup:
            ms.append(e0.elapsed_time(e1))
            its.append(info["iters"])
    nfine = int(np.sum(np.prod(fba.boxes[:, 3:] - fba.boxes[:, :3] + 1, axis=1)))
    ncomp = case.ncell - nfine // 8 + nfine
    t = float(np.mean(ms))
    out = {
        "metric": "composite cell-updates/s of the two-level MAC projection",
        "value": ncomp / (t * 1e-3),
        "unit": "cell-updates/s (uncovered level-0 cells + level-1 cells)",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": t,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ragged {cba.nbox} boxes 320x320x480 level 0 + {fba.nbox} level-1 boxes (ratio 2, "
                               f"{nfine} cells over 40% of the domain), triply periodic, rtol 1e-11",
                   "composite_iterations": its, "r0": info["r0"], "rfinal": info["rfinal"]},
    }
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.build()
        c2, f2 = wl.ragged_two_level((64, 64, 96), 32, fine_z=(20, 58), fine_box=32)
        umc, umf = [], []
        for g, lst in ((c2.geom, umc), (f2.geom, umf)):
            for d in range(3):
                ax = [(np.arange(g.n_cell[e] + (e == d)) + (0.0 if e == d else 0.5)) * g.dx[e] for e in range(3)]
                Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
                lst.append(np.ascontiguousarray(wl.two_level_faces(g, np, X, Y, Z, d)).ravel())
        n2 = int(np.sum(np.prod(f2.boxes[:, 3:] - f2.boxes[:, :3] + 1, axis=1)))
        t0 = time.perf_counter()
        rinfo, _, _ = O.mac_project_two_levels(c2.geom.n_cell, c2.geom.dx, c2.lobc, c2.hibc, f2.boxes, 1.0, umc, umf,
                                               rtol, atol)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": (c2.ncell - n2 // 8 + n2) / dt, "unit": out["unit"],
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
                               "sample": f"one oracle composite projection of the 64x64x96 two-level copy "
                                         f"({rinfo['iters']} iterations), {dt:.2f} s"}
    return out


def _cells_on_device(ba, ncomp, ng, fn):
    return importlib.import_module("kynema-sgf_amd.workloads").cells_on_device(ba, ncomp, ng, fn)


def _tgv_cells(geom):
    return importlib.import_module("kynema-sgf_amd.workloads").tgv_cells(geom)


def levels_step(steps, warmup, no_cpu):
    """Two-level advect + MAC-project steps (levels.TwoLevelStep: per-level
    Godunov, composite MAC projection, fill-patches, flux averaging,
    average_down) on the 1,500-box ragged 320x320x480 level 0 with 600
    level-1 boxes over 40% of the domain; the oracle's two-level step on the
    16^3 Taylor-Green copy with three level-1 boxes is the CPU baseline."""
    import torch

    wl = importlib.import_module("kynema-sgf_amd.workloads")
    lv = importlib.import_module("kynema-sgf_amd.levels")
    case, fba = wl.ragged_two_level()
    S = lv.TwoLevelStep(case, fba)
    vc = _cells_on_device(case.ba, 3, 3, _tgv_cells(case.geom))
    vf = _cells_on_device(fba, 3, 3, _tgv_cells(fba.geom))
    S.df.average_down_to(vf, 3, S.dc, vc, 3, 3)
    S.load(vc, vf)
    del vc, vf
    st = torch.cuda.current_stream()
    ms, its = [], []
    for s in range(warmup + steps):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        S.step()
        e1.record(st)
        torch.cuda.synchronize()
        if s >= warmup:
            ms.append(e0.elapsed_time(e1))
            its.append(S.last_info["iters"])
    nfine = int(np.sum(np.prod(fba.boxes[:, 3:] - fba.boxes[:, :3] + 1, axis=1)))
    ncomp = case.ncell - nfine // 8 + nfine
    t = float(np.mean(ms))
    out = {
        "metric": "composite cell-updates/s of the two-level advect + MAC-project step",
        "value": ncomp / (t * 1e-3),
        "unit": "cell-updates/s (uncovered level-0 cells + level-1 cells)",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": t,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ragged {case.ba.nbox} boxes 320x320x480 level 0 + {fba.nbox} level-1 boxes (ratio 2, "
                               f"{nfine} cells over 40% of the domain), triply periodic Taylor-Green, CFL 0.5, "
                               f"MAC rtol 1e-11", "composite_iterations": its, "last_dt": S.last_dt},
    }
    S.close()
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.build()
        c2, f2 = wl.ragged_two_level((64, 64, 96), 32, fine_z=(20, 58), fine_box=32)
        vel = []
        for ba in (c2.ba, f2):
            g = ba.geom
            k = [2.0 * np.pi / (g.prob_hi[e] - g.prob_lo[e]) for e in range(3)]
            x = g.cell_centers(0)[None, None, :]
            y = g.cell_centers(1)[None, :, None]
            z = g.cell_centers(2)[:, None, None]
            u = np.sin(k[0] * x) * np.cos(k[1] * y) * np.cos(k[2] * z)
            v = -np.cos(k[0] * x) * np.sin(k[1] * y) * np.cos(k[2] * z)
            vel.append(ba.scatter(np.stack(np.broadcast_arrays(u, v, 0.1 * u * v)), 0, 3))
        O.average_down(O.Geom(f2), vel[1], 3, O.Geom(c2.ba), vel[0], 3, 3, 2)
        gp_c = np.zeros(c2.ba.mf_size(0, 3, 0))
        gp_f = np.zeros(f2.mf_size(0, 3, 0))
        n2 = int(np.sum(np.prod(f2.boxes[:, 3:] - f2.boxes[:, :3] + 1, axis=1)))
        t0 = time.perf_counter()
        reps = 0
        vel_c, vel_f = vel
        clock = {}
        while True:
            vel_c, vel_f, rinfo = O.two_level_step(c2, f2, vel_c, vel_f, gp_c, gp_f, clock=clock)
            reps += 1
            if time.perf_counter() - t0 > 10.0 or reps >= 5:
                break
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": (c2.ncell - n2 // 8 + n2) * reps / dt, "unit": out["unit"],
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
                               "sample": f"{reps} oracle two-level steps of the 64x64x96 copy of the layout "
                                         f"({rinfo['iters']} composite iterations), {dt:.2f} s"}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; runs on more GPUs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.copy_probe:
        copy_probe()
        return
    if a.pmc_summary:
        print(pmc_summary(*a.pmc_summary))
        return
    if a.workload == "poisson-batch":
        print(json.dumps(poisson_batch(a.steps, a.warmup)))
        return
    if a.workload == "nodal":
        print(json.dumps(nodal_step(tuple(a.n), a.steps, a.warmup, a.no_cpu, a.cpu_n)))
        return
    if a.workload == "levels":
        print(json.dumps(levels_project(a.steps, a.warmup, a.no_cpu)))
        return
    if a.workload == "levels-step":
        print(json.dumps(levels_step(a.steps, a.warmup, a.no_cpu)))
        return
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    n = tuple(a.n)
    ext = (20.0 * n[0], 20.0 * n[1], 4.0 * n[2])
    case = wl.abl(n, extent=ext, nsteps=a.steps)
    res = {}
    precs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    for p in precs:
        res[p] = run_gpu(case, p, a.steps, a.warmup, profile=(p == precs[-1]))
    head = precs[-1]
    ms, infos, prof, divmax, vel = res[head]
    value = case.ncell / (ms * 1e-3)
    out = {
        "metric": "cell-updates/s per advect+project step at 256^3 (FP64 vs mixed), %HBM roofline",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32mg" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"neutral ABL {n[0]}x{n[1]}x{n[2]}, dx=dy=20 m dz=4 m, uneven boxes "
                               f"{{32,48,64}} ({case.ba.nbox} boxes), CFL 0.9, PPM, MAC mg_rtol 1e-11",
                   "precision": head, "mlmg_iters": [i["iters"] for i in infos],
                   "mlmg_r0": [i["r0"] for i in infos], "mlmg_rfinal": [i["rfinal"] for i in infos],
                   "post_projection_max_div": divmax},
    }
    if "fp64" in res and head != "fp64":
        ms64, infos64, _, div64, vel64 = res["fp64"]
        out["fp64"] = {"value": case.ncell / (ms64 * 1e-3), "ms_per_step": ms64,
                       "mlmg_iters": [i["iters"] for i in infos64], "post_projection_max_div": div64}
        out["speedup_mixed_vs_fp64"] = ms64 / ms
        out["mixed_vs_fp64"] = accuracy(case, 10)
    if prof and prof.get("launches"):
        avg_s = prof["ms"] * 1e-3 / prof["launches"]
        bpl = prof["bytes"] / prof["launches"]
        ach = bpl / avg_s / 1e9
        traffic = pmc_traffic() if (n == (256, 256, 256) and head == "mixed") else None
        out["roofline"] = {"bound": "hbm", "kernel": prof["name"], "achieved": ach, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                           "launches": prof["launches"], "avg_us": avg_s * 1e6,
                           "bytes_per_launch": bpl}
    if not a.no_cpu:
        out["cpu_baseline"], (c2, reps, vref, pref) = cpu_baseline(a.cpu_n)
        out["vs_oracle"] = vs_oracle(c2, reps, vref, pref)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
