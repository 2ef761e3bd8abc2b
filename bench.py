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
    p.add_argument("--no-theta", action="store_true", help="skip the run with the potential temperature advected")
    p.add_argument("--cpu-n", type=int, nargs=3, default=[128, 128, 128])
    p.add_argument("--workload", choices=["abl", "poisson-batch", "nodal", "levels", "levels-step", "fixed-dt"],
                   default="abl",
                   help="poisson-batch: 64 independent periodic Poisson problems of 32^3..128^3 (zero-mean "
                        "N(0,1) right-hand sides, seed 5) solved as one batch (hipGraph replay and eager) and one "
                        "after another; nodal: the full incflo step (advect + MAC projection + ApplyProjection's "
                        "nodal projection) on the same ABL box next to the step without it; each prints its own "
                        "JSON line")
    p.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                   help="summarise four rocprofv3 --pmc output directories (FETCH_SIZE and WRITE_SIZE of a "
                        "1-step mixed run with KSGF_SWEEP_PLANES fixed, and of --copy-probe) into HBM bytes per "
                        "launch of the level-0 half sweep; prints text for profiles/")
    p.add_argument("--trace-summary", nargs=2, metavar=("TRACE_DIR", "STEPS"),
                   help="summarise a rocprofv3 --kernel-trace output directory of STEPS advect + project steps "
                        "by (kernel, multigrid level): calls, average and total time per step; prints text for "
                        "profiles/")
    p.add_argument("--trace-grid-summary", nargs=2, metavar=("TRACE_DIR", "STEPS"),
                   help="summarise every launch of a rocprofv3 --kernel-trace output directory by (kernel, grid "
                        "size): calls, average time and total time divided by STEPS; prints text for profiles/")
    p.add_argument("--pmc-grid-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR",
                                                            "COPY_WRITE_DIR"),
                   help="summarise rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs (and the two of --copy-probe) "
                        "by (kernel, grid size): HBM MB read and written per launch, scaled by the copy probe; "
                        "prints text for profiles/")
    p.add_argument("--pmc-kernels", nargs="+", metavar="DIR",
                   help="FETCH_DIR WRITE_DIR SQ_DIR COPY_FETCH_DIR COPY_WRITE_DIR [SQ2_DIR]: summarise rocprofv3 "
                        "--pmc runs (FETCH_SIZE; WRITE_SIZE; SQ_INSTS_VALU, SQ_INSTS_SALU, SQ_INSTS_VMEM_RD, "
                        "SQ_INSTS_VMEM_WR, SQ_INSTS_LDS, SQ_LDS_BANK_CONFLICT, SQ_WAVES; optionally SQ_WAVE_CYCLES, "
                        "SQ_ACTIVE_INST_VALU, SQ_ACTIVE_INST_LDS, SQ_ACTIVE_INST_VMEM, SQ_WAIT_INST_ANY) per (kernel, "
                        "multigrid level): HBM bytes per launch (scaled by the copy probe), instructions per wave and "
                        "busy fractions; prints text for profiles/")
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


THETA_BC = [[0, 0, 2, 0, 0, 2]]  # potential temperature: periodic x/y, zero gradient (foextrap) at both walls


def run_gpu(case, precision, steps, warmup, profile, theta=False, prof_steps=2):
    """W untimed steps, then K steps bracketed by synchronisation, with an
    event between steps on the stream the host layer launches on (torch's
    current stream): the mean (ms) and the per-step times (step_ms).  With
    theta the potential temperature is advected as a passive scalar.  With
    profile, prof_steps more steps run with kernel timing on (each level-0
    half sweep between its own HIP events; the MAC solve then runs its
    host-tested path instead of the graph replay the timed steps use): the
    roofline's kernel time."""
    import torch

    ksgf = importlib.import_module("kynema-sgf_amd")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    opts = ksgf.default_options(**case.mlmg)
    opts.precision = 1 if precision == "mixed" else 0
    S = hs.HostStepper(case, options=opts)
    S.load()
    if theta:
        S.enable_scalar(np.asarray(THETA_BC, dtype=np.int32))
        S.load_scalar(case.extra["theta"][None])
    torch.cuda.synchronize()
    for _ in range(warmup):
        S.step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    infos = []
    evs[0].record(stream)
    for s in range(steps):
        S.step()
        evs[s + 1].record(stream)
        infos.append(dict(S.last_info))
    torch.cuda.synchronize()
    step_ms = [evs[s].elapsed_time(evs[s + 1]) for s in range(steps)]
    ms = evs[0].elapsed_time(evs[-1]) / steps
    prof = None
    if profile:
        ksgf.prof_reset()
        for _ in range(prof_steps):
            S.step()
        torch.cuda.synchronize()
        prof = ksgf.prof_read()
    # one more (untimed) step with the post-projection divergence check
    S.step(check_divergence=True)
    divmax = S.last_divmax
    vel = S.velocity()
    S.close()
    return ms, infos, prof, divmax, vel, step_ms


def step_bytes_per_cell(info, precision):
    """Algorithmic HBM bytes per cell of one advect + MAC-project step: 616 B
    for the streaming calls (CFL 24, ExtrapVelToFaces 72, MAC rhs 32,
    correction 56, fluxes 224, divergence 96, predictor 112) plus per V-cycle
    166.9 B in FP64, or 83.4 B of FP32 V-cycle and 40 B of FP64 residual and
    correction per defect-correction iteration in mixed precision."""
    nv = info["iters"]
    if precision == "mixed":
        return 616.0 + nv * 83.4 + max(info.get("outer_iters", nv), nv) * 40.0
    return 616.0 + nv * 166.9


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


PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_kernels_abl256_mixed.txt")


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


# Multigrid level of each launch, from the launch order of one solve: the
# outer residual starts at level 0, a restriction moves one level down, the
# first launch of a level's post-smoothing (prolongation) returns to the level
# it came from.
_MG_DOWN = ("k_resid_restrict_rt", "k_resid_restrict4", "k_restrict")
_MG_UP = ("k_pk_prolong_color", "k_pk_prolong_color4", "k_pk_prolong_add", "k_mg_prolong_red", "k_mg_prolong_add")
_MG_RESET = ("k_update_residual", "k_update_residual4")
_MG_KERNELS = ("k_pk_half", "k_pk_half4", "k_mg_smooth", "k_mg_ghosts", "k_fill", "k_coarse_vcycle") + _MG_DOWN + \
    _MG_UP + _MG_RESET


def _short(name):
    import re

    n = name.replace("(anonymous namespace)::", "")
    n = re.sub(r"\(.*", "", n).replace("void ", "").replace("ksgf::", "")
    return n.strip()


def _levels(rows):
    """rows: (dispatch id, short kernel name, ...) sorted by dispatch id;
    returns the multigrid level of each row ('-' for other kernels)"""
    lev, stack, out = 0, [], []
    for r in rows:
        k = r[1]
        base = k.split("<")[0]
        if base in _MG_RESET:
            lev, stack = 0, []
            out.append("0")
        elif base in _MG_UP:
            lev = stack.pop() if stack else lev
            out.append(str(lev))
        elif base in _MG_DOWN:
            out.append(str(lev))
            stack.append(lev)
            lev += 1
        elif base == "k_coarse_vcycle":
            out.append(f"{lev}+")
        elif base in _MG_KERNELS:
            out.append(str(lev))
        else:
            out.append("-")
    return out


def _trace_rows(d):
    import csv
    import glob

    rows = []
    for path in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        for r in csv.DictReader(open(path)):
            g = int(r["Grid_Size"]) if "Grid_Size" in r else \
                int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
            rows.append((int(r["Dispatch_Id"]), _short(r["Kernel_Name"]), g,
                         (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3))
    rows.sort()
    return rows


def _step_window(rows):
    """the launches of the steps: from the first CFL reduction to the last
    predictor update (the state upload, the projector's autotuner and the
    final download fall outside)"""
    names = [r[1] for r in rows]
    first = next((i for i, n in enumerate(names) if n.startswith("k_cfl")), 0)
    last = max((i for i, n in enumerate(names) if n == "k_predictor"), default=len(rows) - 1)
    return first, last + 1


def trace_summary(d, steps):
    import collections

    rows = _trace_rows(d)
    a, b = _step_window(rows)
    rows = rows[a:b]
    # the steps in the window: one predictor update each
    steps = sum(1 for r in rows if r[1] == "k_predictor") or steps
    lv = _levels(rows)
    acc = collections.defaultdict(list)
    for r, l in zip(rows, lv):
        acc[(r[1], l)].append(r[3])
    tot = sum(r[3] for r in rows)
    lines = [f"{'kernel':54s} {'level':>5s} {'calls':>6s} {'avg us':>9s} {'ms/step':>8s} {'%':>6s}"]
    for (k, l), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        t = sum(v)
        if t < 0.002 * tot:
            continue
        lines.append(f"{k[:54]:54s} {l:>5s} {len(v):6d} {t / len(v):9.1f} {t * 1e-3 / steps:8.3f} {100 * t / tot:6.2f}")
    lines.append(f"total kernel time {tot * 1e-3 / steps:.3f} ms/step over {steps} steps")
    return "\n".join(lines)


def trace_grid_summary(d, steps):
    """every launch of a kernel trace grouped by (kernel, grid threads)"""
    import collections

    rows = _trace_rows(d)
    acc = collections.defaultdict(list)
    for r in rows:
        acc[(r[1], r[2])].append(r[3])
    tot = sum(r[3] for r in rows)
    lines = [f"{'kernel':48s} {'grid':>10s} {'calls':>6s} {'avg us':>9s} {'ms/step':>8s} {'%':>6s}"]
    for (k, g), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        t = sum(v)
        if t < 0.002 * tot:
            continue
        lines.append(f"{k[:48]:48s} {g:10d} {len(v):6d} {t / len(v):9.1f} {t * 1e-3 / steps:8.3f} {100 * t / tot:6.2f}")
    lines.append(f"total kernel time {tot * 1e-3:.2f} ms = {tot * 1e-3 / steps:.2f} ms/step over {steps} steps")
    return "\n".join(lines)


def _pmc_dispatches(d):
    """per dispatch: (kernel, grid threads) and the summed counter values"""
    import collections
    import csv
    import glob

    per = collections.defaultdict(dict)
    key = {}
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(path)):
            i = int(r["Dispatch_Id"])
            key[i] = (_short(r["Kernel_Name"]), int(r["Grid_Size"]))
            per[i][r["Counter_Name"]] = per[i].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return [(key[i], per[i]) for i in sorted(per)]


def pmc_grid_summary(fdir, wdir, cfdir, cwdir):
    """mean HBM MB read / written per launch by (kernel, grid), each counter
    scaled so that the copy probe's 1 GiB reads and writes come out exact"""
    import collections

    true_kib = (1 << 27) * 8 / 1024.0
    scale = []
    for d, c in ((cfdir, "FETCH_SIZE"), (cwdir, "WRITE_SIZE")):
        vals = [x[c] for k, x in _pmc_dispatches(d) if "MulFunctor" in k[0]]
        scale.append(true_kib / (sum(vals) / len(vals)))
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for d, c, sc in ((fdir, "FETCH_SIZE", scale[0]), (wdir, "WRITE_SIZE", scale[1])):
        for k, x in _pmc_dispatches(d):
            if c in x:
                acc[k][c].append(x[c] * sc * 1024.0)
    lines = [f"copy probe scale: FETCH_SIZE x{scale[0]:.4f}, WRITE_SIZE x{scale[1]:.4f}",
             f"{'kernel':48s} {'grid':>10s} {'calls':>6s} {'read MB':>9s} {'write MB':>9s}"]

    def m(v):
        return sum(v) / len(v) if v else float("nan")

    for (k, g), c in sorted(acc.items(), key=lambda kv: -m(kv[1].get("FETCH_SIZE", [0]))):
        if k.startswith("__amd") or "MulFunctor" in k:
            continue
        lines.append(f"{k[:48]:48s} {g:10d} {len(c.get('FETCH_SIZE', [])):6d} "
                     f"{m(c.get('FETCH_SIZE', [])) / 1e6:9.2f} {m(c.get('WRITE_SIZE', [])) / 1e6:9.2f}")
    return "\n".join(lines)


def pmc_kernels(fdir, wdir, sqdir, cfdir, cwdir, sq2dir=None):
    """per (kernel, level): mean FETCH_SIZE / WRITE_SIZE bytes per launch
    (scaled by the copy probe) and SQ instruction counts per launch; with
    sq2dir (SQ_WAVE_CYCLES, SQ_ACTIVE_INST_VALU / _LDS / _VMEM,
    SQ_WAIT_INST_ANY) also the fractions of wave time issuing VALU, LDS and
    memory instructions and waiting on a dependency"""
    import collections
    import csv
    import glob

    def rows(d):
        per = collections.defaultdict(dict)
        names = {}
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(path)):
                i = int(r["Dispatch_Id"])
                names[i] = _short(r["Kernel_Name"])
                per[i][r["Counter_Name"]] = per[i].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        ids = sorted(per)
        return [(i, names[i]) for i in ids], [per[i] for i in ids]

    true_kib = (1 << 27) * 8 / 1024.0
    scale = []
    for d, c in ((cfdir, "FETCH_SIZE"), (cwdir, "WRITE_SIZE")):
        k, v = rows(d)
        vals = [x[c] for kk, x in zip(k, v) if "MulFunctor" in kk[1]]
        scale.append(true_kib / (sum(vals) / len(vals)))
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for d, sc in ((fdir, scale[0]), (wdir, scale[1]), (sqdir, None)) + (((sq2dir, None),) if sq2dir else ()):
        k, v = rows(d)
        a, b = _step_window(k)
        k, v = k[a:b], v[a:b]
        lv = _levels(k)
        for kk, x, l in zip(k, v, lv):
            for c, val in x.items():
                out[(kk[1], l)][c].append(val * sc * 1024.0 if sc else val)
    lines = [f"copy probe scale: FETCH_SIZE x{scale[0]:.4f}, WRITE_SIZE x{scale[1]:.4f} "
             f"(1 GiB read + 1 GiB written by torch.mul(x, 1.0))",
             f"{'kernel':48s} {'level':>5s} {'calls':>5s} {'read MB':>9s} {'write MB':>9s} {'VALU/wave':>10s} "
             f"{'SALU/wave':>10s} {'VMEM rd/wave':>12s} {'VMEM wr/wave':>12s} {'LDS/wave':>9s} {'LDS bank cf':>11s} "
             f"{'waves':>9s}" + (f" {'VALU busy':>9s} {'LDS busy':>8s} {'VMEM busy':>9s} {'dep wait':>8s}" if sq2dir else "")]

    def m(v):
        return sum(v) / len(v) if v else float("nan")

    for (k, l), c in sorted(out.items(), key=lambda kv: -m(kv[1].get("FETCH_SIZE", [0])) - m(kv[1].get("WRITE_SIZE",
                                                                                                         [0]))):
        if k.startswith("__amd") or "MulFunctor" in k or "elementwise" in k:
            continue
        w = m(c.get("SQ_WAVES", []))
        lines.append(f"{k[:48]:48s} {l:>5s} {len(c.get('FETCH_SIZE', [])):5d} {m(c.get('FETCH_SIZE', [])) / 1e6:9.2f} "
                     f"{m(c.get('WRITE_SIZE', [])) / 1e6:9.2f} {m(c.get('SQ_INSTS_VALU', [])) / w:10.1f} "
                     f"{m(c.get('SQ_INSTS_SALU', [])) / w:10.1f} {m(c.get('SQ_INSTS_VMEM_RD', [])) / w:12.1f} "
                     f"{m(c.get('SQ_INSTS_VMEM_WR', [])) / w:12.1f} {m(c.get('SQ_INSTS_LDS', [])) / w:9.1f} "
                     f"{m(c.get('SQ_LDS_BANK_CONFLICT', [])) / w:11.1f} {w:9.0f}" +
                     ((lambda wc: f" {m(c.get('SQ_ACTIVE_INST_VALU', [])) / wc:9.3f} "
                                  f"{m(c.get('SQ_ACTIVE_INST_LDS', [])) / wc:8.3f} "
                                  f"{m(c.get('SQ_ACTIVE_INST_VMEM', [])) / wc:9.3f} "
                                  f"{m(c.get('SQ_WAIT_INST_ANY', [])) / wc:8.3f}")(m(c.get('SQ_WAVE_CYCLES', [float('nan')])))
                      if sq2dir else ""))
    return "\n".join(lines)


def pmc_traffic():
    """HBM bytes per level-0 half-sweep launch from the committed counter
    summary (read + write MB of the level-0 k_pk_half4 rows, weighted by
    their launches), or None."""
    tot = n = 0.0
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith("k_pk_half4<"):
                name_end = line.index(">") + 1
                rest = line[name_end:].split()
                if rest[0] == "0":
                    calls, rd, wr = int(rest[1]), float(rest[2]), float(rest[3])
                    tot += calls * (rd + wr) * 1e6
                    n += calls
    except OSError:
        return None
    return tot / n if n else None


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


def fixed_dt_step(n, steps, warmup):
    """Fixed-dt steps (time.fixed_dt, the setting of most reference decks)
    eager and replayed from hipGraphs with one host read per step
    (HostStepper.enable_graph), on the Taylor-Green 64^3 case with 32^3 boxes
    when n is 64^3, else the neutral ABL box of size n; mixed-precision MAC
    solve.  dt = the first adaptive step's.  Both runs must end bit-identical."""
    import torch

    ksgf = importlib.import_module("kynema-sgf_amd")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    if tuple(n) == (64, 64, 64):
        case = wl.taylor_green(64, 32)
        name = "Taylor-Green 64^3, 8 boxes of 32^3"
    else:
        case = wl.abl(tuple(n), extent=(20.0 * n[0], 20.0 * n[1], 4.0 * n[2]), nsteps=steps)
        name = f"neutral ABL {n[0]}x{n[1]}x{n[2]}, uneven boxes {{32,48,64}}"
    res = {}
    vels = {}
    for graph in (False, True):
        opts = ksgf.default_options(**case.mlmg)
        opts.precision = 1
        S = hs.HostStepper(case, options=opts)
        S.load()
        if graph:
            S.enable_graph(True)
        S.step()
        dt = S.last_dt
        for _ in range(warmup + 2):  # the first fixed-dt steps capture the graphs
            S.step(dt)
        torch.cuda.synchronize()
        t = time.perf_counter()
        its = []
        for _ in range(steps):
            S.step(dt)
            its.append(S.last_info["iters"])
        torch.cuda.synchronize()
        res[graph] = ((time.perf_counter() - t) / steps * 1e3, its, dt)
        vels[graph] = S.velocity()
        S.close()
    ms_g, ms_e = res[True][0], res[False][0]
    return {
        "metric": "cell-updates/s per fixed-dt advect + MAC-project step, replayed from hipGraphs",
        "value": case.ncell / (ms_g * 1e-3),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_g,
        "higher_is_better": True,
        "dtype": "f64+f32mg",
        "data": "synthetic",
        "config": {"workload": f"{name}, fixed dt {res[True][2]:.6g}, PPM, MAC mixed precision rtol 1e-11",
                   "mlmg_iters": res[True][1]},
        "ms_per_step_eager": ms_e,
        "speedup_graph_vs_eager": ms_e / ms_g,
        "bit_identical": bool(np.array_equal(vels[True], vels[False])),
    }


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
    """Two-level advect + MAC-project steps (the C++ TwoLevelPredictor:
    per-level Godunov, composite MAC projection, fill-patches, flux
    averaging, average_down) on the 1,500-box ragged 320x320x480 level 0 with
    600 level-1 boxes over 40% of the domain; the oracle's two-level step on a
    64x64x96 copy of the layout is the CPU baseline."""
    import torch

    ksgf = importlib.import_module("kynema-sgf_amd")
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    case, fba = wl.ragged_two_level()
    dc, df = ksgf.DeviceBoxArray(case.ba), ksgf.DeviceBoxArray(fba, covers_domain=False)
    vc = _cells_on_device(case.ba, 3, 3, _tgv_cells(case.geom))
    vf = _cells_on_device(fba, 3, 3, _tgv_cells(fba.geom))
    df.average_down_to(vf, 3, dc, vc, 3, 3)
    S = hs.HostTwoLevel(case, fba)
    S.load(vc, vf)
    del vc, vf, dc, df
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
    if a.trace_summary:
        print(trace_summary(a.trace_summary[0], int(a.trace_summary[1])))
        return
    if a.pmc_grid_summary:
        print(pmc_grid_summary(*a.pmc_grid_summary))
        return
    if a.trace_grid_summary:
        print(trace_grid_summary(a.trace_grid_summary[0], int(a.trace_grid_summary[1])))
        return
    if a.pmc_kernels:
        if len(a.pmc_kernels) not in (5, 6):
            sys.exit("--pmc-kernels: FETCH_DIR WRITE_DIR SQ_DIR COPY_FETCH_DIR COPY_WRITE_DIR [SQ2_DIR]")
        print(pmc_kernels(*a.pmc_kernels))
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
    if a.workload == "fixed-dt":
        print(json.dumps(fixed_dt_step(tuple(a.n), a.steps, a.warmup)))
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
    ms, infos, prof, divmax, vel, step_ms = res[head]
    value = case.ncell / (ms * 1e-3)
    med = float(np.median(step_ms))
    bpc = float(np.mean([step_bytes_per_cell(i, head) for i in infos]))
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
        "median_ms_per_step": med,
        "value_median": case.ncell / (med * 1e-3),
        "step_ms": step_ms,
    }
    if "fp64" in res and head != "fp64":
        ms64, infos64, _, div64, vel64, step64 = res["fp64"]
        med64 = float(np.median(step64))
        out["fp64"] = {"value": case.ncell / (ms64 * 1e-3), "ms_per_step": ms64, "median_ms_per_step": med64,
                       "mlmg_iters": [i["iters"] for i in infos64], "post_projection_max_div": div64,
                       "full_step_gbs": case.ncell * float(np.mean([step_bytes_per_cell(i, "fp64") for i in infos64]))
                       / (med64 * 1e-3) / 1e9}
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
    # the whole step against the HBM roofline: the algorithmic bytes of the
    # logged V-cycle counts over the median step
    ach = case.ncell * bpc / (med * 1e-3) / 1e9
    out.setdefault("roofline", {})["full_step"] = {
        "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
        "bytes_per_cell": bpc, "bytes_per_step": case.ncell * bpc, "ms": med}
    # the same step with the potential temperature advected (C2 and C3 carry it)
    if "theta" in case.extra and not a.no_theta:
        msT, infosT, _, _, _, stepT = run_gpu(case, head, a.steps, a.warmup, profile=False, theta=True)
        medT = float(np.median(stepT))
        out["with_theta"] = {"value": case.ncell / (msT * 1e-3), "ms_per_step": msT, "median_ms_per_step": medT,
                             "mlmg_iters": [i["iters"] for i in infosT]}
    if not a.no_cpu:
        out["cpu_baseline"], (c2, reps, vref, pref) = cpu_baseline(a.cpu_n)
        out["vs_oracle"] = vs_oracle(c2, reps, vref, pref)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
