"""Rebuild the golden fixtures' cases (tests/golden/*.npz) without the
workload generators."""
import importlib
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["tgv8_step", "wall16_step"]


def load(name):
    lay = importlib.import_module("kynema-sgf_amd.layout")
    wl = importlib.import_module("kynema-sgf_amd.workloads")
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    n = tuple(int(v) for v in z["n_cell"])
    geom = lay.Geometry(n, (0.0, 0.0, 0.0), tuple(float(v) for v in z["prob_hi"]),
                        tuple(int(v) for v in z["periodic"]))
    ba = lay.BoxArray(geom, z["boxes"])
    case = wl.Case(name, geom, ba, z["vel"], z["rho"], z["gp"], z["bcrec"], z["bcval"], tuple(z["lobc"]),
                   tuple(z["hibc"]), float(z["cfl"]), 1, mg_rtol=float(z["mg_rtol"]), mg_atol=float(z["mg_atol"]))
    return case, z
