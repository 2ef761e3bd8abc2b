"""ksgf_cfl_max_noalloc (the CFL reduction with a file-scope device workspace,
for callers that cannot allocate) against ksgf_cfl_max through the C ABI:
equal results on two calls in a row, the second on a smaller input, so the
shared workspace's leftover partials from the first call cannot leak into
the second; ngrow without a launch table and an input beyond the workspace
are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CELL = 0


def _vel(dba, ba, seed, scale):
    rng = np.random.default_rng(seed)
    return dba.upload(scale * rng.standard_normal(ba.mf_size(CELL, 3, 3)))


def test_noalloc_cfl_equals_cfl_max(ksgf, workloads, gpu):
    lay = __import__("importlib").import_module("kynema-sgf_amd.layout")
    big = workloads.abl((96, 80, 64), extent=(1920.0, 1600.0, 256.0), sizes=(16, 32))
    small = workloads.taylor_green(16, 8)
    L = ksgf.load()
    for case, seed, scale in ((big, 1, 3.0), (small, 2, 0.5)):
        print(f"noalloc CFL on {case.name} ({case.ba.nbox} boxes)", flush=True)
        dba = ksgf.DeviceBoxArray(case.ba)
        v = _vel(dba, case.ba, seed, scale)
        want = dba.cfl_max(v, 3)
        got = dba.cfl_max_noalloc(v, 3)
        assert got == want, (got, want)
        # the same field with its valid region only
        v0 = dba.upload(case.ba.scatter(case.ba.gather(v.cpu().numpy(), CELL, 3, 3), CELL, 0))
        assert dba.cfl_max_noalloc(v0, 0) == want
        out = C.c_double()
        rc = L.ksgf_cfl_max_noalloc(dba.h, C.cast(C.c_void_p(v.data_ptr()), C.POINTER(C.c_double)), 4,
                                    C.byref(out), ksgf._stream())
        assert rc == ksgf.KSGF_ERR_INVALID and b"launch table" in L.ksgf_last_error()
        dba.close()


def test_noalloc_cfl_rejects_beyond_workspace(ksgf, gpu):
    """a 1024 x 1024 x 512 BoxArray (537 M cells, two workspaces' worth) is
    refused on the host: the velocity pointer is never read"""
    lay = __import__("importlib").import_module("kynema-sgf_amd.layout")
    geom = lay.Geometry((1024, 1024, 512), (0.0, 0.0, 0.0), (1.0, 1.0, 0.5), (1, 1, 1))
    ba = lay.BoxArray(geom, lay.chop_uniform(geom.n_cell, 512))
    dba = ksgf.DeviceBoxArray(ba)
    import torch

    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    with pytest.raises(ksgf.KsgfError) as e:
        dba.cfl_max_noalloc(dummy, 0)
    assert e.value.code == ksgf.KSGF_ERR_INVALID and "workspace" in str(e.value)
    dba.close()
