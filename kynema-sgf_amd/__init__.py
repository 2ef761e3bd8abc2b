"""kynema-sgf_amd: MI355X (gfx950) implementation of Kynema-SGF's Godunov PPM
advection + MAC projection step behind the AMReX-Hydro MultiFab-level calls.

The product is libksgf.so (HIP kernels + the C ABI of include/ksgf.h).
This module is its ctypes binding for Python callers (tests, bench): device
buffers are torch CUDA tensors (HIP on ROCm) passed by data pointer, work is
enqueued on torch's current stream.  torch is imported before the library is
loaded so both share one HIP runtime.

There is no CPU fallback: every entry point raises if libksgf.so is missing
or a call fails.
"""
from __future__ import annotations

import ctypes as C
import os

from . import layout  # noqa: F401
from .layout import CELL, XFACE, YFACE, ZFACE, BoxArray, Geometry  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libksgf.so")

_dp = C.c_void_p
_ip = C.POINTER(C.c_int)


class KsgfGeometry(C.Structure):
    _fields_ = [("domain_lo", C.c_int * 3), ("domain_hi", C.c_int * 3), ("prob_lo", C.c_double * 3),
                ("prob_hi", C.c_double * 3), ("is_periodic", C.c_int * 3)]


class KsgfBCRec(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class KsgfMLMGOptions(C.Structure):
    _fields_ = [("maxiter", C.c_int), ("num_pre_smooth", C.c_int), ("num_post_smooth", C.c_int),
                ("num_final_smooth", C.c_int), ("num_bottom_smooth", C.c_int), ("bottom_maxiter", C.c_int),
                ("bottom_rtol", C.c_double), ("bottom_atol", C.c_double), ("do_semicoarsening", C.c_int),
                ("max_coarsening_level", C.c_int), ("precision", C.c_int)]


class KsgfMLMGInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer_iters", C.c_int), ("r0", C.c_double), ("rfinal", C.c_double),
                ("converged", C.c_int), ("nlevels", C.c_int), ("bottom_iters", C.c_int), ("bottom_fail", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# every symbol include/ksgf.h declares (checked by the CPU test suite)
EXPORTS = [
    "ksgf_last_error", "ksgf_version", "ksgf_boxarray_create", "ksgf_boxarray_destroy", "ksgf_multifab_size",
    "ksgf_fill_boundary", "ksgf_fill_physbc", "ksgf_set_inflow_velocity", "ksgf_ppm_predict_state_on_face",
    "ksgf_extrap_vel_to_faces", "ksgf_compute_fluxes_from_state", "ksgf_compute_divergence",
    "ksgf_mlmg_default_options", "ksgf_macproj_create", "ksgf_macproj_set_beta", "ksgf_macproj_project",
    "ksgf_macproj_get_phi", "ksgf_macproj_destroy", "ksgf_macproj_solve", "ksgf_cfl_max", "ksgf_cfl_max_noalloc", "ksgf_force_cfl_max", "ksgf_macproj_set_beta_faces",
    "ksgf_average_cellcenter_to_face", "ksgf_invert",
    "ksgf_predictor_update",
    "ksgf_icns_src_term", "ksgf_scale_by_rho", "ksgf_prof_reset", "ksgf_prof_read", "ksgf_mac_divergence_max",
    "ksgf_poisson_batch_create", "ksgf_poisson_batch_solve", "ksgf_poisson_batch_destroy",
    "ksgf_adjoint_create", "ksgf_adjoint_forward", "ksgf_adjoint_backward", "ksgf_adjoint_destroy",
    "ksgf_plotfile_write", "ksgf_plotfile_read", "ksgf_vismf_write", "ksgf_vismf_read",
    "ksgf_checkpoint_write_header", "ksgf_nodalproj_create", "ksgf_nodalproj_set_sigma", "ksgf_nodalproj_node_dims",
    "ksgf_nodalproj_project", "ksgf_nodalproj_get_phi", "ksgf_nodalproj_compute_rhs", "ksgf_nodalproj_solve",
    "ksgf_nodalproj_destroy", "ksgf_set_bndry", "ksgf_incflo_add_gradp", "ksgf_abeclap_create",
    "ksgf_abeclap_set_scalars", "ksgf_abeclap_solve", "ksgf_abeclap_apply", "ksgf_abeclap_destroy",
    "ksgf_icns_src_term_divtau", "ksgf_boxarray_create_level", "ksgf_average_down", "ksgf_average_down_faces",
    "ksgf_fillpatch_two_levels", "ksgf_macproj_levels_create", "ksgf_macproj_levels_project",
    "ksgf_macproj_levels_get_phi", "ksgf_macproj_levels_destroy", "ksgf_fillpatch_faces_two_levels",
    "ksgf_tensorop_create", "ksgf_tensorop_set_scalars", "ksgf_tensorop_set_acoeffs",
    "ksgf_tensorop_set_shear_viscosity", "ksgf_tensorop_apply", "ksgf_tensorop_solve", "ksgf_tensorop_destroy",
    "ksgf_average_velocity_eta_to_faces", "ksgf_macproj_project_begin", "ksgf_macproj_project_finish",
    "ksgf_macproj_project_correct",
]

_lib = None


# status codes of include/ksgf.h
KSGF_OK = 0
KSGF_ERR_INVALID = -1
KSGF_ERR_UNSUPPORTED = -2
KSGF_ERR_HIP = -3
KSGF_ERR_NOMEM = -4
KSGF_ERR_NOT_CONVERGED = 1
KSGF_ERR_DIVERGED = -5


class KsgfError(RuntimeError):
    """A failed call; ``code`` is the ksgf status code where there is one."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def load(path=None):
    """Load libksgf.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (one HIP runtime for torch and the library)

    path = path or LIBPATH
    if not os.path.exists(path):
        raise KsgfError(f"{path} is not built: run python kynema-sgf_amd/build.py")
    L = C.CDLL(path)
    L.ksgf_last_error.restype = C.c_char_p
    L.ksgf_version.restype = C.c_char_p
    L.ksgf_boxarray_create.argtypes = [C.POINTER(KsgfGeometry), _ip, C.c_int, C.POINTER(C.c_void_p)]
    L.ksgf_boxarray_destroy.argtypes = [C.c_void_p]
    L.ksgf_multifab_size.restype = C.c_longlong
    L.ksgf_multifab_size.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.ksgf_fill_boundary.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_fill_physbc.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.POINTER(KsgfBCRec),
                                   C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_set_inflow_velocity.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.POINTER(KsgfBCRec),
                                           C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_ppm_predict_state_on_face.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int,
                                                 C.POINTER(KsgfBCRec), C.c_double, C.c_int, _dp, _dp, C.c_void_p]
    L.ksgf_extrap_vel_to_faces.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.POINTER(KsgfBCRec), C.c_double,
                                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_compute_fluxes_from_state.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                                 _dp, _dp, _dp, _dp, _dp, C.c_double, C.POINTER(KsgfBCRec), _ip,
                                                 C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_compute_divergence.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_void_p]
    L.ksgf_mlmg_default_options.argtypes = [C.POINTER(KsgfMLMGOptions)]
    L.ksgf_mlmg_default_options.restype = None
    L.ksgf_macproj_create.argtypes = [C.c_void_p, _ip, _ip, C.c_double, C.POINTER(KsgfMLMGOptions),
                                      C.POINTER(C.c_void_p)]
    L.ksgf_macproj_set_beta.argtypes = [C.c_void_p, C.c_double]
    L.ksgf_macproj_project.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double, C.POINTER(KsgfMLMGInfo),
                                       C.c_void_p]
    L.ksgf_macproj_project_begin.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_void_p]
    L.ksgf_macproj_project_finish.argtypes = [C.c_void_p, C.POINTER(KsgfMLMGInfo), C.c_void_p]
    L.ksgf_macproj_project_correct.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_void_p]
    L.ksgf_macproj_solve.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.POINTER(KsgfMLMGInfo),
                                     C.c_void_p]
    L.ksgf_macproj_get_phi.argtypes = [C.c_void_p, _dp, C.c_void_p]
    L.ksgf_poisson_batch_create.argtypes = [C.c_int, _ip, C.POINTER(C.c_double), _ip, _ip, C.POINTER(C.c_double),
                                            C.POINTER(KsgfMLMGOptions), C.POINTER(C.c_void_p)]
    L.ksgf_poisson_batch_solve.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_double,
                                           C.c_double, C.c_int, C.POINTER(KsgfMLMGInfo), C.c_void_p]
    L.ksgf_poisson_batch_destroy.argtypes = [C.c_void_p]
    L.ksgf_adjoint_create.argtypes = [_ip, C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double,
                                      C.POINTER(KsgfMLMGOptions), C.POINTER(C.c_void_p)]
    L.ksgf_adjoint_forward.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int, _dp, _dp, _ip,
                                       C.c_void_p]
    L.ksgf_adjoint_backward.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_void_p]
    L.ksgf_adjoint_destroy.argtypes = [C.c_void_p]
    L.ksgf_macproj_destroy.argtypes = [C.c_void_p]
    L.ksgf_macproj_levels_create.argtypes = [C.c_void_p, C.c_void_p, _ip, _ip, C.c_double,
                                             C.POINTER(KsgfMLMGOptions), C.POINTER(C.c_void_p)]
    L.ksgf_macproj_levels_project.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int,
                                              C.c_double, C.c_double, C.POINTER(KsgfMLMGInfo), C.c_void_p]
    L.ksgf_macproj_levels_get_phi.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_void_p]
    L.ksgf_macproj_levels_destroy.argtypes = [C.c_void_p]
    L.ksgf_cfl_max.argtypes = [C.c_void_p, _dp, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_macproj_set_beta_faces.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_void_p]
    L.ksgf_invert.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p]
    L.ksgf_average_cellcenter_to_face.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_void_p]
    L.ksgf_cfl_max_noalloc.argtypes = [C.c_void_p, _dp, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_force_cfl_max.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_predictor_update.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp,
                                        _dp, C.c_double, C.c_double, C.c_int, C.c_void_p]
    L.ksgf_icns_src_term.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_scale_by_rho.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p]
    L.ksgf_mac_divergence_max.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_prof_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.POINTER(C.c_double), C.c_char_p,
                                 C.c_int]
    L.ksgf_plotfile_write.argtypes = [C.c_char_p, C.c_void_p, _dp, C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_double,
                                      C.c_int, C.c_void_p]
    L.ksgf_plotfile_read.argtypes = [C.c_char_p, C.c_void_p, _dp, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_vismf_write.argtypes = [C.c_char_p, C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_vismf_read.argtypes = [C.c_char_p, C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_checkpoint_write_header.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                               C.c_double]
    L.ksgf_nodalproj_create.argtypes = [C.c_void_p, _ip, _ip, C.c_double, C.POINTER(KsgfMLMGOptions),
                                        C.POINTER(C.c_void_p)]
    L.ksgf_nodalproj_set_sigma.argtypes = [C.c_void_p, C.c_double]
    L.ksgf_nodalproj_node_dims.argtypes = [C.c_void_p, _ip]
    L.ksgf_nodalproj_project.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_double,
                                         C.POINTER(KsgfMLMGInfo), C.c_void_p]
    L.ksgf_nodalproj_get_phi.argtypes = [C.c_void_p, _dp, C.c_void_p]
    L.ksgf_nodalproj_compute_rhs.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_void_p]
    L.ksgf_nodalproj_solve.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.POINTER(KsgfMLMGInfo),
                                       C.c_void_p]
    L.ksgf_nodalproj_destroy.argtypes = [C.c_void_p]
    L.ksgf_set_bndry.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p]
    L.ksgf_abeclap_create.argtypes = [C.c_void_p, _ip, _ip, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(KsgfMLMGOptions), C.POINTER(C.c_void_p)]
    L.ksgf_abeclap_set_scalars.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.ksgf_abeclap_solve.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.POINTER(KsgfMLMGInfo), C.c_void_p]
    L.ksgf_abeclap_apply.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_abeclap_destroy.argtypes = [C.c_void_p]
    L.ksgf_tensorop_create.argtypes = [C.c_void_p, _ip, _ip, C.c_void_p, C.POINTER(C.c_void_p)]
    L.ksgf_tensorop_set_scalars.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.ksgf_tensorop_set_acoeffs.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p]
    L.ksgf_tensorop_set_shear_viscosity.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_void_p]
    L.ksgf_tensorop_apply.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.c_void_p]
    L.ksgf_tensorop_solve.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_double, C.c_void_p,
                                      C.c_void_p]
    L.ksgf_tensorop_destroy.argtypes = [C.c_void_p]
    L.ksgf_average_velocity_eta_to_faces.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_void_p]
    L.ksgf_boxarray_create_level.argtypes = [C.POINTER(KsgfGeometry), _ip, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.ksgf_average_down.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.ksgf_average_down_faces.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_void_p, _dp, _dp, _dp, C.c_int,
                                          C.c_int, C.c_int, C.c_void_p]
    L.ksgf_fillpatch_two_levels.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p, _dp, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(KsgfBCRec), C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_fillpatch_faces_two_levels.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_void_p, _dp, C.c_int,
                                                  C.c_int, C.POINTER(KsgfBCRec), C.POINTER(C.c_double), C.c_void_p]
    L.ksgf_icns_src_term_divtau.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_void_p]
    L.ksgf_incflo_add_gradp.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double,
                                        C.c_void_p]
    _lib = L
    return L


def prof_reset():
    check(load().ksgf_prof_reset(), "prof_reset")


def prof_read():
    ms, n, b = C.c_double(), C.c_longlong(), C.c_double()
    name = C.create_string_buffer(128)
    check(load().ksgf_prof_read(C.byref(ms), C.byref(n), C.byref(b), name, 128), "prof_read")
    return {"ms": ms.value, "launches": n.value, "bytes": b.value, "name": name.value.decode()}


def check(rc, what="ksgf call", allow=(0,)):
    if rc not in allow:
        msg = load().ksgf_last_error().decode()
        raise KsgfError(f"{what} failed ({rc}): {msg}", rc)
    return rc


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    if t is None:
        return None
    import torch

    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous(), "expected a contiguous float64 device tensor"
    return C.c_void_p(t.data_ptr())


def bcrecs(bcrec):
    import numpy as np

    bcrec = np.asarray(bcrec, dtype=np.int32).reshape(-1, 6)
    arr = (KsgfBCRec * len(bcrec))()
    for n, row in enumerate(bcrec):
        for d in range(3):
            arr[n].lo[d] = int(row[d])
            arr[n].hi[d] = int(row[3 + d])
    return arr


def default_options(**kw):
    o = KsgfMLMGOptions()
    load().ksgf_mlmg_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


DT_GROWTH = 0.1  # time.max_dt_growth default (SimTime.H:248)


def set_current_cfl(cfl, conv_cfl, force_cfl, dt_prev=0.0, dt_growth=DT_GROWTH):
    """SimTime::set_current_cfl (SimTime.cpp:162-205) for an adaptive step
    after initialisation, without a diffusive CFL (Godunov with implicit or
    Crank-Nicolson diffusion) and without output-time clipping:
    dt = 2 cfl / (C + sqrt(C^2 + 4 F)), then at most (1 + dt_growth) dt_prev
    once a previous step exists.  Raises where the reference aborts (CFL
    below machine epsilon)."""
    import math

    eps = 2.220446049250313e-16
    cd = conv_cfl + 0.0
    unit = cd + math.sqrt((cd * cd) + (4.0 * force_cfl))
    if unit < eps:
        raise KsgfError("CFL is below machine epsilon and the time step is adaptive: set a fixed dt")
    dt = 2.0 * cfl / max(unit, eps)
    if dt_prev > 0.0:
        dt = min(dt, (1.0 + dt_growth) * dt_prev)
    return dt


class DeviceBoxArray:
    """ksgf_boxarray handle for a layout.BoxArray."""

    def __init__(self, ba: BoxArray, covers_domain=True):
        import numpy as np

        L = load()
        self.ba = ba
        g = KsgfGeometry()
        for d in range(3):
            g.domain_lo[d] = ba.geom.dlo[d]
            g.domain_hi[d] = ba.geom.dhi[d]
            g.prob_lo[d] = ba.geom.prob_lo[d]
            g.prob_hi[d] = ba.geom.prob_hi[d]
            g.is_periodic[d] = int(ba.geom.periodic[d])
        self._boxes = np.ascontiguousarray(ba.boxes, dtype=np.int32)
        h = C.c_void_p()
        check(L.ksgf_boxarray_create_level(C.byref(g), self._boxes.ctypes.data_as(_ip), ba.nbox, int(covers_domain),
                                           C.byref(h)), "ksgf_boxarray_create")
        self.h = h

    def size(self, itype, ncomp, ngrow):
        n = load().ksgf_multifab_size(self.h, itype, ncomp, ngrow)
        if n < 0:
            raise KsgfError("ksgf_multifab_size failed")
        return int(n)

    def alloc(self, itype, ncomp, ngrow):
        import torch

        return torch.zeros(self.size(itype, ncomp, ngrow), dtype=torch.float64, device="cuda")

    def upload(self, flat):
        import torch

        return torch.from_numpy(flat).to("cuda")

    def close(self):
        if self.h:
            load().ksgf_boxarray_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- Hydro-level calls --------------------------------------------------
    def fill_boundary(self, mf, itype, ncomp, ngrow):
        check(load().ksgf_fill_boundary(self.h, _ptr(mf), itype, ncomp, ngrow, _stream()), "fill_boundary")

    def fill_physbc(self, mf, itype, ncomp, ngrow, bcrec, bcval):
        import numpy as np

        v = np.ascontiguousarray(bcval, dtype=np.float64)
        check(load().ksgf_fill_physbc(self.h, _ptr(mf), itype, ncomp, ngrow, bcrecs(bcrec),
                                      v.ctypes.data_as(C.POINTER(C.c_double)), _stream()), "fill_physbc")

    def fill_ghosts(self, mf, itype, ncomp, ngrow, bcrec, bcval):
        self.fill_boundary(mf, itype, ncomp, ngrow)
        self.fill_physbc(mf, itype, ncomp, ngrow, bcrec, bcval)

    def set_inflow_velocity(self, mac, ngrow, bcrec, bcval):
        import numpy as np

        v = np.ascontiguousarray(bcval, dtype=np.float64)
        check(load().ksgf_set_inflow_velocity(self.h, _ptr(mac[0]), _ptr(mac[1]), _ptr(mac[2]), ngrow, bcrecs(bcrec),
                                              v.ctypes.data_as(C.POINTER(C.c_double)), _stream()),
              "set_inflow_velocity")

    def ppm_predict_state(self, d, s, ncomp_s, ng_s, comp, vel_edge, ng_v, bcrec1, dt, limiter):
        Im = self.alloc(CELL, 1, 0)
        Ip = self.alloc(CELL, 1, 0)
        check(load().ksgf_ppm_predict_state_on_face(self.h, d, _ptr(s), ncomp_s, ng_s, comp, _ptr(vel_edge), ng_v,
                                                    bcrecs(bcrec1), dt, limiter, _ptr(Im), _ptr(Ip), _stream()),
              "ppm_predict_state_on_face")
        return Im, Ip

    def extrap_vel_to_faces(self, vel, f, bcrec, dt, use_ppm=1, use_forces_in_trans=0, limiter=1, allow_inflow=0,
                            mac=None):
        mac = mac or [self.alloc(XFACE + d, 1, 1) for d in range(3)]
        check(load().ksgf_extrap_vel_to_faces(self.h, _ptr(vel), _ptr(f), _ptr(mac[0]), _ptr(mac[1]), _ptr(mac[2]),
                                              bcrecs(bcrec), dt, use_ppm, use_forces_in_trans, limiter, allow_inflow,
                                              _stream()), "extrap_vel_to_faces")
        return mac

    def compute_fluxes(self, ncomp, q, q_bc, mac, divu, fq, dt, bcrec, iconserv, use_ppm=1, use_forces_in_trans=0,
                       is_velocity=1, area_weighted=0, limiter=1, allow_inflow=0, known_edge_state=0, faces=None,
                       flux=None):
        import numpy as np

        flux = flux or [self.alloc(XFACE + d, ncomp, 0) for d in range(3)]
        face = faces or [self.alloc(XFACE + d, ncomp, 0) for d in range(3)]
        ic = np.ascontiguousarray(iconserv, dtype=np.int32)
        check(load().ksgf_compute_fluxes_from_state(
            self.h, ncomp, _ptr(q), _ptr(q_bc), _ptr(flux[0]), _ptr(flux[1]), _ptr(flux[2]), _ptr(face[0]),
            _ptr(face[1]), _ptr(face[2]), known_edge_state, _ptr(mac[0]), _ptr(mac[1]), _ptr(mac[2]), _ptr(divu),
            _ptr(fq), dt, bcrecs(bcrec), ic.ctypes.data_as(_ip), use_ppm, use_forces_in_trans, is_velocity,
            area_weighted, limiter, allow_inflow, _stream()), "compute_fluxes_from_state")
        return flux, face

    def compute_divergence(self, flux, ncomp, mult=-1.0, area_weighted=0, out=None):
        div = out if out is not None else self.alloc(CELL, ncomp, 0)
        check(load().ksgf_compute_divergence(self.h, _ptr(div), _ptr(flux[0]), _ptr(flux[1]), _ptr(flux[2]), ncomp,
                                             mult, area_weighted, _stream()), "compute_divergence")
        return div

    def cfl_max(self, vel, ngrow):
        out = C.c_double()
        check(load().ksgf_cfl_max(self.h, _ptr(vel), ngrow, C.byref(out), _stream()), "cfl_max")
        return out.value

    def cfl_max_noalloc(self, vel, ngrow):
        out = C.c_double()
        check(load().ksgf_cfl_max_noalloc(self.h, _ptr(vel), ngrow, C.byref(out), _stream()), "cfl_max_noalloc")
        return out.value

    def average_cellcenter_to_face(self, cc, ng_cc, ng_face=0):
        out = [self.alloc(XFACE + d, 1, ng_face) for d in range(3)]
        check(load().ksgf_average_cellcenter_to_face(self.h, _ptr(cc), ng_cc, _ptr(out[0]), _ptr(out[1]),
                                                     _ptr(out[2]), ng_face, _stream()), "average_cellcenter_to_face")
        return out

    def average_velocity_eta_to_faces(self, cc, ng_cc, ng_face=0):
        """diffusion::average_velocity_eta_to_faces: face means of a cell
        viscosity, the adjacent cell on non-periodic domain faces"""
        out = [self.alloc(XFACE + d, 1, ng_face) for d in range(3)]
        check(load().ksgf_average_velocity_eta_to_faces(self.h, _ptr(cc), ng_cc, _ptr(out[0]), _ptr(out[1]),
                                                        _ptr(out[2]), ng_face, _stream()),
              "average_velocity_eta_to_faces")
        return out

    def invert(self, mf, itype, ncomp, ngrow, nghost, numerator):
        check(load().ksgf_invert(self.h, _ptr(mf), itype, ncomp, ngrow, nghost, numerator, _stream()), "invert")
        return mf

    def force_cfl_max(self, src, ng_src, rho, ng_rho):
        out = C.c_double()
        check(load().ksgf_force_cfl_max(self.h, _ptr(src), ng_src, _ptr(rho), ng_rho, C.byref(out), _stream()),
              "force_cfl_max")
        return out.value

    def mac_divergence_max(self, mac, ngrow=1):
        out = C.c_double()
        check(load().ksgf_mac_divergence_max(self.h, _ptr(mac[0]), _ptr(mac[1]), _ptr(mac[2]), ngrow, C.byref(out),
                                             _stream()), "mac_divergence_max")
        return out.value

    def predictor_update(self, vel_new, ng_new, vel_old, ng_old, rho_old, ng_ro, rho_new, ng_rn, conv, src, diff,
                         kappa, dt, ncomp=3):
        check(load().ksgf_predictor_update(self.h, _ptr(vel_new), ng_new, _ptr(vel_old), ng_old, _ptr(rho_old), ng_ro,
                                           _ptr(rho_new), ng_rn, _ptr(conv), _ptr(src), _ptr(diff), kappa, dt, ncomp,
                                           _stream()), "predictor_update")

    def icns_src_term(self, src, ng_src, gp, rho, ng_rho, scale_by_rho):
        check(load().ksgf_icns_src_term(self.h, _ptr(src), ng_src, _ptr(gp), _ptr(rho), ng_rho, scale_by_rho,
                                        _stream()), "icns_src_term")

    # --- plotfile / VisMF files ---------------------------------------------
    def write_plotfile(self, path, mf, ncomp, ngrow, names, time=0.0, step=0):
        arr = (C.c_char_p * ncomp)(*[n.encode() for n in names])
        check(load().ksgf_plotfile_write(str(path).encode(), self.h, _ptr(mf), ncomp, ngrow, arr, time, step,
                                         _stream()), "plotfile_write")

    def read_plotfile(self, path, mf, ncomp, ngrow):
        check(load().ksgf_plotfile_read(str(path).encode(), self.h, _ptr(mf), ncomp, ngrow, _stream()),
              "plotfile_read")

    def write_vismf(self, prefix, mf, itype, ncomp, ngrow):
        check(load().ksgf_vismf_write(str(prefix).encode(), self.h, _ptr(mf), itype, ncomp, ngrow, _stream()),
              "vismf_write")

    def read_vismf(self, prefix, mf, itype, ncomp, ngrow):
        check(load().ksgf_vismf_read(str(prefix).encode(), self.h, _ptr(mf), itype, ncomp, ngrow, _stream()),
              "vismf_read")

    def write_checkpoint_header(self, path, step, time, dt, dt_nm1=0.0, dt_nm2=0.0):
        check(load().ksgf_checkpoint_write_header(str(path).encode(), self.h, step, time, dt, dt_nm1, dt_nm2),
              "checkpoint_write_header")

    def average_down_to(self, fine, ng_f, crse_dba, crse, ng_c, ncomp, ratio=2):
        """amrex::average_down of this (fine) level's cell MultiFab into crse"""
        check(load().ksgf_average_down(self.h, _ptr(fine), ng_f, crse_dba.h, _ptr(crse), ng_c, ncomp, ratio,
                                       _stream()), "average_down")

    def average_down_faces_to(self, fine3, ng_f, crse_dba, crse3, ng_c, ncomp, ratio=2):
        """amrex::average_down_faces of this level's three face MultiFabs into crse3"""
        check(load().ksgf_average_down_faces(self.h, _ptr(fine3[0]), _ptr(fine3[1]), _ptr(fine3[2]), ng_f, crse_dba.h,
                                             _ptr(crse3[0]), _ptr(crse3[1]), _ptr(crse3[2]), ng_c, ncomp, ratio,
                                             _stream()), "average_down_faces")

    def fillpatch_two_levels(self, fine, ng, crse_dba, crse, ng_c, ncomp, bcrec, bcval, ratio=2):
        """FillPatchTwoLevels of this (fine) level's ghost cells from crse"""
        import numpy as np

        bv = np.ascontiguousarray(bcval, dtype=np.float64)
        check(load().ksgf_fillpatch_two_levels(self.h, _ptr(fine), ng, crse_dba.h, _ptr(crse), ng_c, ncomp, ratio,
                                               bcrecs(bcrec), bv.ctypes.data_as(C.POINTER(C.c_double)), _stream()),
              "fillpatch_two_levels")

    def fillpatch_faces_two_levels(self, fine, itype, ng, crse_dba, crse, ng_c, ncomp, bcrec, bcval):
        """FillPatchTwoLevels of this (fine) level's ghost faces from crse"""
        import numpy as np

        bv = np.ascontiguousarray(bcval, dtype=np.float64)
        check(load().ksgf_fillpatch_faces_two_levels(self.h, _ptr(fine), itype, ng, crse_dba.h, _ptr(crse), ng_c,
                                                     ncomp, bcrecs(bcrec), bv.ctypes.data_as(C.POINTER(C.c_double)),
                                                     _stream()), "fillpatch_faces_two_levels")

    def set_bndry(self, mf, itype, ncomp, ngrow, val=0.0):
        check(load().ksgf_set_bndry(self.h, _ptr(mf), itype, ncomp, ngrow, val, _stream()), "set_bndry")

    def scale_by_rho(self, out, ng_out, a, ng_a, rho, ng_rho, ncomp, ngrow, divide=0):
        check(load().ksgf_scale_by_rho(self.h, _ptr(out), ng_out, _ptr(a), ng_a, _ptr(rho), ng_rho, ncomp, ngrow,
                                       divide, _stream()), "scale_by_rho")


class MacProjector:
    """Hydro::MacProjector over a DeviceBoxArray."""

    def __init__(self, dba: DeviceBoxArray, lobc, hibc, beta=1.0, options=None):
        import numpy as np

        L = load()
        self.dba = dba
        self._lo = np.ascontiguousarray(lobc, dtype=np.int32)
        self._hi = np.ascontiguousarray(hibc, dtype=np.int32)
        self.options = options or default_options()
        h = C.c_void_p()
        check(L.ksgf_macproj_create(dba.h, self._lo.ctypes.data_as(_ip), self._hi.ctypes.data_as(_ip), beta,
                                    C.byref(self.options), C.byref(h)), "ksgf_macproj_create")
        self.h = h
        self.info = None

    def update_beta(self, beta):
        check(load().ksgf_macproj_set_beta(self.h, beta), "macproj_set_beta")

    def update_beta_faces(self, b3, ngrow=0):
        """beta on the faces: three face MultiFabs (device tensors)"""
        check(load().ksgf_macproj_set_beta_faces(self.h, _ptr(b3[0]), _ptr(b3[1]), _ptr(b3[2]), ngrow, _stream()),
              "macproj_set_beta_faces")

    def project(self, mac, rtol, atol):
        info = KsgfMLMGInfo()
        check(load().ksgf_macproj_project(self.h, _ptr(mac[0]), _ptr(mac[1]), _ptr(mac[2]), rtol, atol, C.byref(info),
                                          _stream()), "macproj_project", allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def project_begin(self, mac, rtol, atol, ncycles):
        """div u_mac and ncycles device-tested iterations (no host
        synchronisation: capturable)"""
        check(load().ksgf_macproj_project_begin(self.h, _ptr(mac[0]), _ptr(mac[1]), _ptr(mac[2]), rtol, atol,
                                                int(ncycles), _stream()), "macproj_project_begin")

    def project_finish(self):
        info = KsgfMLMGInfo()
        check(load().ksgf_macproj_project_finish(self.h, C.byref(info), _stream()), "macproj_project_finish",
              allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def project_correct(self, mac):
        check(load().ksgf_macproj_project_correct(self.h, _ptr(mac[0]), _ptr(mac[1]), _ptr(mac[2]), _stream()),
              "macproj_project_correct")

    def solve(self, rhs, phi, rtol, atol):
        info = KsgfMLMGInfo()
        check(load().ksgf_macproj_solve(self.h, _ptr(rhs), _ptr(phi), rtol, atol, C.byref(info), _stream()),
              "macproj_solve", allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def phi(self):
        import torch

        n = int(torch.tensor(self.dba.ba.geom.n_cell).prod())
        out = torch.zeros(n, dtype=torch.float64, device="cuda")
        check(load().ksgf_macproj_get_phi(self.h, _ptr(out), _stream()), "macproj_get_phi")
        return out

    def close(self):
        if self.h:
            load().ksgf_macproj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MacProjectorLevels:
    """Hydro::MacProjector over two AMR levels (coarse DeviceBoxArray
    covering the domain, fine DeviceBoxArray at ratio 2): composite solve."""

    def __init__(self, crse: DeviceBoxArray, fine: DeviceBoxArray, lobc, hibc, beta=1.0, options=None):
        import numpy as np

        L = load()
        self.crse, self.fine = crse, fine
        self._lo = np.ascontiguousarray(lobc, dtype=np.int32)
        self._hi = np.ascontiguousarray(hibc, dtype=np.int32)
        self.options = options or default_options()
        h = C.c_void_p()
        check(L.ksgf_macproj_levels_create(crse.h, fine.h, self._lo.ctypes.data_as(_ip),
                                           self._hi.ctypes.data_as(_ip), beta, C.byref(self.options), C.byref(h)),
              "ksgf_macproj_levels_create")
        self.h = h
        self.info = None

    def project(self, mac_c, mac_f, ng, rtol, atol):
        info = KsgfMLMGInfo()
        pc = (C.c_void_p * 3)(*[_ptr(a) for a in mac_c])
        pf = (C.c_void_p * 3)(*[_ptr(a) for a in mac_f])
        check(load().ksgf_macproj_levels_project(self.h, pc, pf, ng, rtol, atol, C.byref(info), _stream()),
              "macproj_levels_project", allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def phi(self):
        """(coarse domain-wide, fine cell MultiFab with one ghost layer)"""
        import torch

        nc = int(torch.tensor(self.crse.ba.geom.n_cell).prod())
        pc = torch.zeros(nc, dtype=torch.float64, device="cuda")
        pf = torch.zeros(self.fine.ba.mf_size(0, 1, 1), dtype=torch.float64, device="cuda")
        check(load().ksgf_macproj_levels_get_phi(self.h, _ptr(pc), _ptr(pf), 1, _stream()), "macproj_levels_get_phi")
        return pc, pf

    def close(self):
        if self.h:
            load().ksgf_macproj_levels_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TensorOp:
    """amrex::MLTensorOp + MLMG over a DeviceBoxArray (the default ICNS viscous
    operator): alpha a u - beta div(eta (grad u + grad u^T) - (2/3) eta (div u)
    I) on 3-component cell MultiFabs.  lobc / hibc: (3 components, 3
    directions) of KSGF_LO_* values."""

    def __init__(self, dba: DeviceBoxArray, lobc, hibc, options=None):
        import numpy as np

        L = load()
        self.dba = dba
        self._lo = np.ascontiguousarray(lobc, dtype=np.int32).ravel()
        self._hi = np.ascontiguousarray(hibc, dtype=np.int32).ravel()
        if self._lo.size != 9 or self._hi.size != 9:
            raise ValueError("TensorOp: lobc / hibc need 3 components x 3 directions")
        self.options = options or default_options()
        h = C.c_void_p()
        check(L.ksgf_tensorop_create(dba.h, self._lo.ctypes.data_as(_ip), self._hi.ctypes.data_as(_ip),
                                     C.byref(self.options), C.byref(h)), "ksgf_tensorop_create")
        self.h = h
        self.info = None

    def set_scalars(self, alpha, beta):
        check(load().ksgf_tensorop_set_scalars(self.h, alpha, beta), "tensorop_set_scalars")

    def set_acoeffs(self, a, ngrow):
        check(load().ksgf_tensorop_set_acoeffs(self.h, _ptr(a), ngrow, _stream()), "tensorop_set_acoeffs")

    def set_shear_viscosity(self, faces, ngrow):
        check(load().ksgf_tensorop_set_shear_viscosity(self.h, _ptr(faces[0]), _ptr(faces[1]), _ptr(faces[2]), ngrow,
                                                       _stream()), "tensorop_set_shear_viscosity")

    def apply(self, out, ng_out, x, ng_x):
        check(load().ksgf_tensorop_apply(self.h, _ptr(out), ng_out, _ptr(x), ng_x, _stream()), "tensorop_apply")

    def solve(self, sol, ng_sol, rhs, ng_rhs, rtol, atol):
        info = KsgfMLMGInfo()
        check(load().ksgf_tensorop_solve(self.h, _ptr(sol), ng_sol, _ptr(rhs), ng_rhs, rtol, atol, C.byref(info),
                                         _stream()), "tensorop_solve", allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def close(self):
        if self.h:
            load().ksgf_tensorop_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ABecLaplacian:
    """amrex::MLABecLaplacian (constant a, b) + MLMG over a DeviceBoxArray:
    alpha a phi - beta b lap(phi) = rhs, one component at a time."""

    def __init__(self, dba: DeviceBoxArray, lobc, hibc, alpha, a, beta, b, options=None):
        import numpy as np

        L = load()
        self.dba = dba
        self._lo = np.ascontiguousarray(lobc, dtype=np.int32)
        self._hi = np.ascontiguousarray(hibc, dtype=np.int32)
        self.options = options or default_options()
        h = C.c_void_p()
        check(L.ksgf_abeclap_create(dba.h, self._lo.ctypes.data_as(_ip), self._hi.ctypes.data_as(_ip), alpha, a, beta,
                                    b, C.byref(self.options), C.byref(h)), "ksgf_abeclap_create")
        self.h = h
        self.info = None

    def set_scalars(self, alpha, beta):
        check(load().ksgf_abeclap_set_scalars(self.h, alpha, beta), "abeclap_set_scalars")

    def solve(self, sol, ng_sol, rhs, ng_rhs, ncomp, comp, rtol, atol):
        info = KsgfMLMGInfo()
        check(load().ksgf_abeclap_solve(self.h, _ptr(sol), ng_sol, _ptr(rhs), ng_rhs, ncomp, comp, rtol, atol,
                                        C.byref(info), _stream()), "abeclap_solve", allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def apply(self, out, ng_out, x, ng_x, ncomp, comp):
        check(load().ksgf_abeclap_apply(self.h, _ptr(out), ng_out, _ptr(x), ng_x, ncomp, comp, _stream()),
              "abeclap_apply")

    def close(self):
        if self.h:
            load().ksgf_abeclap_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NodalProjector:
    """Hydro::NodalProjector with a constant sigma over a DeviceBoxArray."""

    def __init__(self, dba: DeviceBoxArray, lobc, hibc, sigma=1.0, options=None):
        import numpy as np

        L = load()
        self.dba = dba
        self._lo = np.ascontiguousarray(lobc, dtype=np.int32)
        self._hi = np.ascontiguousarray(hibc, dtype=np.int32)
        self.options = options or default_options()
        h = C.c_void_p()
        check(L.ksgf_nodalproj_create(dba.h, self._lo.ctypes.data_as(_ip), self._hi.ctypes.data_as(_ip), sigma,
                                      C.byref(self.options), C.byref(h)), "ksgf_nodalproj_create")
        self.h = h
        N = (C.c_int * 3)()
        check(L.ksgf_nodalproj_node_dims(h, N), "nodalproj_node_dims")
        self.N = tuple(N)
        self.info = None

    def set_sigma(self, sigma):
        check(load().ksgf_nodalproj_set_sigma(self.h, sigma), "nodalproj_set_sigma")

    def project(self, vel, ng_vel, rtol, atol, gradphi=None, ng_gp=0):
        info = KsgfMLMGInfo()
        check(load().ksgf_nodalproj_project(self.h, _ptr(vel), ng_vel, _ptr(gradphi), ng_gp, rtol, atol,
                                            C.byref(info), _stream()), "nodalproj_project", allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def compute_rhs(self, vel, ng_vel):
        import torch

        out = torch.zeros(self.N[0] * self.N[1] * self.N[2], dtype=torch.float64, device="cuda")
        check(load().ksgf_nodalproj_compute_rhs(self.h, _ptr(vel), ng_vel, _ptr(out), _stream()),
              "nodalproj_compute_rhs")
        return out

    def solve(self, rhs, phi, rtol, atol):
        info = KsgfMLMGInfo()
        check(load().ksgf_nodalproj_solve(self.h, _ptr(rhs), _ptr(phi), rtol, atol, C.byref(info), _stream()),
              "nodalproj_solve", allow=(0, 1))
        self.info = info.as_dict()
        return self.info

    def phi(self):
        import torch

        out = torch.zeros(self.N[0] * self.N[1] * self.N[2], dtype=torch.float64, device="cuda")
        check(load().ksgf_nodalproj_get_phi(self.h, _ptr(out), _stream()), "nodalproj_get_phi")
        return out

    def close(self):
        if self.h:
            load().ksgf_nodalproj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PoissonBatch:
    """Many independent single-box Poisson problems solved together.

    problems: list of dicts with n (3 ints), dx (3 floats), lobc/hibc (3 ints
    each, KSGF_LO_*) and beta.  solve(rhs_list, phi_list) takes flat float64
    device tensors of n0*n1*n2 each (Fortran order)."""

    def __init__(self, problems, options=None):
        import numpy as np

        L = load()
        self.np = len(problems)
        self.n = [tuple(int(v) for v in p["n"]) for p in problems]
        n = np.ascontiguousarray([p["n"] for p in problems], dtype=np.int32).ravel()
        dx = np.ascontiguousarray([p["dx"] for p in problems], dtype=np.float64).ravel()
        lo = np.ascontiguousarray([p["lobc"] for p in problems], dtype=np.int32).ravel()
        hi = np.ascontiguousarray([p["hibc"] for p in problems], dtype=np.int32).ravel()
        beta = np.ascontiguousarray([p.get("beta", 1.0) for p in problems], dtype=np.float64)
        self.options = options or default_options()
        h = C.c_void_p()
        dp = C.POINTER(C.c_double)
        check(L.ksgf_poisson_batch_create(self.np, n.ctypes.data_as(_ip), dx.ctypes.data_as(dp),
                                          lo.ctypes.data_as(_ip), hi.ctypes.data_as(_ip), beta.ctypes.data_as(dp),
                                          C.byref(self.options), C.byref(h)), "ksgf_poisson_batch_create")
        self.h = h
        self.infos = None

    def solve(self, rhs, phi, rtol, atol, use_graph=True):
        if len(rhs) != self.np or len(phi) != self.np:
            raise ValueError("one rhs and one phi per problem")
        for p in range(self.np):
            m = self.n[p][0] * self.n[p][1] * self.n[p][2]
            for t in (rhs[p], phi[p]):
                if t.numel() != m or not t.is_contiguous() or str(t.dtype) != "torch.float64":
                    raise ValueError(f"problem {p}: contiguous float64 arrays of {m} cells expected")
        r = (C.c_void_p * self.np)(*[_ptr(t) for t in rhs])
        f = (C.c_void_p * self.np)(*[_ptr(t) for t in phi])
        infos = (KsgfMLMGInfo * self.np)()
        check(load().ksgf_poisson_batch_solve(self.h, r, f, rtol, atol, int(use_graph), infos, _stream()),
              "poisson_batch_solve", allow=(0, 1))
        self.infos = [i.as_dict() for i in infos]
        return self.infos

    def close(self):
        if self.h:
            load().ksgf_poisson_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
