"""ctypes binding of libksgf_host.so: the C++ GodunovPredictor (the Godunov
branch of incflo::ApplyPredictor without diffusion, optionally followed by
incflo::ApplyProjection) driven from Python.  Work goes on torch's current
stream."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import HERE, KsgfBCRec, KsgfError, KsgfGeometry, KsgfMLMGInfo, KsgfMLMGOptions, bcrecs, default_options, load

HOSTLIB = os.path.join(HERE, "libksgf_host.so")
_h = None


def host_lib():
    global _h
    if _h is None:
        load()  # libksgf.so first (one HIP runtime, shared with torch)
        if not os.path.exists(HOSTLIB):
            raise KsgfError(f"{HOSTLIB} is not built: run python kynema-sgf_amd/build.py")
        L = C.CDLL(HOSTLIB)
        L.ksgf_host_last_error.restype = C.c_char_p
        L.ksgf_host_stepper_create.argtypes = [C.POINTER(KsgfGeometry), C.POINTER(C.c_int), C.c_int,
                                               C.POINTER(KsgfBCRec), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int), C.c_double, C.c_double, C.c_double, C.c_double,
                                               C.POINTER(KsgfMLMGOptions), C.POINTER(C.c_void_p)]
        L.ksgf_host_stepper_set_stream.argtypes = [C.c_void_p, C.c_void_p]
        L.ksgf_host_two_level_create.argtypes = [C.POINTER(KsgfGeometry), C.POINTER(C.c_int), C.c_int,
                                                 C.POINTER(C.c_int), C.c_int, C.POINTER(KsgfBCRec),
                                                 C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                 C.c_double, C.c_double, C.c_double, C.c_double,
                                                 C.POINTER(KsgfMLMGOptions), C.c_int, C.POINTER(C.c_void_p)]
        L.ksgf_host_two_level_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                    C.c_int]
        L.ksgf_host_two_level_step.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double),
                                               C.POINTER(KsgfMLMGInfo)]
        L.ksgf_host_two_level_get_velocity.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ksgf_host_two_level_destroy.argtypes = [C.c_void_p]
        L.ksgf_host_stepper_set_state.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                  C.POINTER(C.c_double)]
        L.ksgf_host_stepper_step.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(KsgfMLMGInfo),
                                             C.POINTER(C.c_double)]
        L.ksgf_host_stepper_get_velocity.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ksgf_host_stepper_enable_graph.argtypes = [C.c_void_p, C.c_int]
        L.ksgf_host_stepper_get_phi.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ksgf_host_stepper_enable_scalar.argtypes = [C.c_void_p, C.POINTER(KsgfBCRec), C.POINTER(C.c_double)]
        L.ksgf_host_stepper_set_scalar.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ksgf_host_stepper_get_scalar.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ksgf_host_stepper_write_plotfile.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.c_int]
        L.ksgf_host_stepper_enable_nodal_projection.argtypes = [C.c_void_p, C.c_double, C.c_double,
                                                                C.POINTER(KsgfMLMGOptions)]
        L.ksgf_host_stepper_nodal_info.argtypes = [C.c_void_p, C.POINTER(KsgfMLMGInfo)]
        L.ksgf_host_stepper_get_gradp.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ksgf_host_stepper_get_nodal_phi.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ksgf_host_stepper_enable_diffusion.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double,
                                                         C.POINTER(KsgfMLMGOptions), C.c_int]
        L.ksgf_host_stepper_diffusion_info.argtypes = [C.c_void_p, C.POINTER(KsgfMLMGInfo)]
        L.ksgf_host_stepper_destroy.argtypes = [C.c_void_p]
        L.ksgf_host_mac_project_density.argtypes = [C.POINTER(KsgfGeometry), C.POINTER(C.c_int), C.c_int,
                                                    C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                                    C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                    C.POINTER(C.c_double), C.c_double, C.c_double,
                                                    C.POINTER(KsgfMLMGOptions), C.POINTER(KsgfMLMGInfo),
                                                    C.POINTER(C.c_double)]
        _h = L
    return _h


def _chk(rc, what, allow=(0,)):
    if rc not in allow:
        raise KsgfError(f"{what} failed ({rc}): {host_lib().ksgf_host_last_error().decode()}", rc)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class HostStepper:
    def __init__(self, case, options=None):
        import torch

        L = host_lib()
        self.case = case
        g = KsgfGeometry()
        for d in range(3):
            g.domain_lo[d] = case.geom.dlo[d]
            g.domain_hi[d] = case.geom.dhi[d]
            g.prob_lo[d] = case.geom.prob_lo[d]
            g.prob_hi[d] = case.geom.prob_hi[d]
            g.is_periodic[d] = int(case.geom.periodic[d])
        self._boxes = np.ascontiguousarray(case.ba.boxes, dtype=np.int32)
        self._bcval = np.ascontiguousarray(case.bcval, dtype=np.float64)
        self._lo = np.ascontiguousarray(case.lobc, dtype=np.int32)
        self._hi = np.ascontiguousarray(case.hibc, dtype=np.int32)
        self.options = options or default_options(**case.mlmg)
        h = C.c_void_p()
        _chk(L.ksgf_host_stepper_create(C.byref(g), self._boxes.ctypes.data_as(C.POINTER(C.c_int)), case.ba.nbox,
                                        bcrecs(case.bcrec), _dp(self._bcval), self._lo.ctypes.data_as(C.POINTER(C.c_int)),
                                        self._hi.ctypes.data_as(C.POINTER(C.c_int)), case.rho0, case.cfl, case.mg_rtol,
                                        case.mg_atol, C.byref(self.options), C.byref(h)), "stepper_create")
        self.h = h
        _chk(L.ksgf_host_stepper_set_stream(h, C.c_void_p(torch.cuda.current_stream().cuda_stream)), "set_stream")
        self.last_info = None
        self.last_dt = None

    def load(self, vel=None, rho=None, gp=None):
        c = self.case
        v = np.ascontiguousarray(c.vel if vel is None else vel, dtype=np.float64)
        r = np.ascontiguousarray(c.rho if rho is None else rho, dtype=np.float64)
        g = np.ascontiguousarray(c.gp if gp is None else gp, dtype=np.float64)
        _chk(host_lib().ksgf_host_stepper_set_state(self.h, _dp(v), _dp(r), _dp(g)), "set_state")

    def enable_graph(self, on=True):
        """replay fixed-dt steps (step(dt) with dt > 0) from hipGraphs"""
        _chk(host_lib().ksgf_host_stepper_enable_graph(self.h, int(bool(on))), "enable_graph")

    def step(self, dt=-1.0, check_divergence=False):
        info = KsgfMLMGInfo()
        dtu = C.c_double()
        dv = C.c_double(-1.0)
        _chk(host_lib().ksgf_host_stepper_step(self.h, dt if dt is not None else -1.0, C.byref(dtu), C.byref(info),
                                               C.byref(dv) if check_divergence else None), "step", allow=(0, 1))
        self.last_info = info.as_dict()
        self.last_dt = dtu.value
        self.last_divmax = dv.value if check_divergence else None
        return dtu.value

    def enable_scalar(self, bc=None, bcval=None):
        """a passive scalar advected with the step (1 BCRec, 6 values)"""
        bc = np.zeros((1, 6), dtype=np.int32) if bc is None else np.asarray(bc, dtype=np.int32).reshape(1, 6)
        self._sbcval = np.zeros(6) if bcval is None else np.ascontiguousarray(bcval, dtype=np.float64).reshape(6)
        _chk(host_lib().ksgf_host_stepper_enable_scalar(self.h, bcrecs(bc), _dp(self._sbcval)), "enable_scalar")

    def load_scalar(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64)
        _chk(host_lib().ksgf_host_stepper_set_scalar(self.h, _dp(a)), "set_scalar")

    def scalar(self):
        n = self.case.geom.n_cell
        out = np.zeros((1, n[2], n[1], n[0]))
        _chk(host_lib().ksgf_host_stepper_get_scalar(self.h, _dp(out)), "get_scalar")
        return out

    def velocity(self):
        n = self.case.geom.n_cell
        out = np.zeros((3, n[2], n[1], n[0]))
        _chk(host_lib().ksgf_host_stepper_get_velocity(self.h, _dp(out)), "get_velocity")
        return out

    def phi(self):
        """the MAC projection's phi of the last step, (n2, n1, n0)"""
        n = self.case.geom.n_cell
        out = np.zeros((n[2], n[1], n[0]))
        _chk(host_lib().ksgf_host_stepper_get_phi(self.h, _dp(out)), "get_phi")
        return out

    def enable_nodal_projection(self, rtol, atol, options=None):
        """end every step with incflo::ApplyProjection (constant density)"""
        self._nopts = options or default_options(**self.case.mlmg)
        _chk(host_lib().ksgf_host_stepper_enable_nodal_projection(self.h, rtol, atol, C.byref(self._nopts)),
             "enable_nodal_projection")
        self.nodal = True

    def enable_diffusion(self, mu, diff_type, rtol, atol, options=None, tensor=True):
        """viscous term, constant mu: diff_type 1 Crank-Nicolson, 2 Implicit;
        tensor: MLTensorOp (the reference's default) or one MLABecLaplacian per
        component"""
        self._dopts = options or default_options(**self.case.mlmg)
        _chk(host_lib().ksgf_host_stepper_enable_diffusion(self.h, mu, diff_type, rtol, atol, C.byref(self._dopts),
                                                           int(bool(tensor))),
             "enable_diffusion")

    def diffusion_info(self):
        infos = (KsgfMLMGInfo * 3)()
        _chk(host_lib().ksgf_host_stepper_diffusion_info(self.h, infos), "diffusion_info")
        return [i.as_dict() for i in infos]

    def nodal_info(self):
        info = KsgfMLMGInfo()
        _chk(host_lib().ksgf_host_stepper_nodal_info(self.h, C.byref(info)), "nodal_info")
        return info.as_dict()

    def gradp(self):
        n = self.case.geom.n_cell
        out = np.zeros((3, n[2], n[1], n[0]))
        _chk(host_lib().ksgf_host_stepper_get_gradp(self.h, _dp(out)), "get_gradp")
        return out

    def nodal_phi(self):
        """the last nodal projection's phi, domain-wide nodes (flat, x fastest)"""
        n = self.case.geom.n_cell
        N = [n[d] if self.case.lobc[d] == 0 else n[d] + 1 for d in range(3)]
        out = np.zeros(N[0] * N[1] * N[2])
        _chk(host_lib().ksgf_host_stepper_get_nodal_phi(self.h, _dp(out)), "get_nodal_phi")
        return out

    def write_plotfile(self, path, time=0.0, step=0):
        """plotfile of velocity, gp, density (and the scalar when enabled)"""
        _chk(host_lib().ksgf_host_stepper_write_plotfile(self.h, str(path).encode(), time, step), "write_plotfile")

    def close(self):
        if self.h:
            host_lib().ksgf_host_stepper_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mac_project_density(ba, lobc, hibc, rho, mac, factor=1.0, rtol=1e-11, atol=0.0, options=None):
    """MacProjOp's variable-density projection through the C++ Hydro layer
    (face beta = factor / rho_face): rho (nz, ny, nx), mac = domain-wide face
    arrays (corrected copies are returned with phi and the MLMG info)"""
    L = host_lib()
    g = KsgfGeometry()
    for d in range(3):
        g.domain_lo[d] = ba.geom.dlo[d]
        g.domain_hi[d] = ba.geom.dhi[d]
        g.prob_lo[d] = ba.geom.prob_lo[d]
        g.prob_hi[d] = ba.geom.prob_hi[d]
        g.is_periodic[d] = int(ba.geom.periodic[d])
    boxes = np.ascontiguousarray(ba.boxes, dtype=np.int32)
    lo = np.ascontiguousarray(lobc, dtype=np.int32)
    hi = np.ascontiguousarray(hibc, dtype=np.int32)
    r = np.ascontiguousarray(rho, dtype=np.float64)
    out = [np.ascontiguousarray(m, dtype=np.float64).copy() for m in mac]
    phi = np.zeros(r.size)
    info = KsgfMLMGInfo()
    o = options or default_options()
    _chk(L.ksgf_host_mac_project_density(C.byref(g), boxes.ctypes.data_as(C.POINTER(C.c_int)), ba.nbox,
                                         lo.ctypes.data_as(C.POINTER(C.c_int)), hi.ctypes.data_as(C.POINTER(C.c_int)),
                                         _dp(r), factor, _dp(out[0]), _dp(out[1]), _dp(out[2]), rtol, atol,
                                         C.byref(o), C.byref(info), _dp(phi)), "mac_project_density")
    return out, phi, info.as_dict()


class HostTwoLevel:
    """The C++ two-level step driver (ksgf::amr::TwoLevelPredictor) through
    libksgf_host.so: case is the level-0 workload, fba the level-1 BoxArray
    (boxes aligned to 2).  States are host arrays in each level's FAB layout
    (velocity with 3 ghost layers, gp with none)."""

    def __init__(self, case, fba, options=None, limiter=1):
        L = host_lib()
        self.case = case
        self.fba = fba
        g = KsgfGeometry()
        for d in range(3):
            g.domain_lo[d] = case.geom.dlo[d]
            g.domain_hi[d] = case.geom.dhi[d]
            g.prob_lo[d] = case.geom.prob_lo[d]
            g.prob_hi[d] = case.geom.prob_hi[d]
            g.is_periodic[d] = int(case.geom.periodic[d])
        self._b0 = np.ascontiguousarray(case.ba.boxes, dtype=np.int32)
        self._b1 = np.ascontiguousarray(fba.boxes, dtype=np.int32)
        self._bcval = np.ascontiguousarray(case.bcval, dtype=np.float64)
        self._lo = np.ascontiguousarray(case.lobc, dtype=np.int32)
        self._hi = np.ascontiguousarray(case.hibc, dtype=np.int32)
        self.options = options or default_options(**case.mlmg)
        ip = C.POINTER(C.c_int)
        h = C.c_void_p()
        _chk(L.ksgf_host_two_level_create(C.byref(g), self._b0.ctypes.data_as(ip), case.ba.nbox,
                                          self._b1.ctypes.data_as(ip), fba.nbox, bcrecs(case.bcrec), _dp(self._bcval),
                                          self._lo.ctypes.data_as(ip), self._hi.ctypes.data_as(ip), case.rho0,
                                          case.cfl, case.mg_rtol, case.mg_atol, C.byref(self.options), int(limiter),
                                          C.byref(h)), "two_level_create")
        self.h = h
        self.last_info = None
        self.last_dt = None
        self._n = [case.ba.mf_size(0, 3, 3), fba.mf_size(0, 3, 3)]

    def load(self, vel_c, vel_f, gp_c=None, gp_f=None):
        """host arrays, or device tensors (all of one kind) in the FAB layout"""
        import torch

        dev = torch.is_tensor(vel_c)
        keep = []

        def ptr(x):
            if x is None:
                return None
            if dev:
                x = x.contiguous()
                keep.append(x)
                return C.c_void_p(x.data_ptr())
            x = np.ascontiguousarray(x, dtype=np.float64)
            keep.append(x)
            return C.c_void_p(x.ctypes.data)

        if dev:
            torch.cuda.synchronize()
        _chk(host_lib().ksgf_host_two_level_set_state(self.h, ptr(vel_c), ptr(vel_f), ptr(gp_c), ptr(gp_f), int(dev)),
             "two_level_set_state")

    def step(self, dt=None):
        info = KsgfMLMGInfo()
        dtu = C.c_double()
        _chk(host_lib().ksgf_host_two_level_step(self.h, -1.0 if dt is None else dt, C.byref(dtu), C.byref(info)),
             "two_level_step", allow=(0, 1))
        self.last_info = info.as_dict()
        self.last_dt = dtu.value
        return self.last_dt

    def velocity(self):
        """(level 0, level 1) velocity in the FAB layout, 3 ghost layers"""
        v0, v1 = np.zeros(self._n[0]), np.zeros(self._n[1])
        _chk(host_lib().ksgf_host_two_level_get_velocity(self.h, _dp(v0), _dp(v1)), "two_level_get_velocity")
        return v0, v1

    def close(self):
        if self.h:
            host_lib().ksgf_host_two_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
