"""ctypes binding of the CPU FP64 oracle (liboracle.so, built from
ksgf_oracle.c by the Makefile beside this file).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the bench's
cpu_baseline leg as the checker and the CPU baseline.  The product library
(libksgf.so) never loads it.

Also provides ``advect_project_step``: the Godunov branch of
incflo::ApplyPredictor (incflo_advance.cpp:186-438) restricted to the
advect + MAC-project part (sources, ExtrapVelToFaces, MAC projection, face
ghost fill, momentum fluxes, convective term, predictor update), evaluated
with the oracle functions on numpy MultiFabs.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

CELL, XFACE, YFACE, ZFACE = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class KoGeom(C.Structure):
    _fields_ = [("nbox", C.c_int), ("boxes", _ip), ("dlo", C.c_int * 3), ("dhi", C.c_int * 3),
                ("dx", C.c_double * 3), ("periodic", C.c_int * 3)]


class KoBC(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class KoOpts(C.Structure):
    _fields_ = [("maxiter", C.c_int), ("nu1", C.c_int), ("nu2", C.c_int), ("nuf", C.c_int), ("nub", C.c_int),
                ("bottom_maxiter", C.c_int), ("bottom_rtol", C.c_double), ("bottom_atol", C.c_double),
                ("semicoarsening", C.c_int), ("max_coarsening_level", C.c_int)]


class KoInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("r0", C.c_double), ("rfinal", C.c_double), ("converged", C.c_int),
                ("nlevels", C.c_int), ("bottom_iters", C.c_int), ("bottom_fail", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def build(force=False):
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(os.path.join(HERE, f)) for f in ("ksgf_oracle.c", "ksgf_oracle.h")):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.ko_mf_size.restype = C.c_long
        L.ko_mf_size.argtypes = [C.POINTER(KoGeom), C.c_int, C.c_int, C.c_int]
        L.ko_fill_boundary.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, C.c_int, C.c_int]
        L.ko_fill_physbc.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, C.c_int, C.c_int, C.POINTER(KoBC), _dp]
        L.ko_ppm_predict_state.argtypes = [C.POINTER(KoGeom), C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int,
                                           C.POINTER(KoBC), C.c_double, C.c_int, _dp, _dp]
        L.ko_extrap_vel_to_faces.argtypes = [C.POINTER(KoGeom), _dp, _dp, _dp, _dp, _dp, C.POINTER(KoBC), C.c_double,
                                             C.c_int, C.c_int, C.c_int, C.c_int]
        L.ko_compute_fluxes_from_state.argtypes = [C.POINTER(KoGeom), C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                                   C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, C.POINTER(KoBC), _ip,
                                                   C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ko_compute_divergence.argtypes = [C.POINTER(KoGeom), _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int]
        L.ko_mlmg_default_opts.argtypes = [C.POINTER(KoOpts)]
        L.ko_mac_project.argtypes = [C.POINTER(KoGeom), _dp, _dp, _dp, _ip, _ip, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(KoOpts), C.POINTER(KoInfo), _dp]
        L.ko_poisson_solve.argtypes = [_ip, _dp, _ip, _ip, C.c_double, _dp, _dp, C.c_double, C.c_double,
                                       C.POINTER(KoOpts), C.POINTER(KoInfo)]
        L.ko_vb_solve.argtypes = [_ip, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                  C.POINTER(KoOpts), C.POINTER(KoInfo)]
        L.ko_nodal_dims.argtypes = [_ip, _ip, _ip]
        L.ko_nodal_rhs.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, _ip, _ip, _dp]
        L.ko_nodal_solve.argtypes = [_ip, _dp, _ip, _ip, C.c_double, _dp, _dp, C.c_double, C.c_double,
                                     C.POINTER(KoOpts), C.POINTER(KoInfo)]
        L.ko_nodal_correct.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, _dp, _dp, _ip, C.c_double]
        L.ko_nodal_project.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, _ip, _ip, C.c_double, C.c_double,
                                       C.c_double, C.POINTER(KoOpts), C.POINTER(KoInfo), _dp, _dp]
        L.ko_abeclap_solve.argtypes = [_ip, _dp, _ip, _ip, C.c_double, C.c_double, _dp, _dp, C.c_double, C.c_double,
                                       C.POINTER(KoOpts), C.POINTER(KoInfo)]
        L.ko_abeclap_apply.argtypes = [_ip, _dp, _ip, _ip, C.c_double, C.c_double, _dp, _dp]
        L.ko_tensor_apply.argtypes = [_ip, _dp, _ip, _ip, C.c_double, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.ko_tensor_solve.argtypes = [_ip, _dp, _ip, _ip, C.c_double, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp,
                                      C.c_double, C.c_double, C.c_int, C.POINTER(KoInfo)]
        L.ko_average_down.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, C.POINTER(KoGeom), _dp, C.c_int, C.c_int,
                                      C.c_int]
        L.ko_average_down_faces.argtypes = [C.POINTER(KoGeom), C.POINTER(_dp), C.c_int, C.POINTER(KoGeom),
                                            C.POINTER(_dp), C.c_int, C.c_int, C.c_int]
        L.ko_fillpatch_two_levels.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, C.POINTER(KoGeom), _dp, C.c_int,
                                              C.c_int, C.c_int, C.POINTER(KoBC), _dp]
        L.ko_fillpatch_faces_two_levels.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, C.c_int, C.POINTER(KoGeom), _dp,
                                                    C.c_int, C.c_int, C.POINTER(KoBC), _dp]
        L.ko_tl_apply.argtypes = [_ip, _dp, _ip, _ip, _ip, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.ko_mac_project_two_levels.argtypes = [_ip, _dp, _ip, _ip, _ip, C.c_int, C.c_double, C.POINTER(_dp),
                                                C.POINTER(_dp), C.c_double, C.c_double, C.POINTER(KoOpts),
                                                C.POINTER(KoInfo), _dp, _dp]
        L.ko_cfl_max.restype = C.c_double
        L.ko_cfl_max.argtypes = [C.POINTER(KoGeom), _dp, C.c_int]
        L.ko_force_cfl_max.restype = C.c_double
        L.ko_force_cfl_max.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, _dp, C.c_int]
        L.ko_set_plant.argtypes = [C.c_int]
        L.ko_predictor_update.argtypes = [C.POINTER(KoGeom), _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                          _dp, _dp, _dp, C.c_double, C.c_double, C.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Geom:
    """Keeps the ctypes geometry and the box array alive together."""

    def __init__(self, ba):
        self.ba = ba
        self.boxes = np.ascontiguousarray(ba.boxes, dtype=np.int32)
        g = KoGeom()
        g.nbox = ba.nbox
        g.boxes = self.boxes.ctypes.data_as(_ip)
        for d in range(3):
            g.dlo[d] = ba.geom.dlo[d]
            g.dhi[d] = ba.geom.dhi[d]
            g.dx[d] = ba.geom.dx[d]
            g.periodic[d] = int(ba.geom.periodic[d])
        self.g = g

    @property
    def ref(self):
        return C.byref(self.g)


def bc_structs(bcrec):
    bcrec = np.asarray(bcrec, dtype=np.int32).reshape(-1, 6)
    arr = (KoBC * len(bcrec))()
    for n, row in enumerate(bcrec):
        for d in range(3):
            arr[n].lo[d] = int(row[d])
            arr[n].hi[d] = int(row[3 + d])
    return arr


def default_opts(**kw):
    o = KoOpts()
    lib().ko_mlmg_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# --------------------------------------------------------------------------
# thin wrappers
# --------------------------------------------------------------------------
def fill_boundary(G, mf, itype, ncomp, ngrow):
    lib().ko_fill_boundary(G.ref, _p(mf), itype, ncomp, ngrow)


def fill_physbc(G, mf, itype, ncomp, ngrow, bcrec, bcval):
    bcval = np.ascontiguousarray(bcval, dtype=np.float64)
    lib().ko_fill_physbc(G.ref, _p(mf), itype, ncomp, ngrow, bc_structs(bcrec), _p(bcval))


def fill_ghosts(G, mf, itype, ncomp, ngrow, bcrec, bcval):
    fill_boundary(G, mf, itype, ncomp, ngrow)
    fill_physbc(G, mf, itype, ncomp, ngrow, bcrec, bcval)


def ppm_predict_state(G, d, s, ncomp_s, ng_s, comp, vel_edge, ng_v, bcrec1, dt, limiter):
    n = G.ba.mf_size(CELL, 1, 0)
    Im = np.zeros(n)
    Ip = np.zeros(n)
    lib().ko_ppm_predict_state(G.ref, d, _p(s), ncomp_s, ng_s, comp, _p(vel_edge), ng_v, bc_structs(bcrec1), dt,
                               limiter, _p(Im), _p(Ip))
    return Im, Ip


def extrap_vel_to_faces(G, vel, f, bcrec, dt, use_ppm=1, use_forces_in_trans=0, limiter=1, allow_inflow=0):
    ba = G.ba
    mac = [np.zeros(ba.mf_size(XFACE + d, 1, 1)) for d in range(3)]
    rc = lib().ko_extrap_vel_to_faces(G.ref, _p(vel), _p(f), _p(mac[0]), _p(mac[1]), _p(mac[2]), bc_structs(bcrec),
                                      dt, use_ppm, use_forces_in_trans, limiter, allow_inflow)
    if rc != 0:
        raise RuntimeError(f"ko_extrap_vel_to_faces: {rc}")
    return mac


def compute_fluxes(G, ncomp, q, q_bc, mac, divu, fq, dt, bcrec, iconserv, use_ppm=1, use_forces_in_trans=0,
                   is_velocity=1, area_weighted=0, limiter=1, allow_inflow=0, known_edge_state=0, faces=None):
    ba = G.ba
    flux = [np.zeros(ba.mf_size(XFACE + d, ncomp, 0)) for d in range(3)]
    face = faces if faces is not None else [np.zeros(ba.mf_size(XFACE + d, ncomp, 0)) for d in range(3)]
    ic = np.ascontiguousarray(iconserv, dtype=np.int32)
    rc = lib().ko_compute_fluxes_from_state(G.ref, ncomp, _p(q), _p(q_bc), _p(flux[0]), _p(flux[1]), _p(flux[2]),
                                            _p(face[0]), _p(face[1]), _p(face[2]), known_edge_state, _p(mac[0]),
                                            _p(mac[1]), _p(mac[2]), _p(divu), _p(fq), dt, bc_structs(bcrec),
                                            ic.ctypes.data_as(_ip), use_ppm, use_forces_in_trans, is_velocity,
                                            area_weighted, limiter, allow_inflow)
    if rc != 0:
        raise RuntimeError(f"ko_compute_fluxes_from_state: {rc}")
    return flux, face


def compute_divergence(G, flux, ncomp, mult=-1.0, area_weighted=0):
    div = np.zeros(G.ba.mf_size(CELL, ncomp, 0))
    lib().ko_compute_divergence(G.ref, _p(div), _p(flux[0]), _p(flux[1]), _p(flux[2]), ncomp, mult, area_weighted)
    return div


def mac_project(G, mac, lobc, hibc, beta, rtol, atol, opts=None, want_phi=False):
    opts = opts or default_opts()
    info = KoInfo()
    lo = np.ascontiguousarray(lobc, dtype=np.int32)
    hi = np.ascontiguousarray(hibc, dtype=np.int32)
    phi = np.zeros(int(np.prod(G.ba.geom.n_cell))) if want_phi else None
    st = lib().ko_mac_project(G.ref, _p(mac[0]), _p(mac[1]), _p(mac[2]), lo.ctypes.data_as(_ip),
                              hi.ctypes.data_as(_ip), beta, rtol, atol, C.byref(opts), C.byref(info), _p(phi))
    return dict(info.as_dict(), status=st), phi


def poisson_solve(n, dx, lobc, hibc, beta, rhs, rtol, atol, opts=None):
    opts = opts or default_opts()
    info = KoInfo()
    nn = np.ascontiguousarray(n, dtype=np.int32)
    dd = np.ascontiguousarray(dx, dtype=np.float64)
    lo = np.ascontiguousarray(lobc, dtype=np.int32)
    hi = np.ascontiguousarray(hibc, dtype=np.int32)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).ravel()
    phi = np.zeros_like(rhs)
    st = lib().ko_poisson_solve(nn.ctypes.data_as(_ip), _p(dd), lo.ctypes.data_as(_ip), hi.ctypes.data_as(_ip),
                                beta, _p(rhs), _p(phi), rtol, atol, C.byref(opts), C.byref(info))
    return phi, dict(info.as_dict(), status=st)


def vb_solve(n, dx, lobc, hibc, b3, rhs, rtol, atol, opts=None):
    """div(beta grad phi) = rhs with face coefficients b3 = (bx, by, bz),
    domain-wide (x fastest; bx has n0 + 1 faces along x); returns (phi,
    info)."""
    opts = opts or default_opts()
    info = KoInfo()
    phi = np.zeros(int(np.prod(n)))
    b3 = [np.ascontiguousarray(b, dtype=np.float64).ravel() for b in b3]
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).ravel()
    nn, lo, hi = _iarr(n), _iarr(lobc), _iarr(hibc)
    dd = np.ascontiguousarray(dx, dtype=np.float64)
    rc = lib().ko_vb_solve(nn[1], _p(dd), lo[1], hi[1], _p(b3[0]), _p(b3[1]), _p(b3[2]), _p(rhs), _p(phi), rtol,
                           atol, C.byref(opts), C.byref(info))
    if rc == -2:
        raise ValueError("face-coefficient solve: odd periodic extent")
    return phi, dict(info.as_dict(), status=rc)


def mac_project_vb(ba, mac, lobc, hibc, b3, rtol, atol, opts=None):
    """MacProjector with face coefficients on the domain-wide face velocities
    mac = (u (nz, ny, nx+1), v, w) and b3 of the same shapes: rhs = div u
    (k_mac_rhs's expression), the face-coefficient solve, then
    u -= (beta_face * (1/dx)) * (phi_hi - phi_lo) on every face (periodic wrap;
    Neumann sides unchanged; Dirichlet sides with the ghost -phi).  Returns
    (corrected faces, phi, info)."""
    g = ba.geom
    n = tuple(int(v) for v in g.n_cell)
    dx = g.dx
    u, v, w = (np.asarray(a, dtype=np.float64) for a in mac)
    rhs = (u[:, :, 1:] - u[:, :, :-1]) * (1.0 / dx[0]) + (v[:, 1:, :] - v[:, :-1, :]) * (1.0 / dx[1]) + \
        (w[1:, :, :] - w[:-1, :, :]) * (1.0 / dx[2])
    phi, info = vb_solve(n, dx, lobc, hibc, b3, rhs, rtol, atol, opts)
    phi = phi.reshape(n[2], n[1], n[0])
    out = []
    for d, a in enumerate((u, v, w)):
        ax = 2 - d
        a = a.copy()
        lo_ghost = {0: np.take(phi, [-1], axis=ax), 1: np.take(phi, [0], axis=ax),
                    2: -np.take(phi, [0], axis=ax)}[lobc[d]]
        hi_ghost = {0: np.take(phi, [0], axis=ax), 1: np.take(phi, [-1], axis=ax),
                    2: -np.take(phi, [-1], axis=ax)}[hibc[d]]
        ext = np.concatenate([lo_ghost, phi, hi_ghost], axis=ax)
        sl_hi = [slice(None)] * 3
        sl_lo = [slice(None)] * 3
        sl_hi[ax] = slice(1, None)
        sl_lo[ax] = slice(None, -1)
        bd = np.asarray(b3[d], dtype=np.float64) * (1.0 / dx[d])
        corr = bd * (ext[tuple(sl_hi)] - ext[tuple(sl_lo)])
        if lobc[d] == 1:
            s0 = [slice(None)] * 3
            s0[ax] = 0
            corr[tuple(s0)] = 0.0
        if hibc[d] == 1:
            s1 = [slice(None)] * 3
            s1[ax] = -1
            corr[tuple(s1)] = 0.0
        out.append(a - corr)
    return out, phi, info


def _iarr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def nodal_dims(n, lobc):
    return tuple(n[d] if lobc[d] == 0 else n[d] + 1 for d in range(3))


def nodal_rhs(G, vel, ng, lobc, hibc):
    """nodal divergence of a cell velocity MultiFab (ghosts as ApplyProjection
    leaves them), domain-wide nodes, Fortran order"""
    N = nodal_dims(G.ba.geom.n_cell, lobc)
    rhs = np.zeros(int(np.prod(N)))
    lo, plo = _iarr(lobc)
    hi, phi_ = _iarr(hibc)
    lib().ko_nodal_rhs(G.ref, _p(vel), ng, plo, phi_, _p(rhs))
    return rhs


def nodal_solve(n, dx, lobc, hibc, sigma, rhs, rtol, atol, opts=None):
    opts = opts or default_opts()
    info = KoInfo()
    nn, pn = _iarr(n)
    dd = np.ascontiguousarray(dx, dtype=np.float64)
    lo, plo = _iarr(lobc)
    hi, phi_ = _iarr(hibc)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).ravel()
    phi = np.zeros_like(rhs)
    st = lib().ko_nodal_solve(pn, _p(dd), plo, phi_, sigma, _p(rhs), _p(phi), rtol, atol, C.byref(opts),
                              C.byref(info))
    return phi, dict(info.as_dict(), status=st)


def nodal_project(G, vel, ng, lobc, hibc, sigma, rtol, atol, opts=None):
    """NodalProjector::project with a constant sigma; vel corrected in place.
    Returns (info, phi (domain-wide nodes), gradphi (cell MultiFab, 3 comp, ngrow 0))"""
    opts = opts or default_opts()
    info = KoInfo()
    N = nodal_dims(G.ba.geom.n_cell, lobc)
    phi = np.zeros(int(np.prod(N)))
    gp = np.zeros(G.ba.mf_size(CELL, 3, 0))
    lo, plo = _iarr(lobc)
    hi, phi_ = _iarr(hibc)
    st = lib().ko_nodal_project(G.ref, _p(vel), ng, plo, phi_, sigma, rtol, atol, C.byref(opts), C.byref(info),
                                _p(phi), _p(gp))
    return dict(info.as_dict(), status=st), phi, gp


def abeclap_solve(n, dx, lobc, hibc, alpha_a, beta_b, rhs, rtol, atol, opts=None):
    """alpha_a phi - beta_b lap(phi) = rhs (domain-wide, Fortran order)"""
    opts = opts or default_opts()
    info = KoInfo()
    nn, pn = _iarr(n)
    dd = np.ascontiguousarray(dx, dtype=np.float64)
    lo, plo = _iarr(lobc)
    hi, phi_ = _iarr(hibc)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).ravel()
    phi = np.zeros_like(rhs)
    st = lib().ko_abeclap_solve(pn, _p(dd), plo, phi_, alpha_a, beta_b, _p(rhs), _p(phi), rtol, atol,
                                C.byref(opts), C.byref(info))
    return phi, dict(info.as_dict(), status=st)


def abeclap_apply(n, dx, lobc, hibc, alpha_a, beta_b, x):
    nn, pn = _iarr(n)
    dd = np.ascontiguousarray(dx, dtype=np.float64)
    lo, plo = _iarr(lobc)
    hi, phi_ = _iarr(hibc)
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.zeros_like(x)
    lib().ko_abeclap_apply(pn, _p(dd), plo, phi_, alpha_a, beta_b, _p(x), _p(y))
    return y


def _tensor_args(n, dx, lobc, hibc, a, eta3):
    nn = np.ascontiguousarray(n, dtype=np.int32)
    dd = np.ascontiguousarray(dx, dtype=np.float64)
    lo = np.ascontiguousarray(lobc, dtype=np.int32).ravel()
    hi = np.ascontiguousarray(hibc, dtype=np.int32).ravel()
    assert lo.size == 9 and hi.size == 9
    a = None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
    e = [np.ascontiguousarray(b, dtype=np.float64).ravel() for b in eta3]
    return nn, dd, lo, hi, a, e


def tensor_apply(n, dx, lobc, hibc, alpha, a, beta, eta3, u):
    """MLTensorOp apply: alpha a u - beta div(eta (grad u + grad u^T) -
    (2/3) eta (div u) I).  u: (3, ncell) domain-wide (x fastest); a: cell
    array or None (1); eta3: face viscosities (x faces (nz, ny, nx+1), ...);
    lobc / hibc: (3 components, 3 directions) of KO_LO_* values."""
    nn, dd, lo, hi, a, e = _tensor_args(n, dx, lobc, hibc, a, eta3)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(3, -1)
    out = np.empty_like(u)
    lib().ko_tensor_apply(nn.ctypes.data_as(_ip), _p(dd), lo.ctypes.data_as(_ip), hi.ctypes.data_as(_ip), alpha,
                          _p(a), beta, _p(e[0]), _p(e[1]), _p(e[2]), _p(u), _p(out))
    return out


def tensor_solve(n, dx, lobc, hibc, alpha, a, beta, eta3, rhs, rtol, atol, maxiter=20000):
    """L u = rhs for the MLTensorOp operator (BiCGStab from zero; returns
    (u (3, ncell), info))"""
    nn, dd, lo, hi, a, e = _tensor_args(n, dx, lobc, hibc, a, eta3)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(3, -1)
    sol = np.zeros_like(rhs)
    info = KoInfo()
    st = lib().ko_tensor_solve(nn.ctypes.data_as(_ip), _p(dd), lo.ctypes.data_as(_ip), hi.ctypes.data_as(_ip),
                               alpha, _p(a), beta, _p(e[0]), _p(e[1]), _p(e[2]), _p(rhs), _p(sol), rtol, atol,
                               int(maxiter), C.byref(info))
    return sol, dict(info.as_dict(), status=st)


def velocity_eta_to_faces(cc, periodic):
    """diffusion::average_velocity_eta_to_faces (incflo_diffusion.cpp:108-128)
    on a domain-wide cell array cc (nz, ny, nx): each face the mean
    0.5 (below + above) of its cells (periodic wrap on a periodic side), a
    face on a non-periodic domain side the value of its one cell
    (fixup_eta_on_domain_faces, :130-200)."""
    cc = np.asarray(cc, dtype=np.float64)
    out = []
    for d in range(3):
        ax = 2 - d
        lo = np.take(cc, [-1], axis=ax) if periodic[d] else np.take(cc, [0], axis=ax)
        hi = np.take(cc, [0], axis=ax) if periodic[d] else np.take(cc, [-1], axis=ax)
        below = np.concatenate([lo, cc], axis=ax)
        above = np.concatenate([cc, hi], axis=ax)
        f = 0.5 * (below + above)
        if not periodic[d]:
            sl0 = [slice(None)] * 3
            sl0[ax] = slice(0, 1)
            sl1 = [slice(None)] * 3
            sl1[ax] = slice(-1, None)
            f[tuple(sl0)] = np.take(cc, [0], axis=ax)
            f[tuple(sl1)] = np.take(cc, [-1], axis=ax)
        out.append(np.ascontiguousarray(f))
    return out


def average_down(Gf, fine, ng_f, Gc, crse, ng_c, ncomp, ratio):
    """amrex::average_down into crse (in place)"""
    lib().ko_average_down(Gf.ref, _p(fine), ng_f, Gc.ref, _p(crse), ng_c, ncomp, ratio)


def average_down_faces(Gf, fine3, ng_f, Gc, crse3, ng_c, ncomp, ratio):
    """amrex::average_down_faces into crse3 (three face MultiFabs, in place)"""
    f = (_dp * 3)(*[_p(a) for a in fine3])
    c = (_dp * 3)(*[_p(a) for a in crse3])
    lib().ko_average_down_faces(Gf.ref, f, ng_f, Gc.ref, c, ng_c, ncomp, ratio)


def fillpatch_two_levels(Gf, fine, ng, Gc, crse, ng_c, ncomp, ratio, bcrec, bcval):
    """FillPatchTwoLevels of the fine level's ghost cells (in place)"""
    bcval = np.ascontiguousarray(bcval, dtype=np.float64)
    lib().ko_fillpatch_two_levels(Gf.ref, _p(fine), ng, Gc.ref, _p(crse), ng_c, ncomp, ratio, bc_structs(bcrec),
                                  _p(bcval))


def fillpatch_faces_two_levels(Gf, fine, itype, ng, Gc, crse, ng_c, ncomp, bcrec, bcval):
    """FillPatchTwoLevels of the fine level's ghost faces (in place)"""
    bcval = np.ascontiguousarray(bcval, dtype=np.float64)
    lib().ko_fillpatch_faces_two_levels(Gf.ref, _p(fine), itype, ng, Gc.ref, _p(crse), ng_c, ncomp,
                                        bc_structs(bcrec), _p(bcval))


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def _fboxes(fboxes):
    fb = np.ascontiguousarray(np.asarray(fboxes, dtype=np.int32).reshape(-1, 6))
    return fb, fb.ctypes.data_as(_ip)


def tl_apply(n, dxc, lobc, hibc, fboxes, beta, phic, phif):
    """composite two-level operator div(beta grad phi): (Lc, Lf) domain-wide"""
    fb, fp = _fboxes(fboxes)
    phic = np.ascontiguousarray(phic, dtype=np.float64)
    phif = np.ascontiguousarray(phif, dtype=np.float64)
    Lc, Lf = np.zeros_like(phic), np.zeros_like(phif)
    lib().ko_tl_apply(_i3(n), _d3(dxc), _i3(lobc), _i3(hibc), fp, fb.shape[0], beta, _p(phic), _p(phif), _p(Lc),
                      _p(Lf))
    return Lc, Lf


def mac_project_two_levels(n, dxc, lobc, hibc, fboxes, beta, umc, umf, rtol=1e-11, atol=0.0, opts=None):
    """composite two-level MAC projection; umc/umf: three domain-wide face
    arrays per level, corrected in place.  Returns (info, phic, phif)."""
    fb, fp = _fboxes(fboxes)
    for a in list(umc) + list(umf):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    o = opts if opts is not None else default_opts()
    info = KoInfo()
    nc = int(np.prod(n))
    phic, phif = np.zeros(nc), np.zeros(8 * nc)
    ret = lib().ko_mac_project_two_levels(_i3(n), _d3(dxc), _i3(lobc), _i3(hibc), fp, fb.shape[0], beta,
                                          (_dp * 3)(*[_p(a) for a in umc]), (_dp * 3)(*[_p(a) for a in umf]), rtol,
                                          atol, C.byref(o), C.byref(info), _p(phic), _p(phif))
    return dict(info.as_dict(), status=ret), phic, phif


def set_plant(bits):
    """Plant a mistake in the oracle (0 = none): bit 0 drops the transverse
    terms of ExtrapVelToFaces, bit 1 doubles them, bit 2 flips their sign,
    bit 3 drops the transverse terms of ComputeFluxesOnBoxFromState.  Used
    only by tests that show a check fails on a wrong oracle."""
    lib().ko_set_plant(int(bits))


def force_cfl_max(G, src, ng_src, rho, ng_rho):
    return lib().ko_force_cfl_max(C.byref(G.g), _p(src), ng_src, _p(rho), ng_rho)


DT_GROWTH = 0.1  # time.max_dt_growth default (SimTime.H:248)


def set_current_cfl(cfl, conv_cfl, force_cfl, dt_prev=0.0, dt_growth=DT_GROWTH):
    """SimTime::set_current_cfl (SimTime.cpp:162-205) for an adaptive step
    after initialisation, without a diffusive CFL (Godunov with implicit or
    Crank-Nicolson diffusion) and without output-time clipping:
    dt = 2 cfl / (C + sqrt(C^2 + 4 F)), then at most (1 + dt_growth) dt_prev
    when a previous step exists."""
    cd = conv_cfl + 0.0
    unit = cd + math.sqrt((cd * cd) + (4.0 * force_cfl))
    if unit < np.finfo(float).eps:
        raise ValueError("CFL is below machine epsilon and the time step is adaptive")
    dt = 2.0 * cfl / max(unit, np.finfo(float).eps)
    if dt_prev > 0.0:
        dt = min(dt, (1.0 + dt_growth) * dt_prev)
    return dt


def src_nph_of(ba, gp, rho):
    """The ICNS source term in momentum form, (-gp / rho) rho on valid cells
    (SrcTermOp<ICNS> then multiply_rho, icns_ops.H:98-120, PDEOps.H:94-113):
    what src_term holds when the next step's compute_dt reads it."""
    out = np.zeros(ba.mf_size(CELL, 3, 0))
    for b in range(ba.nbox):
        g_ = ba.fab_view(gp, b, CELL, 3, 0)
        r_ = ba.fab_view(rho, b, CELL, 1, 3)[:, 3:-3, 3:-3, 3:-3]
        ba.fab_view(out, b, CELL, 3, 0)[...] = (-(1.0 * g_) * (1.0 / r_)) * r_
    return out


def compute_dt(case, G, vel, rho, gp, clock=None):
    """incflo::compute_dt + SimTime::set_current_cfl (incflo_compute_dt.cpp:33-220,
    SimTime.cpp:162-205) with use_force_cfl (the default): the convective CFL
    of vel, the forcing CFL of the source term the previous step formed
    (clock["src_nph"]; before a first step, the one of the loaded gp), and
    the dt growth cap against clock["dt_prev"]."""
    clock = {} if clock is None else clock
    src = clock.get("src_nph")
    if src is None:
        src = src_nph_of(case.ba, gp, rho)
    c = cfl_max(G, vel, 3)
    f = force_cfl_max(G, src, 0, rho, 3)
    return set_current_cfl(case.cfl, c, f, clock.get("dt_prev", 0.0))


def cfl_max(G, vel, ng):
    return lib().ko_cfl_max(G.ref, _p(vel), ng)


def predictor_update(G, vel_new, ng_new, vel_old, ng_old, rho_old, ng_ro, rho_new, ng_rn, conv, src, diff, kappa, dt,
                     ncomp=3):
    lib().ko_predictor_update(G.ref, _p(vel_new), ng_new, _p(vel_old), ng_old, _p(rho_old), ng_ro, _p(rho_new), ng_rn,
                              _p(conv), _p(src), _p(diff), kappa, dt, ncomp)


# --------------------------------------------------------------------------
# one advect + MAC-project step (Godunov branch of ApplyPredictor)
# --------------------------------------------------------------------------
def _cellwise(ba, flat, ncomp, ngrow, fn):
    """apply fn(view) to every FAB's (ncomp, nz, ny, nx) view in place"""
    for b in range(ba.nbox):
        fn(ba.fab_view(flat, b, CELL, ncomp, ngrow))


def make_state(case):
    """velocity (ng 3), density (ng 3), gp (ng 0) MultiFabs with ghosts filled"""
    ba = case.ba
    G = Geom(ba)
    vel = ba.scatter(case.vel, CELL, 3)
    fill_ghosts(G, vel, CELL, 3, 3, case.bcrec, case.bcval)
    rho = ba.scatter(case.rho, CELL, 3)
    rbc = np.zeros((1, 6), dtype=np.int32)
    for d in range(3):
        if not case.geom.periodic[d]:
            rbc[0, d] = rbc[0, 3 + d] = 2  # foextrap for density
    fill_ghosts(G, rho, CELL, 1, 3, rbc, np.zeros((6, 1)))
    gp = ba.scatter(case.gp, CELL, 0)
    return G, vel, rho, gp


def case_opts(case):
    kw = {}
    for k, v in case.mlmg.items():
        kw["semicoarsening" if k == "do_semicoarsening" else k] = v
    return default_opts(**kw)


def advect_scalar(case, G, scal, rho, mac, dt, scal_bc, scal_bcval, limiter=1):
    """Godunov advection of a passive scalar by the projected u_mac
    (AdvectionOp<PDE, Godunov>::operator(), AdvOp_Godunov.H:116-300, with
    multiply_rho; ApplyPredictor order incflo_advance.cpp:316-324): q = rho s,
    conservative fluxes (iconserv 1, is_velocity 0, zero forcing and divu),
    conv = -div F, s_new = (rho_o s_o + dt conv) / rho_n (CompRHSOps.H:93-117).
    Returns s_new (ng 3, ghosts filled)."""
    ba = case.ba
    qs = scal.copy()
    for b in range(ba.nbox):
        ba.fab_view(qs, b, CELL, 1, 3)[...] *= ba.fab_view(rho, b, CELL, 1, 3)
    fqs = np.zeros(ba.mf_size(CELL, 1, 1))
    flux, _ = compute_fluxes(G, 1, qs, qs, mac, None, fqs, dt, scal_bc, [1], 1, 0, 0, 0, limiter, 0)
    conv = compute_divergence(G, flux, 1, -1.0, 0)
    s_new = np.zeros_like(scal)
    predictor_update(G, s_new, 3, scal, 3, rho, 3, rho, 3, conv, np.zeros(ba.mf_size(CELL, 1, 0)), None, 0.0, dt,
                     ncomp=1)
    fill_ghosts(G, s_new, CELL, 1, 3, scal_bc, scal_bcval)
    return s_new


def diffusion_bcs(case, c):
    """LinOp BCs of velocity component c for the diffusion solve
    (get_diffuse_tensor_bc, incflo_diffusion.cpp:11-69): int_dir -> periodic,
    ext_dir -> Dirichlet, extrapolation -> Neumann"""
    lo, hi = [], []
    for d in range(3):
        for side, arr in ((0, lo), (1, hi)):
            t = int(case.bcrec[c, 3 * side + d])
            arr.append(0 if t == 0 else (2 if t == 3 else 1))
    return tuple(lo), tuple(hi)


def _diffusion(case, ba, opts, vel, mu, dt_diff, rtol, atol, comp_apply=None, tensor=False):
    """component-wise (rho_0 - dt_diff mu lap) u = rho_0 u (solve) on the
    valid cells of vel (3 comp, ng 3), or with comp_apply the apply
    mu lap(u) (scalars (0, -1)) returned as a (3, ncell) array.  With tensor
    the operator is MLTensorOp's (ICNSDiffTensorOp): a = rho_0, eta = mu on
    every face (average_velocity_eta_to_faces of a constant), one coupled
    solve of the three components (BiCGStab: its iterates are not MLMG's)."""
    n = tuple(int(v) for v in case.geom.n_cell)
    dx = case.geom.dx
    u = ba.gather(vel, CELL, 3, 3)
    if tensor:
        lo = [diffusion_bcs(case, c)[0] for c in range(3)]
        hi = [diffusion_bcs(case, c)[1] for c in range(3)]
        per = tuple(int(case.geom.periodic[d]) for d in range(3))
        eta3 = velocity_eta_to_faces(np.full((n[2], n[1], n[0]), float(mu)), per)
        a = np.full(int(np.prod(n)), float(case.rho0))
        if comp_apply:
            return tensor_apply(n, dx, lo, hi, 0.0, a, -1.0, eta3, u.reshape(3, -1)), []
        x, info = tensor_solve(n, dx, lo, hi, 1.0, a, dt_diff, eta3, (u * case.rho0).reshape(3, -1), rtol, atol)
        return x, [info]
    out = []
    infos = []
    for c in range(3):
        lo, hi = diffusion_bcs(case, c)
        if comp_apply:
            out.append(abeclap_apply(n, dx, lo, hi, 0.0 * case.rho0, -1.0 * mu, u[c].ravel()))
        else:
            rhs = (u[c] * case.rho0).ravel()
            x, info = abeclap_solve(n, dx, lo, hi, 1.0 * case.rho0, dt_diff * mu, rhs, rtol, atol, opts)
            out.append(x)
            infos.append(info)
    return np.stack(out), infos


def two_level_step(case, fba, vel_c, vel_f, gp_c, gp_f, dt=None, opts=None, limiter=1, clock=None):
    """One advect + MAC-project step of two AMR levels (level 1 = fba at
    ratio 2, no subcycling), constant density case.rho0: the Godunov branch of
    ApplyPredictor with every operator looping over levels
    (icns_advection.H:210-217 ExtrapVelToFaces per level, :226 set_inflow,
    :229 MacProjector over both levels, :231-243 face fill-patch and
    FillBoundary, :359-373 fluxes per level, :401-407 fine fluxes averaged
    onto the coarse faces, :418-422 divergence, CompRHSOps.H:93-117 predictor
    per level), then average_down of the new velocity.  vel_c / vel_f: cell
    MultiFabs with 3 ghosts (filled here); gp_*: ngrow 0.  Returns (vel_c
    new, vel_f new, diagnostics).  clock: as for advect_project_step, with
    the source terms of both levels."""
    ba = case.ba
    Gc, Gf = Geom(ba), Geom(fba)
    opts = opts or case_opts(case)
    bcrec = np.asarray(case.bcrec, dtype=np.int32)
    bcval = np.asarray(case.bcval, dtype=np.float64)
    fill_ghosts(Gc, vel_c, CELL, 3, 3, bcrec, bcval)
    fillpatch_two_levels(Gf, vel_f, 3, Gc, vel_c, 3, 3, 2, bcrec, bcval)
    if dt is None:
        clk = {} if clock is None else clock
        c = max(cfl_max(Gc, vel_c, 3), cfl_max(Gf, vel_f, 3))
        srcs = clk.get("src_nph")
        rr = [np.full(b.mf_size(CELL, 1, 3), case.rho0) for b in (ba, fba)]
        if srcs is None:
            srcs = [src_nph_of(ba, gp_c, rr[0]), src_nph_of(fba, gp_f, rr[1])]
        f = max(force_cfl_max(Gc, srcs[0], 0, rr[0], 3), force_cfl_max(Gf, srcs[1], 0, rr[1], 3))
        dt = set_current_cfl(case.cfl, c, f, clk.get("dt_prev", 0.0))
    sbc = bcrec.copy()
    sbc[sbc != 0] = 2
    lev = []
    for G, b, vel, gp in ((Gc, ba, vel_c, gp_c), (Gf, fba, vel_f, gp_f)):
        rho = np.full(b.mf_size(CELL, 1, 3), case.rho0)
        src = np.zeros(b.mf_size(CELL, 3, 1))
        for i in range(b.nbox):
            s_ = b.fab_view(src, i, CELL, 3, 1)
            g_ = b.fab_view(gp, i, CELL, 3, 0)
            rv = b.fab_view(rho, i, CELL, 1, 3)[0, 3:-3, 3:-3, 3:-3]
            s_[:, 1:-1, 1:-1, 1:-1] = -(1.0 * g_) * (1.0 / rv)
        lev.append(dict(G=G, ba=b, vel=vel, gp=gp, rho=rho, src=src))
    fill_ghosts(Gc, lev[0]["src"], CELL, 3, 1, sbc, np.zeros((6, 3)))
    fillpatch_two_levels(Gf, lev[1]["src"], 1, Gc, lev[0]["src"], 1, 3, 2, sbc, np.zeros((6, 3)))
    for L in lev:
        L["mac"] = extrap_vel_to_faces(L["G"], L["vel"], L["src"], bcrec, dt, 1, 0, limiter, 0)
        _set_inflow(case, L["ba"], L["mac"])
    # composite projection on domain-wide face arrays
    nc = case.geom.n_cell
    umc = [np.ascontiguousarray(ba.gather(lev[0]["mac"][d], XFACE + d, 1, 1)[0]).ravel() for d in range(3)]
    umf = [np.ascontiguousarray(fba.gather(lev[1]["mac"][d], XFACE + d, 1, 1)[0]).ravel() for d in range(3)]
    info, _, _ = mac_project_two_levels(nc, case.geom.dx, case.lobc, case.hibc, fba.boxes, 1.0 / case.rho0, umc, umf,
                                        case.mg_rtol, case.mg_atol, opts)
    nf = fba.geom.n_cell
    for d in range(3):
        sc = [1, nc[2], nc[1], nc[0]]
        sc[3 - d] += 1
        sf = [1, nf[2], nf[1], nf[0]]
        sf[3 - d] += 1
        ba.scatter(umc[d].reshape(sc), XFACE + d, 1, out=lev[0]["mac"][d])
        fba.scatter(umf[d].reshape(sf), XFACE + d, 1, out=lev[1]["mac"][d])
    for d in range(3):
        bc1, bv1 = bcrec[d:d + 1], np.ascontiguousarray(bcval[:, d:d + 1])
        fill_boundary(Gc, lev[0]["mac"][d], XFACE + d, 1, 1)
        fill_physbc(Gc, lev[0]["mac"][d], XFACE + d, 1, 1, bc1, bv1)
        fillpatch_faces_two_levels(Gf, lev[1]["mac"][d], XFACE + d, 1, Gc, lev[0]["mac"][d], 1, 1, bc1, bv1)
    for L in lev:
        b = L["ba"]
        q = L["vel"].copy()
        fq = L["src"].copy()
        for i in range(b.nbox):
            r3 = b.fab_view(L["rho"], i, CELL, 1, 3)
            b.fab_view(q, i, CELL, 3, 3)[...] *= r3
            b.fab_view(fq, i, CELL, 3, 1)[...] *= r3[:, 2:-2, 2:-2, 2:-2]
        L["flux"], _ = compute_fluxes(L["G"], 3, q, q.copy(), L["mac"], None, fq, dt, bcrec, [1, 1, 1], 1, 0, 1, 0,
                                      limiter, 0)
    # conservation across the coarse-fine boundary: the fine fluxes onto the
    # coarse faces before the divergence
    average_down_faces(Gf, lev[1]["flux"], 0, Gc, lev[0]["flux"], 0, 3, 2)
    out = []
    for L in lev:
        b = L["ba"]
        conv = compute_divergence(L["G"], L["flux"], 3, -1.0, 0)
        src_nph = src_nph_of(b, L["gp"], L["rho"])
        L["src_nph"] = src_nph
        vn = np.zeros_like(L["vel"])
        predictor_update(L["G"], vn, 3, L["vel"], 3, L["rho"], 3, L["rho"], 3, conv, src_nph, None, 0.0, dt)
        out.append(vn)
    if clock is not None:
        clock["dt_prev"] = dt
        clock["src_nph"] = [L["src_nph"] for L in lev]
    average_down(Gf, out[1], 3, Gc, out[0], 3, 3, 2)
    fill_ghosts(Gc, out[0], CELL, 3, 3, bcrec, bcval)
    fillpatch_two_levels(Gf, out[1], 3, Gc, out[0], 3, 3, 2, bcrec, bcval)
    return out[0], out[1], dict(info, dt=dt)


def _set_inflow(case, ba, mac):
    """set_inflow_velocity (icns_advection.cpp:157-183): the Dirichlet value on
    ext_dir normal faces of the domain"""
    dhi = [int(n) - 1 for n in ba.geom.n_cell]
    for d in range(3):
        if case.geom.periodic[d]:
            continue
        for side in (0, 1):
            if case.bcrec[d, 3 * side + d] != 3:
                continue
            val = case.bcval[3 * side + d, d]
            for b in range(ba.nbox):
                v = ba.fab_view(mac[d], b, XFACE + d, 1, 1)
                lo, hi = ba.boxes[b, d], ba.boxes[b, 3 + d] + 1
                x = 0 if side == 0 else dhi[d] + 1
                if lo <= x <= hi:
                    idx = [slice(1, -1)] * 3
                    idx[2 - d] = x - lo + 1
                    v[(0, *idx)] = val


def advect_project_step(case, G, vel, rho, gp, dt=None, opts=None, limiter=1, scal=None, scal_bc=None,
                        scal_bcval=None, diffusion=None, clock=None):
    """One step; returns (vel_new (ng 3, ghosts filled), diagnostics dict).
    With scal (cell, 1 comp, ng 3) the passive scalar is advected too; the
    new scalar is diag["scal"].  diffusion = dict(mu, type (1 Crank-Nicolson,
    2 Implicit), rtol, atol, tensor (default True: MLTensorOp, the reference's
    use_tensor_operator default; False: one MLABecLaplacian per component)):
    the viscous term with constant mu and density rho_0 as ApplyPredictor
    carries it for Godunov (incflo_advance.cpp:246-291, 380-389).  clock: a dict carried from step to step (dt_prev, src_nph)
    for compute_dt's growth cap and forcing CFL; updated in place."""
    ba = case.ba
    opts = opts or case_opts(case)
    if np.any(ba.gather(rho, CELL, 1, 3) != case.rho0):
        # the step projects with beta = 1/rho_0 (MacProjOp::init_projector(Real),
        # icns_advection.cpp:143-154); the reference's variable-density branch
        # (face beta 1/rho, :234-291) is not restated
        raise ValueError("advect_project_step: density differs from rho_0")
    if dt is None:
        dt = compute_dt(case, G, vel, rho, gp, clock)
    cn = diffusion is not None and diffusion["type"] == 1
    divtau = None
    if cn:
        tau, _ = _diffusion(case, ba, opts, vel, diffusion["mu"], 0.0, 0.0, 0.0, comp_apply=True,
                            tensor=diffusion.get("tensor", True))
        n = case.geom.n_cell
        divtau = ba.scatter(tau.reshape(3, n[2], n[1], n[0]), CELL, 0)
    # sources: src = -gp/rho (velocity form), ghost 1 (foextrap-like fill)
    src = np.zeros(ba.mf_size(CELL, 3, 1))
    for b in range(ba.nbox):
        s = ba.fab_view(src, b, CELL, 3, 1)
        g_ = ba.fab_view(gp, b, CELL, 3, 0)
        r_ = ba.fab_view(rho, b, CELL, 1, 3)[:, 2:-2, 2:-2, 2:-2]
        rv = r_[0, 1:-1, 1:-1, 1:-1]
        if cn:
            s[:, 1:-1, 1:-1, 1:-1] = (-(1.0 * g_) + ba.fab_view(divtau, b, CELL, 3, 0)) * (1.0 / rv)
        else:
            s[:, 1:-1, 1:-1, 1:-1] = -(1.0 * g_) * (1.0 / rv)
    sbc = np.array(case.bcrec)
    sbc[sbc != 0] = 2
    fill_ghosts(G, src, CELL, 3, 1, sbc, np.zeros((6, 3)))
    # face velocities
    mac = extrap_vel_to_faces(G, vel, src, case.bcrec, dt, 1, 0, limiter, 0)
    # set_inflow_velocity on ext_dir normal faces
    for d in range(3):
        if case.geom.periodic[d]:
            continue
        for side in (0, 1):
            if case.bcrec[d, 3 * side + d] != 3:
                continue
            val = case.bcval[3 * side + d, d]
            for b in range(ba.nbox):
                v = ba.fab_view(mac[d], b, XFACE + d, 1, 1)
                lo, hi = ba.boxes[b, d], ba.boxes[b, 3 + d] + 1
                x = 0 if side == 0 else case.geom.dhi[d] + 1
                if lo <= x <= hi:
                    idx = [slice(1, -1)] * 3
                    idx[2 - d] = x - lo + 1
                    v[(0, *idx)] = val
    info, phi = mac_project(G, mac, case.lobc, case.hibc, 1.0 / case.rho0, case.mg_rtol, case.mg_atol, opts,
                            want_phi=True)
    for d in range(3):
        bc1 = case.bcrec[d:d + 1]
        fill_boundary(G, mac[d], XFACE + d, 1, 1)
        fill_physbc(G, mac[d], XFACE + d, 1, 1, bc1, np.ascontiguousarray(case.bcval[:, d:d + 1]))
    scal_new = None
    if scal is not None:
        scal_new = advect_scalar(case, G, scal, rho, mac, dt, scal_bc, scal_bcval, limiter)
    # momentum and forces
    q = vel.copy()
    fq = src.copy()
    for b in range(ba.nbox):
        r3 = ba.fab_view(rho, b, CELL, 1, 3)
        ba.fab_view(q, b, CELL, 3, 3)[...] *= r3
        ba.fab_view(fq, b, CELL, 3, 1)[...] *= r3[:, 2:-2, 2:-2, 2:-2]
    q_nph = q.copy()
    flux, face = compute_fluxes(G, 3, q, q_nph, mac, None, fq, dt, case.bcrec, [1, 1, 1], 1, 0, 1, 0, limiter, 0)
    conv = compute_divergence(G, flux, 3, -1.0, 0)
    # src at n+1/2 in momentum form: -gp (constant density)
    src_nph = src_nph_of(ba, gp, rho)
    if clock is not None:
        clock["dt_prev"] = dt
        clock["src_nph"] = src_nph
    vel_new = np.zeros_like(vel)
    predictor_update(G, vel_new, 3, vel, 3, rho, 3, rho, 3, conv, src_nph, divtau, 0.5 if cn else 0.0, dt)
    dinfo = None
    if diffusion is not None:
        dt_diff = 0.5 * dt if cn else dt
        x, dinfo = _diffusion(case, ba, opts, vel_new, diffusion["mu"], dt_diff, diffusion["rtol"], diffusion["atol"],
                              tensor=diffusion.get("tensor", True))
        n = case.geom.n_cell
        ba.scatter(x.reshape(3, n[2], n[1], n[0]), CELL, 3, out=vel_new)
    fill_ghosts(G, vel_new, CELL, 3, 3, case.bcrec, case.bcval)
    return vel_new, {"dt": dt, "mlmg": info, "phi": phi, "mac": mac, "conv": conv, "flux": flux, "face": face,
                     "scal": scal_new, "diffusion": dinfo}


def apply_projection(case, G, vel, rho, gp, dt, rtol, atol, opts=None):
    """incflo::ApplyProjection, constant density, not incremental
    (incflo_apply_nodal_projection.cpp:149-465): vel += gp * (dt / rho) on
    valid cells, vel.setBndry(0) + FillBoundary, nodal projection with
    sigma = dt / rho_0, gp = grad phi, velocity ghosts refilled.  vel (ng 3)
    and gp (ng 0) are updated in place; returns (info, phi)."""
    ba = case.ba
    opts = opts or case_opts(case)
    for b in range(ba.nbox):
        v = ba.fab_view(vel, b, CELL, 3, 3)
        g = ba.fab_view(gp, b, CELL, 3, 0)
        r = ba.fab_view(rho, b, CELL, 1, 3)
        vv = v[:, 3:-3, 3:-3, 3:-3]
        vv += g * (dt / r[:, 3:-3, 3:-3, 3:-3])
        inner = vv.copy()
        v[...] = 0.0
        v[:, 3:-3, 3:-3, 3:-3] = inner
    fill_boundary(G, vel, CELL, 3, 3)
    info, phi, gpn = nodal_project(G, vel, 3, case.lobc, case.hibc, dt / case.rho0, rtol, atol, opts)
    gp[:] = gpn
    fill_ghosts(G, vel, CELL, 3, 3, case.bcrec, case.bcval)
    return info, phi
