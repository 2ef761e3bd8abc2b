"""One advect + MAC-project step of two AMR levels on the GPU, driven through
the C ABI.

The Godunov branch of incflo::ApplyPredictor with every operator looping over
the levels (level 1 at ratio 2, one dt for both levels, constant density):

  compute_dt: convective and forcing CFL over both levels, dt growth cap
                                                   incflo_compute_dt.cpp:33-220, SimTime.cpp:162-205
  src = -gp/rho; level-1 ghosts by fill-patch      icns_ops.H:98-111
  ExtrapVelToFaces per level                       icns_advection.H:210-217
  set_inflow_velocity                              icns_advection.cpp:157-183
  MacProjector over both levels (composite)        icns_advection.cpp:88-95,331,353
  face ghosts: fillpatch_sibling_fields + FillBoundary   icns_advection.H:231-243
  fluxes per level                                 icns_advection.H:359-373
  fine fluxes averaged onto the coarse faces       icns_advection.H:401-407
  conv = -div F per level                          icns_advection.H:418-422
  predictor per level                              CompRHSOps.H:93-117
  average_down of the new velocity                 incflo_apply_nodal_projection.cpp:457-460 (as for gp)

Level-1 ghost cells come from ksgf_fillpatch_two_levels (conservative linear
interpolation), its ghost faces from ksgf_fillpatch_faces_two_levels.

The C++ host layer's ksgf::amr::TwoLevelPredictor (host/incflo_two_level.H)
makes the same calls in the same order; the two give the same bits.
"""
from __future__ import annotations

import numpy as np

from . import CELL, XFACE, DeviceBoxArray, MacProjectorLevels, default_options, set_current_cfl


class _Level:
    def __init__(self, dba, rho0):
        self.dba = dba
        self.vel = dba.alloc(CELL, 3, 3)
        self.vel_new = dba.alloc(CELL, 3, 3)
        self.rho = dba.alloc(CELL, 1, 3)
        self.rho.fill_(rho0)
        self.gp = dba.alloc(CELL, 3, 0)
        self.gp.zero_()
        self.src = dba.alloc(CELL, 3, 1)
        self.src_nph = dba.alloc(CELL, 3, 0)
        self.src_nph.zero_()
        self.q = dba.alloc(CELL, 3, 3)
        self.fq = dba.alloc(CELL, 3, 1)
        self.mac = [dba.alloc(XFACE + d, 1, 1) for d in range(3)]
        self.flux = [dba.alloc(XFACE + d, 3, 0) for d in range(3)]
        self.face = [dba.alloc(XFACE + d, 3, 0) for d in range(3)]
        self.conv = dba.alloc(CELL, 3, 0)


class TwoLevelStep:
    """case: the level-0 workload (BoxArray, BCs, rho0, CFL, MAC tolerances);
    fba: the level-1 BoxArray (layout.LevelBoxArray, boxes aligned to 2)."""

    def __init__(self, case, fba, options=None, limiter=1):
        self.case = case
        self.limiter = limiter
        self.dc = DeviceBoxArray(case.ba)
        self.df = DeviceBoxArray(fba, covers_domain=False)
        self.fba = fba
        self.lev = [_Level(self.dc, case.rho0), _Level(self.df, case.rho0)]
        self.bcrec = np.asarray(case.bcrec, dtype=np.int32)
        self.bcval = np.asarray(case.bcval, dtype=np.float64)
        self.src_bc = self.bcrec.copy()
        self.src_bc[self.src_bc != 0] = 2  # forcing ghosts: first-order extrapolation
        self.options = options or default_options(**case.mlmg)
        self.mp = MacProjectorLevels(self.dc, self.df, case.lobc, case.hibc, 1.0 / case.rho0, self.options)
        self.last_info = None
        self.last_dt = None
        self.dt_prev = 0.0

    def load(self, vel_c, vel_f, gp_c=None, gp_f=None):
        """Set the levels' velocity MultiFabs (3 ghosts) and pressure
        gradients (ngrow 0) from host arrays in the FAB layout or device
        tensors of the same size."""
        import torch

        L0, L1 = self.lev

        def put(dst, src, dba):
            dst.copy_(src if torch.is_tensor(src) else dba.upload(src))

        put(L0.vel, vel_c, self.dc)
        put(L1.vel, vel_f, self.df)
        if gp_c is not None:
            put(L0.gp, gp_c, self.dc)
        if gp_f is not None:
            put(L1.gp, gp_f, self.df)
        # the forcing compute_dt reads before the first step: the source term
        # (momentum form) of the loaded gp
        for L in self.lev:
            L.dba.icns_src_term(L.src_nph, 0, L.gp, L.rho, 3, 1)
        self.dt_prev = 0.0

    def _fill_vel(self, which):
        L0, L1 = self.lev
        v0, v1 = getattr(L0, which), getattr(L1, which)
        self.dc.fill_ghosts(v0, CELL, 3, 3, self.bcrec, self.bcval)
        self.df.fillpatch_two_levels(v1, 3, self.dc, v0, 3, 3, self.bcrec, self.bcval)

    def step(self, dt=None):
        c = self.case
        dc, df = self.dc, self.df
        L0, L1 = self.lev
        self._fill_vel("vel")
        if dt is None:
            cc = max(dc.cfl_max(L0.vel, 3), df.cfl_max(L1.vel, 3))
            ff = max(L.dba.force_cfl_max(L.src_nph, 0, L.rho, 3) for L in self.lev)
            dt = set_current_cfl(c.cfl, cc, ff, self.dt_prev)
        self.last_dt = dt
        self.dt_prev = dt
        for L in self.lev:
            L.dba.icns_src_term(L.src, 1, L.gp, L.rho, 3, 0)
        dc.fill_ghosts(L0.src, CELL, 3, 1, self.src_bc, np.zeros((6, 3)))
        df.fillpatch_two_levels(L1.src, 1, dc, L0.src, 1, 3, self.src_bc, np.zeros((6, 3)))
        for L in self.lev:
            L.dba.extrap_vel_to_faces(L.vel, L.src, self.bcrec, dt, 1, 0, self.limiter, 0, mac=L.mac)
            L.dba.set_inflow_velocity(L.mac, 1, self.bcrec, self.bcval)
        self.last_info = self.mp.project(L0.mac, L1.mac, 1, c.mg_rtol, c.mg_atol)
        for d in range(3):
            bc1, bv1 = self.bcrec[d:d + 1], np.ascontiguousarray(self.bcval[:, d:d + 1])
            dc.fill_boundary(L0.mac[d], XFACE + d, 1, 1)
            dc.fill_physbc(L0.mac[d], XFACE + d, 1, 1, bc1, bv1)
            df.fillpatch_faces_two_levels(L1.mac[d], XFACE + d, 1, dc, L0.mac[d], 1, 1, bc1, bv1)
        for L in self.lev:
            L.dba.scale_by_rho(L.q, 3, L.vel, 3, L.rho, 3, 3, 3)
            L.dba.scale_by_rho(L.fq, 1, L.src, 1, L.rho, 3, 3, 1)
            L.dba.compute_fluxes(3, L.q, L.q, L.mac, None, L.fq, dt, self.bcrec, [1, 1, 1], 1, 0, 1, 0,
                                 self.limiter, 0, faces=L.face, flux=L.flux)
        # conservation across the coarse-fine boundary: the fine fluxes onto
        # the coarse faces before the coarse divergence
        df.average_down_faces_to(L1.flux, 0, dc, L0.flux, 0, 3)
        for L in self.lev:
            L.dba.compute_divergence(L.flux, 3, -1.0, 0, out=L.conv)
            L.dba.icns_src_term(L.src_nph, 0, L.gp, L.rho, 3, 1)
            L.dba.predictor_update(L.vel_new, 3, L.vel, 3, L.rho, 3, L.rho, 3, L.conv, L.src_nph, None, 0.0, dt)
        df.average_down_to(L1.vel_new, 3, dc, L0.vel_new, 3, 3)
        self._fill_vel("vel_new")
        for L in self.lev:
            L.vel, L.vel_new = L.vel_new, L.vel
        return dt

    def velocity(self):
        """the current velocity MultiFabs (level 0, level 1), ghosts filled"""
        return self.lev[0].vel, self.lev[1].vel

    def close(self):
        self.mp.close()
