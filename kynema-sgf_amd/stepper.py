"""One advect + MAC-project step on the GPU, driven through the C ABI.

Mirrors the Godunov branch of incflo::ApplyPredictor
(incflo_advance.cpp:186-438) up to the predictor update, without the
diffusion solve and the nodal projection:

  compute_dt (conv + forcing CFL, growth cap) incflo_compute_dt.cpp:33-220, SimTime.cpp:162-205
  src = -gp/rho, ghost fill                  icns_ops.H:98-111, incflo_advance.cpp:283-291
  ExtrapVelToFaces                           icns_advection.H:211-216
  set_inflow_velocity                        icns_advection.cpp:157-183
  MacProjector::project                      icns_advection.cpp:331-353
  u/v/w_mac ghost fill                       icns_advection.H:233-243
  q = rho u, fq = rho f, fluxes              icns_advection.H:281-374
  conv = -div F                              icns_advection.H:418-422
  u_new = (rho_o u_o + dt (conv + src))/rho  CompRHSOps.H:93-117

With a passive scalar (scalar=True), the scalar is advected by the projected
u_mac after the face fill and before the momentum, as ApplyPredictor orders
the scalar equations (incflo_advance.cpp:316-324; AdvOp_Godunov.H:116-300
with multiply_rho): q = rho s, conservative fluxes, conv = -div F,
s_new = (rho_o s_o + dt conv) / rho_n.
"""
from __future__ import annotations

import numpy as np

from . import (CELL, KSGF_ERR_UNSUPPORTED, XFACE, DeviceBoxArray, KsgfError, MacProjector, default_options,
               set_current_cfl)


class AdvectProjectStep:
    def __init__(self, case, options=None, limiter=1, scalar=False, scal_bc=None, scal_bcval=None):
        import torch

        self.case = case
        self.limiter = limiter
        ba = case.ba
        self.dba = dba = DeviceBoxArray(ba)
        self.vel = dba.alloc(CELL, 3, 3)
        self.vel_new = dba.alloc(CELL, 3, 3)
        self.rho = dba.alloc(CELL, 1, 3)
        self.gp = dba.alloc(CELL, 3, 0)
        self.src = dba.alloc(CELL, 3, 1)
        self.src_nph = dba.alloc(CELL, 3, 0)
        self.q = dba.alloc(CELL, 3, 3)
        self.fq = dba.alloc(CELL, 3, 1)
        self.mac = [dba.alloc(XFACE + d, 1, 1) for d in range(3)]
        self.flux = [dba.alloc(XFACE + d, 3, 0) for d in range(3)]
        self.face = [dba.alloc(XFACE + d, 3, 0) for d in range(3)]
        self.conv = dba.alloc(CELL, 3, 0)
        self.bcrec = np.asarray(case.bcrec, dtype=np.int32)
        self.bcval = np.asarray(case.bcval, dtype=np.float64)
        self.src_bc = self.bcrec.copy()
        self.src_bc[self.src_bc != 0] = 2  # forcing ghosts: first-order extrapolation
        self.rho_bc = np.zeros((1, 6), dtype=np.int32)
        for d in range(3):
            if not case.geom.periodic[d]:
                self.rho_bc[0, d] = self.rho_bc[0, 3 + d] = 2
        self.options = options or default_options(**case.mlmg)
        self.mp = MacProjector(dba, case.lobc, case.hibc, 1.0 / case.rho0, self.options)
        self.last_info = None
        self.last_dt = None
        self.dt_prev = 0.0
        self.src_nph.zero_()
        self._torch = torch
        self.scalar = scalar
        if scalar:
            self.s = dba.alloc(CELL, 1, 3)
            self.s_new = dba.alloc(CELL, 1, 3)
            self.qs = dba.alloc(CELL, 1, 3)
            self.fqs = dba.alloc(CELL, 1, 1)
            self.fqs.zero_()
            self.flux_s = [dba.alloc(XFACE + d, 1, 0) for d in range(3)]
            self.face_s = [dba.alloc(XFACE + d, 1, 0) for d in range(3)]
            self.conv_s = dba.alloc(CELL, 1, 0)
            self.src_s = dba.alloc(CELL, 1, 0)
            self.src_s.zero_()
            self.scal_bc = np.zeros((1, 6), dtype=np.int32) if scal_bc is None else np.asarray(scal_bc, np.int32)
            self.scal_bcval = np.zeros((6, 1)) if scal_bcval is None else np.asarray(scal_bcval, np.float64)

    def load(self, vel=None, rho=None, gp=None):
        """Upload domain-wide fields (defaults: the case's) and fill ghosts."""
        ba, dba, c = self.case.ba, self.dba, self.case
        vel = c.vel if vel is None else vel
        rho = c.rho if rho is None else rho
        gp = c.gp if gp is None else gp
        if np.any(np.asarray(rho) != c.rho0):
            # the projections use beta = 1/rho_0 (and sigma = dt/rho_0); the
            # reference switches to face coefficients 1/rho once density
            # varies (icns_advection.cpp:234-291)
            raise KsgfError("AdvectProjectStep.load: density differs from rho_0 (variable-density projections are "
                            "not provided)", KSGF_ERR_UNSUPPORTED)
        self.vel.copy_(dba.upload(ba.scatter(vel, CELL, 3)))
        dba.fill_ghosts(self.vel, CELL, 3, 3, self.bcrec, self.bcval)
        self.rho.copy_(dba.upload(ba.scatter(rho, CELL, 3)))
        dba.fill_ghosts(self.rho, CELL, 1, 3, self.rho_bc, np.zeros((6, 1)))
        self.gp.copy_(dba.upload(ba.scatter(gp, CELL, 0)))
        # the forcing compute_dt reads before the first step: the source term
        # of the loaded gp (as the initial iterations leave src_term)
        dba.icns_src_term(self.src_nph, 0, self.gp, self.rho, 3, 1)
        self.dt_prev = 0.0

    def load_scalar(self, s):
        """Upload a domain-wide passive scalar (1, nz, ny, nx) and fill ghosts."""
        self.s.copy_(self.dba.upload(self.case.ba.scatter(s, CELL, 3)))
        self.dba.fill_ghosts(self.s, CELL, 1, 3, self.scal_bc, self.scal_bcval)

    def step(self, dt=None, swap=True):
        c, dba = self.case, self.dba
        if dt is None:
            cc = dba.cfl_max(self.vel, 3)
            ff = dba.force_cfl_max(self.src_nph, 0, self.rho, 3)
            dt = set_current_cfl(c.cfl, cc, ff, self.dt_prev)
        self.last_dt = dt
        self.dt_prev = dt
        dba.icns_src_term(self.src, 1, self.gp, self.rho, 3, 0)
        dba.fill_ghosts(self.src, CELL, 3, 1, self.src_bc, np.zeros((6, 3)))
        dba.extrap_vel_to_faces(self.vel, self.src, self.bcrec, dt, 1, 0, self.limiter, 0, mac=self.mac)
        dba.set_inflow_velocity(self.mac, 1, self.bcrec, self.bcval)
        self.last_info = self.mp.project(self.mac, c.mg_rtol, c.mg_atol)
        for d in range(3):
            dba.fill_boundary(self.mac[d], XFACE + d, 1, 1)
            dba.fill_physbc(self.mac[d], XFACE + d, 1, 1, self.bcrec[d:d + 1],
                            np.ascontiguousarray(self.bcval[:, d:d + 1]))
        if self.scalar:
            dba.scale_by_rho(self.qs, 3, self.s, 3, self.rho, 3, 1, 3)
            dba.compute_fluxes(1, self.qs, self.qs, self.mac, None, self.fqs, dt, self.scal_bc, [1], 1, 0, 0, 0,
                               self.limiter, 0, faces=self.face_s, flux=self.flux_s)
            dba.compute_divergence(self.flux_s, 1, -1.0, 0, out=self.conv_s)
            dba.predictor_update(self.s_new, 3, self.s, 3, self.rho, 3, self.rho, 3, self.conv_s, self.src_s, None,
                                 0.0, dt, ncomp=1)
            dba.fill_ghosts(self.s_new, CELL, 1, 3, self.scal_bc, self.scal_bcval)
            if swap:
                self.s, self.s_new = self.s_new, self.s
        dba.scale_by_rho(self.q, 3, self.vel, 3, self.rho, 3, 3, 3)
        dba.scale_by_rho(self.fq, 1, self.src, 1, self.rho, 3, 3, 1)
        dba.compute_fluxes(3, self.q, self.q, self.mac, None, self.fq, dt, self.bcrec, [1, 1, 1], 1, 0, 1, 0,
                           self.limiter, 0, faces=self.face, flux=self.flux)
        dba.compute_divergence(self.flux, 3, -1.0, 0, out=self.conv)
        dba.icns_src_term(self.src_nph, 0, self.gp, self.rho, 3, 1)
        dba.predictor_update(self.vel_new, 3, self.vel, 3, self.rho, 3, self.rho, 3, self.conv, self.src_nph, None,
                             0.0, dt)
        dba.fill_ghosts(self.vel_new, CELL, 3, 3, self.bcrec, self.bcval)
        if swap:
            self.vel, self.vel_new = self.vel_new, self.vel
        return dt
