"""CPU-taped reference of the advect + project step on a triply periodic
one-box domain (density 1, zero pressure gradient, fixed dt) -- TEST
INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU adjoint,
never by the library.

The step is restated with torch tensor operations (float64, CPU) so that
torch.autograd records a tape of it; the gradient of any scalar function
of the final fields is then the reverse sweep of that tape.  Every branch is
written with torch.where on the forward values, so the taped derivative is
the derivative of the branch the forward pass takes (limiter, upwind and
Burgers selections frozen), the convention of the library's adjoint.  At a
tie, min/max select their first argument, as fmin/fmax do.

It follows the oracle's C restatement (ksgf_oracle.c) operation for
operation, in the same operand order:
  ko_extrap_vel_to_faces        ExtrapVelToFaces (icns_advection.H:211-216)
  ko_compute_fluxes_from_state  ComputeFluxesOnBoxFromState, conservative
                                (icns_advection.H:359-373, AdvOp_Godunov.H:231-250)
  ko_compute_divergence         ComputeDivergence, mult -1 (icns_advection.H:418-422)
  ko_predictor_update           density 1, no source (CompRHSOps.H:93-117)
except the MAC projection (MacProjector::project, icns_advection.cpp:331),
which is solved exactly with the FFT of the same 7-point operator instead of
by MLMG; the tape therefore matches the oracle to the MLMG tolerance.
Pinned by tests/test_tape.py against the oracle's step.

Layout: scalar (n2, n1, n0), velocity (3, n2, n1, n0); a face array of
direction d is indexed by its high cell (face i = low face of cell i).
"""
from __future__ import annotations

import math

import torch

SMALL_VEL = 1.0e-8


def _ax(d):
    return 2 - d


def sh(a, d, s):
    """a at index + s along direction d (periodic)."""
    return torch.roll(a, shifts=-s, dims=_ax(d) + (a.dim() - 3))


def vabs(x):
    return torch.where(x < 0.0, -x, x)


def vmin(a, b):
    return torch.where(b < a, b, a)


def vmax(a, b):
    return torch.where(b > a, b, a)


def vanleer(a, b, c):
    dsc = 0.5 * (c - a)
    dsl = 2.0 * (b - a)
    dsr = 2.0 * (c - b)
    m = vmin(vabs(dsc), vmin(vabs(dsl), vabs(dsr)))
    return torch.where(dsl * dsr > 0.0, torch.copysign(m, dsc), torch.zeros_like(m))


def clip2(e, a, b):
    e = vmax(e, vmin(a, b))
    return vmin(e, vmax(a, b))


def mono(s0, sm, sp):
    flat = (sp - s0) * (s0 - sm) <= 0.0
    c2 = vabs(sp - s0) >= 2.0 * vabs(sm - s0)
    c3 = vabs(sm - s0) >= 2.0 * vabs(sp - s0)
    sp_n = torch.where(flat, s0, torch.where(c2, 3.0 * s0 - 2.0 * sm, sp))
    sm_n = torch.where(flat, s0, torch.where(~c2 & c3, 3.0 * s0 - 2.0 * sp, sm))
    return sm_n, sp_n


def ppm_trace(s, d, um, up, dtdx):
    """(Im, Ip) of every cell along d (van Leer limited PPM, periodic);
    um/up: advecting velocity at the low/high face of the cell."""
    sm2, sm1, s0, sp1, sp2 = sh(s, d, -2), sh(s, d, -1), s, sh(s, d, 1), sh(s, d, 2)
    d0 = vanleer(sm2, sm1, s0)
    d1 = vanleer(sm1, s0, sp1)
    d2 = vanleer(s0, sp1, sp2)
    e1 = clip2(0.5 * (sm1 + s0) - (d1 - d0) / 6.0, sm1, s0)
    e2 = clip2(0.5 * (s0 + sp1) - (d2 - d1) / 6.0, s0, sp1)
    sm, sp = mono(s0, e1, e2)
    s6 = 6.0 * s0 - 3.0 * (sm + sp)
    sig_p = vabs(up) * dtdx
    Ip = torch.where(up > SMALL_VEL, sp - (0.5 * sig_p) * ((sp - sm) - (1.0 - (2.0 / 3.0) * sig_p) * s6), s0)
    sig_m = vabs(um) * dtdx
    Im = torch.where(um < -SMALL_VEL, sm + (0.5 * sig_m) * ((sp - sm) + (1.0 - (2.0 / 3.0) * sig_m) * s6), s0)
    return Im, Ip


def burgers(lo, hi):
    st = torch.where((lo + hi) >= 0.0, lo, hi)
    ltm = ((lo <= 0.0) & (hi >= 0.0)) | (vabs(lo + hi) < SMALL_VEL)
    return torch.where(ltm, torch.zeros_like(st), st)


def upwind(lo, hi, u):
    return torch.where(vabs(u) < SMALL_VEL, 0.5 * (lo + hi), torch.where(u >= 0.0, lo, hi))


def extrap_vel_to_faces(u, dt, dx):
    """ko_extrap_vel_to_faces, periodic, no forces in the transverse terms
    (their f = 0 contributes nothing)."""
    dtdx = [dt / dx[d] for d in range(3)]
    c6 = [dt / (6.0 * dx[d]) for d in range(3)]
    c4 = [dt / (4.0 * dx[d]) for d in range(3)]
    flo, fhi, fed, uad = {}, {}, {}, {}
    for d in range(3):
        ud = u[d]
        for n in range(3):
            Im, Ip = ppm_trace(u[n], d, ud, ud, dtdx[d])
            flo[d, n] = sh(Ip, d, -1)
            fhi[d, n] = Im
        uad[d] = burgers(flo[d, d], fhi[d, d])
        for n in range(3):
            fed[d, n] = upwind(flo[d, n], fhi[d, n], uad[d])
    crn = {}
    for d in range(3):
        for t in range(3):
            if t == d:
                continue
            e = 3 - d - t
            # q = f - e_d; qt = q + e_t; pt = f + e_t
            ua_q, ua_qt = sh(uad[t], d, -1), sh(sh(uad[t], d, -1), t, 1)
            ed_q, ed_qt = sh(fed[t, e], d, -1), sh(sh(fed[t, e], d, -1), t, 1)
            lov = flo[d, e] - c6[t] * (ua_qt + ua_q) * (ed_qt - ed_q)
            hiv = fhi[d, e] - c6[t] * (sh(uad[t], t, 1) + uad[t]) * (sh(fed[t, e], t, 1) - fed[t, e])
            crn[d, t] = upwind(lov, hiv, uad[d])
    mac = []
    for d in range(3):
        t1, t2 = (d + 1) % 3, (d + 2) % 3
        A, B = crn[t1, t2], crn[t2, t1]
        q = lambda a: sh(a, d, -1)  # noqa: E731
        stl = (flo[d, d] - c4[t1] * (sh(q(uad[t1]), t1, 1) + q(uad[t1])) * (sh(q(A), t1, 1) - q(A))
               - c4[t2] * (sh(q(uad[t2]), t2, 1) + q(uad[t2])) * (sh(q(B), t2, 1) - q(B)))
        sth = (fhi[d, d] - c4[t1] * (sh(uad[t1], t1, 1) + uad[t1]) * (sh(A, t1, 1) - A)
               - c4[t2] * (sh(uad[t2], t2, 1) + uad[t2]) * (sh(B, t2, 1) - B))
        mac.append(burgers(stl, sth))
    return mac


def fluxes(q, mac, dt, dx):
    """Conservative ko_compute_fluxes_from_state of q (ncomp, n2, n1, n0),
    periodic, no divu, no forcing; returns the 3 flux arrays."""
    ncomp = q.shape[0]
    dtdx = [dt / dx[d] for d in range(3)]
    c3 = [dt / (3.0 * dx[d]) for d in range(3)]
    c2 = [dt / (2.0 * dx[d]) for d in range(3)]
    flo, fhi, fed = {}, {}, {}
    for d in range(3):
        um, up = mac[d], sh(mac[d], d, 1)
        for n in range(ncomp):
            Im, Ip = ppm_trace(q[n], d, um, up, dtdx[d])
            flo[d, n] = sh(Ip, d, -1)
            fhi[d, n] = Im
            fed[d, n] = upwind(flo[d, n], fhi[d, n], mac[d])
    crn = {}
    for d in range(3):
        for t in range(3):
            if t == d:
                continue
            w = mac[t]
            wql, wqh = sh(w, d, -1), sh(sh(w, d, -1), t, 1)
            wpl, wph = w, sh(w, t, 1)
            for n in range(ncomp):
                e = fed[t, n]
                eql, eqh = sh(e, d, -1), sh(sh(e, d, -1), t, 1)
                epl, eph = e, sh(e, t, 1)
                ql, qh = sh(q[n], d, -1), q[n]
                lov = flo[d, n] - c3[t] * (wqh * eqh - wql * eql) + c3[t] * ql * (wqh - wql)
                hiv = fhi[d, n] - c3[t] * (wph * eph - wpl * epl) + c3[t] * qh * (wph - wpl)
                crn[d, t, n] = upwind(lov, hiv, mac[d])
    out = []
    for d in range(3):
        t1, t2 = (d + 1) % 3, (d + 2) % 3
        u = mac[d]
        v1, v2 = mac[t1], mac[t2]
        v1ql, v1qh, v1pl, v1ph = sh(v1, d, -1), sh(sh(v1, d, -1), t1, 1), v1, sh(v1, t1, 1)
        v2ql, v2qh, v2pl, v2ph = sh(v2, d, -1), sh(sh(v2, d, -1), t2, 1), v2, sh(v2, t2, 1)
        fl = []
        for n in range(ncomp):
            A, B = crn[t1, t2, n], crn[t2, t1, n]
            a_ql, a_qh, a_pl, a_ph = sh(A, d, -1), sh(sh(A, d, -1), t1, 1), A, sh(A, t1, 1)
            b_ql, b_qh, b_pl, b_ph = sh(B, d, -1), sh(sh(B, d, -1), t2, 1), B, sh(B, t2, 1)
            ql, qh = sh(q[n], d, -1), q[n]
            stl = (flo[d, n] - c2[t1] * (v1qh * a_qh - v1ql * a_ql) - c2[t2] * (v2qh * b_qh - v2ql * b_ql)
                   + c2[t1] * ql * (v1qh - v1ql) + c2[t2] * ql * (v2qh - v2ql))
            sth = (fhi[d, n] - c2[t1] * (v1ph * a_ph - v1pl * a_pl) - c2[t2] * (v2ph * b_ph - v2pl * b_pl)
                   + c2[t1] * qh * (v1ph - v1pl) + c2[t2] * qh * (v2ph - v2pl))
            fl.append(u * upwind(stl, sth, u))
        out.append(torch.stack(fl))
    return out


def divergence(flux, dx):
    """-div F (ComputeDivergence with mult -1, not area weighted)."""
    r = (sh(flux[0], 0, 1) - flux[0]) * (1.0 / dx[0]) + (sh(flux[1], 1, 1) - flux[1]) * (1.0 / dx[1]) + \
        (sh(flux[2], 2, 1) - flux[2]) * (1.0 / dx[2])
    return -1.0 * r


def project(mac, dx, beta=1.0):
    """Exact periodic MAC projection: div(beta grad phi) = div u with the
    7-point operator, u -= beta grad phi."""
    rhs = (sh(mac[0], 0, 1) - mac[0]) * (1.0 / dx[0]) + (sh(mac[1], 1, 1) - mac[1]) * (1.0 / dx[1]) + \
        (sh(mac[2], 2, 1) - mac[2]) * (1.0 / dx[2])
    n2, n1, n0 = rhs.shape
    lam = torch.zeros(rhs.shape, dtype=torch.float64)
    for d, nd in enumerate((n0, n1, n2)):
        k = torch.arange(nd, dtype=torch.float64)
        ev = (2.0 * torch.cos(2.0 * math.pi * k / nd) - 2.0) * (beta / (dx[d] * dx[d]))
        shape = [1, 1, 1]
        shape[_ax(d)] = nd
        lam = lam + ev.reshape(shape)
    rh = torch.fft.fftn(rhs)
    keep = torch.ones(rhs.shape, dtype=torch.float64)
    keep[0, 0, 0] = 0.0  # the mean of phi (the operator's null space)
    lam[0, 0, 0] = 1.0
    phi = torch.fft.ifftn(rh / lam * keep).real
    return [mac[d] - beta * (phi - sh(phi, d, -1)) * (1.0 / dx[d]) for d in range(3)]


def step(u, s, dt, dx):
    """(u, s) -> (u', s'): the step of the GPU stepper with density 1."""
    mac = project(extrap_vel_to_faces(u, dt, dx), dx)
    s_new = s + dt * divergence(fluxes(s[None], mac, dt, dx), dx)[0]
    u_new = u + dt * divergence(fluxes(u, mac, dt, dx), dx)
    return u_new, s_new


def run(u0, s0, nsteps, dt, dx):
    u, s = u0, s0
    for _ in range(nsteps):
        u, s = step(u, s, dt, dx)
    return u, s


def gradient(u0, s0, nsteps, dt, dx, weights=None):
    """J = ||s_N||_2 (or, with weights = (wu, ws), J = <wu, u_N> + <ws, s_N>)
    and its gradient with respect to (u0, s0), taped."""
    u = u0.detach().clone().requires_grad_(True)
    s = s0.detach().clone().requires_grad_(True)
    uN, sN = run(u, s, nsteps, dt, dx)
    J = sN.square().sum().sqrt() if weights is None else (weights[0] * uN).sum() + (weights[1] * sN).sum()
    gu, gs = torch.autograd.grad(J, (u, s))
    return float(J.detach()), gu, gs
