"""Synthetic workloads of the advect+project step (seeded, decomposition
independent: fields are generated domain-wide from a counter-based Philox
stream and then scattered into the boxes).

Velocity boundary conditions follow BCInterface.cpp:200-260:
  periodic     -> int_dir (all components)
  no_slip_wall -> ext_dir (all components, value 0)
  slip_wall    -> tangential hoextrap, normal ext_dir (value 0)
MAC projection domain BCs follow icns_advection.cpp:21-41: periodic stays
periodic, walls are Neumann.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .layout import (BC_EXT_DIR, BC_HOEXTRAP, BC_INT_DIR, LO_NEUMANN, LO_PERIODIC, BoxArray, Geometry,
                     LevelBoxArray, chop_bricks, chop_random, chop_uniform)


@dataclasses.dataclass
class Case:
    name: str
    geom: Geometry
    ba: BoxArray
    vel: np.ndarray          # (3, nz, ny, nx) cell-centred velocity
    rho: np.ndarray          # (1, nz, ny, nx)
    gp: np.ndarray           # (3, nz, ny, nx) pressure gradient
    bcrec: np.ndarray        # (3, 6) int32 velocity BCRecs (lo xyz, hi xyz)
    bcval: np.ndarray        # (6, 3) Dirichlet values [orientation][comp]
    lobc: tuple
    hibc: tuple
    cfl: float
    nsteps: int
    mg_rtol: float = 1e-11
    mg_atol: float = 2.220446049250313e-16 * 1e2
    rho0: float = 1.0
    mlmg: dict = dataclasses.field(default_factory=dict)  # mac_proj.* overrides
    extra: dict = dataclasses.field(default_factory=dict)

    @property
    def ncell(self):
        return int(np.prod([int(n) for n in self.geom.n_cell]))


def _philox(seed):
    return np.random.Generator(np.random.Philox(seed))


def _periodic_bcs():
    bc = np.zeros((3, 6), dtype=np.int32)  # int_dir everywhere
    return bc, np.zeros((6, 3)), (LO_PERIODIC,) * 3, (LO_PERIODIC,) * 3


def _abl_bcs():
    """periodic x/y, zlo no_slip_wall, zhi slip_wall."""
    bc = np.zeros((3, 6), dtype=np.int32)
    for n in range(3):
        bc[n, 2] = BC_EXT_DIR                         # zlo no-slip: all ext_dir
        bc[n, 5] = BC_EXT_DIR if n == 2 else BC_HOEXTRAP  # zhi slip: normal ext_dir
    return bc, np.zeros((6, 3)), (LO_PERIODIC, LO_PERIODIC, LO_NEUMANN), (LO_PERIODIC, LO_PERIODIC, LO_NEUMANN)


def taylor_green(n=64, max_grid_size=32, noise=1e-3, seed=1, cfl=0.5, nsteps=1):
    """Single-level n^3 unit cube, triply periodic; u = sin(2 pi x) cos(2 pi y)
    cos(2 pi z), v = -cos sin cos, w = 0 (2 pi / L as TaylorGreenVortex.cpp:45-51)
    plus U(-noise, noise) on every component."""
    geom = Geometry((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))
    ba = BoxArray(geom, chop_uniform(geom.n_cell, max_grid_size))
    x = geom.cell_centers(0)[None, None, :]
    y = geom.cell_centers(1)[None, :, None]
    z = geom.cell_centers(2)[:, None, None]
    tp = 2.0 * np.pi
    u = np.sin(tp * x) * np.cos(tp * y) * np.cos(tp * z)
    v = -np.cos(tp * x) * np.sin(tp * y) * np.cos(tp * z)
    w = np.zeros_like(u)
    vel = np.stack(np.broadcast_arrays(u, v, w)).astype(np.float64)
    if noise:
        vel += _philox(seed).uniform(-noise, noise, size=vel.shape)
    bc, bcval, lobc, hibc = _periodic_bcs()
    shape = (1, n, n, n)
    return Case(f"tgv{n}", geom, ba, np.ascontiguousarray(vel), np.ones(shape), np.zeros((3, n, n, n)), bc, bcval,
                lobc, hibc, cfl, nsteps)


def ctv_exact(geom, t, u0=1.0, v0=1.0, omega=0.0):
    """The convecting Taylor vortex at time t on the cell centres of geom
    (ConvectingTaylorVortex.cpp:17-83): u = u0 - cos(pi (x - u0 t)) sin(pi (y - v0 t)) e^{-2 omega t},
    v = v0 + sin(pi (x - u0 t)) cos(pi (y - v0 t)) e^{-2 omega t}, w = 0, and the
    pressure gradient gpx = pi/2 sin(2 pi (x - u0 t)) e^{-4 omega t}, gpy alike.
    Returns (vel, gp), each (3, nz, ny, nx)."""
    n = geom.n_cell
    x = geom.cell_centers(0)[None, None, :]
    y = geom.cell_centers(1)[None, :, None]
    pi = np.pi
    ex2, ex4 = np.exp(-2.0 * omega * t), np.exp(-4.0 * omega * t)
    u = u0 - np.cos(pi * (x - u0 * t)) * np.sin(pi * (y - v0 * t)) * ex2
    v = v0 + np.sin(pi * (x - u0 * t)) * np.cos(pi * (y - v0 * t)) * ex2
    gx = 0.5 * pi * np.sin(2.0 * pi * (x - u0 * t)) * ex4 + 0.0 * y
    gy = 0.5 * pi * np.sin(2.0 * pi * (y - v0 * t)) * ex4 + 0.0 * x
    sh = (int(n[2]), int(n[1]), int(n[0]))
    z = np.zeros(sh)
    vel = np.stack([np.broadcast_to(u, sh), np.broadcast_to(v, sh), z])
    gp = np.stack([np.broadcast_to(gx, sh), np.broadcast_to(gy, sh), z])
    return np.ascontiguousarray(vel), np.ascontiguousarray(gp)


def convecting_taylor_vortex(n=32, nz=4, u0=1.0, v0=1.0, max_grid_size=32, cfl=0.5):
    """The reference's convecting Taylor vortex verification case
    (test/test_files/ctv_godunov_plm: [0, 2]^3, triply periodic, CTV.u0 =
    CTV.v0 = 1, activate_pressure so gp starts exact), inviscid, on an
    n x n x nz slab of cubic cells (the field does not depend on z).
    extra['exact'](t) gives the exact (vel, gp) at time t."""
    h = 2.0 / n
    geom = Geometry((n, n, nz), (0.0, 0.0, 0.0), (2.0, 2.0, h * nz), (1, 1, 1))
    ba = BoxArray(geom, chop_uniform(geom.n_cell, max_grid_size))
    vel, gp = ctv_exact(geom, 0.0, u0, v0)
    bc, bcval, lobc, hibc = _periodic_bcs()
    c = Case(f"ctv{n}", geom, ba, vel, np.ones((1, nz, n, n)), gp, bc, bcval, lobc, hibc, cfl, 1)
    c.extra["exact"] = lambda t: ctv_exact(geom, t, u0, v0)
    return c


def ragged(n_cell=(320, 320, 480), nbox=1500, edges=(8, 16, 24, 32, 48, 64), box_seed=4, noise=1e-3, seed=1,
           cfl=0.5, nsteps=1):
    """Ragged single-level layout: ``nbox`` boxes with edges drawn from
    ``edges`` per axis (brick wall: rows and slabs do not line up), over a
    triply periodic domain of unit cell size 1/n_cell[0]; Taylor-Green plus
    U(-noise, noise) velocity as in taylor_green()."""
    h = 1.0 / n_cell[0]
    geom = Geometry(tuple(n_cell), (0.0, 0.0, 0.0), tuple(h * n for n in n_cell), (1, 1, 1))
    ba = BoxArray(geom, chop_bricks(n_cell, edges, _philox(box_seed), nbox))
    x = geom.cell_centers(0)[None, None, :]
    y = geom.cell_centers(1)[None, :, None]
    z = geom.cell_centers(2)[:, None, None]
    tp = 2.0 * np.pi
    u = np.sin(tp * x) * np.cos(tp * y) * np.cos(tp * z)
    v = -np.cos(tp * x) * np.sin(tp * y) * np.cos(tp * z)
    vel = np.stack(np.broadcast_arrays(u, v, np.zeros_like(u))).astype(np.float64)
    if noise:
        vel += _philox(seed).uniform(-noise, noise, size=vel.shape)
    bc, bcval, lobc, hibc = _periodic_bcs()
    nx, ny, nz = n_cell
    return Case(f"ragged{ba.nbox}_{nx}x{ny}x{nz}", geom, ba, np.ascontiguousarray(vel), np.ones((1, nz, ny, nx)),
                np.zeros((3, nz, ny, nx)), bc, bcval, lobc, hibc, cfl, nsteps)


def two_level_faces(geom, xp, X, Y, Z, d, amp=0.1):
    """Face velocity of the two-level workload at positions (X, Y, Z): the
    Taylor-Green field with wavenumbers 2 pi / L per axis (periodic on the
    domain) plus amp * grad(cos cos cos), a divergent part for the
    projection to remove.  xp: numpy or torch."""
    k = [2.0 * np.pi / (geom.prob_hi[e] - geom.prob_lo[e]) for e in range(3)]
    sx, cx = xp.sin(k[0] * X), xp.cos(k[0] * X)
    sy, cy = xp.sin(k[1] * Y), xp.cos(k[1] * Y)
    sz, cz = xp.sin(k[2] * Z), xp.cos(k[2] * Z)
    if d == 0:
        return sx * cy * cz - amp * k[0] * sx * cy * cz
    if d == 1:
        return -cx * sy * cz - amp * k[1] * cx * sy * cz
    return -amp * k[2] * cx * cy * sz


def ragged_two_level(n_cell=(320, 320, 480), nbox=1500, fine_z=(96, 288), fine_box=64, box_seed=4):
    """The ragged layout (ragged()) as AMR level 0 and a level-1 BoxArray at
    ratio 2 over the coarse slab z in [fine_z[0], fine_z[1]) (40% of the
    domain by default), chopped into fine boxes of up to fine_box cells per
    edge (aligned to 2).  Returns (case, fine LevelBoxArray)."""
    case = ragged(n_cell, nbox, box_seed=box_seed, noise=0.0)
    nf = tuple(2 * v for v in n_cell)
    fg = Geometry(nf, (0.0, 0.0, 0.0), case.geom.prob_hi, (1, 1, 1))
    lo = (0, 0, 2 * fine_z[0])
    hi = (nf[0], nf[1], 2 * fine_z[1])
    edges = [list(range(lo[d], hi[d], fine_box)) + [hi[d]] for d in range(3)]
    boxes = [[x0, y0, z0, x1 - 1, y1 - 1, z1 - 1]
             for z0, z1 in zip(edges[2][:-1], edges[2][1:])
             for y0, y1 in zip(edges[1][:-1], edges[1][1:])
             for x0, x1 in zip(edges[0][:-1], edges[0][1:])]
    return case, LevelBoxArray(fg, boxes)


def abl(n_cell=(256, 256, 128), extent=(5120.0, 5120.0, 1024.0), sizes=(32, 48, 64), cfl=0.9, nsteps=10,
        seed=2, box_seed=3, sigma=0.5, z_perturb=300.0):
    """Neutral ABL box: periodic x/y, no-slip bottom, slip top; mean log law
    u(z) = (0.4/0.4) ln(z/0.1) along x, seeded N(0, sigma^2) perturbations of
    all three components below z_perturb; uneven boxes with edges drawn from
    ``sizes``.  Potential temperature 300 K with 0.003 K/m above 700 m is
    carried as a passive field (extra['theta'])."""
    geom = Geometry(tuple(n_cell), (0.0, 0.0, 0.0), tuple(extent), (1, 1, 0))
    ba = BoxArray(geom, chop_random(n_cell, list(sizes), _philox(box_seed)))
    nx, ny, nz = n_cell
    z = geom.cell_centers(2)
    u = (0.4 / 0.4) * np.log(z / 0.1)
    vel = np.zeros((3, nz, ny, nx))
    vel[0] = u[:, None, None]
    pert = _philox(seed).normal(0.0, sigma, size=vel.shape)
    mask = (z < z_perturb)[None, :, None, None]
    vel += np.where(mask, pert, 0.0)
    theta = 300.0 + 0.003 * np.maximum(z - 700.0, 0.0)
    bc, bcval, lobc, hibc = _abl_bcs()
    c = Case(f"abl{nx}x{ny}x{nz}", geom, ba, np.ascontiguousarray(vel), np.ones((1, nz, ny, nx)),
             np.zeros((3, nz, ny, nx)), bc, bcval, lobc, hibc, cfl, nsteps)
    c.extra["theta"] = np.broadcast_to(theta[:, None, None], (nz, ny, nx)).copy()
    # cell aspect ratio dx/dz = 2.5 (256x256x128) or 5 (256^3): standard
    # coarsening with piecewise-constant prolongation stalls or diverges, so
    # the deck turns on mac_proj.do_semicoarsening (LPInfo option,
    # MLMGOptions.cpp:24)
    c.mlmg["do_semicoarsening"] = 1
    return c


def random_case(n_cell, boxes=None, seed=7, periodic=(1, 1, 1), extent=(1.0, 1.0, 1.0), amp=1.0, gp_amp=0.0,
                rho_var=0.0, cfl=0.5):
    """Seeded random velocity (and optionally pressure gradient / density)
    for parity tests on small grids."""
    geom = Geometry(tuple(n_cell), (0.0, 0.0, 0.0), tuple(extent), tuple(periodic))
    if boxes is None:
        boxes = chop_uniform(n_cell, max(n_cell))
    ba = BoxArray(geom, boxes)
    nx, ny, nz = n_cell
    g = _philox(seed)
    vel = amp * g.uniform(-1.0, 1.0, size=(3, nz, ny, nx))
    gp = gp_amp * g.uniform(-1.0, 1.0, size=(3, nz, ny, nx))
    rho = 1.0 + rho_var * g.uniform(0.0, 1.0, size=(1, nz, ny, nx))
    if all(periodic):
        bc, bcval, lobc, hibc = _periodic_bcs()
    else:
        bc = np.zeros((3, 6), dtype=np.int32)
        lobc, hibc = [LO_PERIODIC] * 3, [LO_PERIODIC] * 3
        for d in range(3):
            if not periodic[d]:
                lobc[d] = hibc[d] = LO_NEUMANN
                for n in range(3):
                    bc[n, d] = BC_EXT_DIR
                    bc[n, 3 + d] = BC_EXT_DIR if n == d else BC_HOEXTRAP
        bcval = np.zeros((6, 3))
        lobc, hibc = tuple(lobc), tuple(hibc)
    return Case(f"rand{nx}x{ny}x{nz}", geom, ba, np.ascontiguousarray(vel), rho, gp, bc, bcval, lobc, hibc, cfl, 1)


def poisson_problems(nprob=64, edges=(32, 48, 64, 80, 96, 112, 128), seed=5, bcs=None):
    """Independent single-box Poisson problems: cube edges drawn from
    `edges`, unit extent, beta 1, right-hand sides N(0,1) minus their mean
    (the mean is removed for every singular problem: all periodic/Neumann).
    bcs: list of (lobc, hibc) cycled over the problems (default periodic).
    Returns (problems, rhs list of flat float64 arrays, Fortran order)."""
    g = _philox(seed)
    bcs = bcs or [((LO_PERIODIC,) * 3, (LO_PERIODIC,) * 3)]
    probs, rhs = [], []
    for q in range(nprob):
        e = int(edges[int(g.integers(0, len(edges)))])
        lo, hi = bcs[q % len(bcs)]
        n = (e, e, e)
        f = g.standard_normal(e * e * e)
        if 2 not in tuple(lo) + tuple(hi):  # no Dirichlet side: singular
            f -= f.mean()
        probs.append(dict(n=n, dx=(1.0 / e,) * 3, lobc=tuple(lo), hibc=tuple(hi), beta=1.0))
        rhs.append(f)
    return probs, rhs


def differentiation(n=128, noise=1e-3, seed=1, sigma=0.1, cfl=0.5, nsteps=20, center=None, blob_seed=5):
    """Periodic unit cube n^3: Taylor-Green plus U(-noise, noise) velocity
    (as taylor_green()), a Gaussian blob s = exp(-|x - x0|^2 / (2 sigma^2))
    as the passive scalar (extra['scalar'], (1, n, n, n)), and the fixed
    time step of CFL `cfl` for the initial velocity (extra['dt'],
    dt = cfl / max_d max |u_d| / dx).  x0 is the box centre moved by a
    seeded U(-1/2, 1/2) cell in each direction (blob_seed), unless `center`
    is given: with the blob exactly at the centre (a face of an even grid)
    mirror cells hold equal values and the PPM limiters sit exactly on their
    switch points, where the step has no derivative."""
    c = taylor_green(n, n, noise=noise, seed=seed, cfl=cfl, nsteps=nsteps)
    if center is None:
        center = tuple(0.5 + (_philox(blob_seed).uniform(-0.5, 0.5, size=3) / n))
    x = [c.geom.cell_centers(d) for d in range(3)]
    r2 = (((x[0] - center[0]) ** 2)[None, None, :] + ((x[1] - center[1]) ** 2)[None, :, None]
          + ((x[2] - center[2]) ** 2)[:, None, None])
    c.extra["scalar"] = np.exp(-r2 / (2.0 * sigma * sigma))[None].copy()
    h = 1.0 / n
    cmax = max(float(np.max(np.abs(c.vel[d]))) / h for d in range(3))
    c.extra["dt"] = cfl / cmax
    c.name = f"diff{n}"
    return c


def cells_on_device(ba, ncomp, ng, fn):
    """A cell MultiFab (ncomp, ngrow ng) built on the device, each FAB
    (ghosts included) filled with fn(c, X, Y, Z) at its cell centres (torch
    tensors): large layouts without a domain-wide host array."""
    import torch

    out = torch.empty(ba.mf_size(0, ncomp, ng), dtype=torch.float64, device="cuda")
    lo, n = ba.fab_lo_n(0, ng)
    off = ba.fab_offsets(0, ncomp, ng)
    h = ba.geom.dx
    for b in range(ba.nbox):
        ax = [(int(lo[b, e]) + torch.arange(int(n[b, e]), dtype=torch.float64, device="cuda") + 0.5) * h[e]
              for e in range(3)]
        Z, Y, X = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        out[int(off[b]):int(off[b + 1])] = torch.stack([fn(c, X, Y, Z) for c in range(ncomp)]).reshape(-1)
    return out


def tgv_cells(geom):
    """fn(c, X, Y, Z) of the two-level workload's velocity: Taylor-Green with
    wavenumbers 2 pi / L per axis and a w component 0.1 u v, periodic on the
    domain"""
    import torch

    k = [2.0 * np.pi / (geom.prob_hi[e] - geom.prob_lo[e]) for e in range(3)]

    def fn(c, X, Y, Z):
        u = torch.sin(k[0] * X) * torch.cos(k[1] * Y) * torch.cos(k[2] * Z)
        v = -torch.cos(k[0] * X) * torch.sin(k[1] * Y) * torch.cos(k[2] * Z)
        return u if c == 0 else (v if c == 1 else 0.1 * u * v)

    return fn
