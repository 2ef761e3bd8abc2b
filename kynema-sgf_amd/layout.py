"""BoxArray / MultiFab layout arithmetic (host side, numpy only).

The layout is the one documented in include/ksgf.h: a MultiFab is one flat
float64 array holding one FAB per box in BoxArray order, no padding; FAB b
covers box b converted to the index type (nodal in direction d for d-faces)
and grown by ngrow, stored Fortran order (x fastest) with the component
index slowest.  This module computes offsets/shapes and scatters/gathers
domain-wide numpy arrays to/from that layout, so tests and workloads can be
written against plain (ncomp, nz, ny, nx) arrays.
"""
from __future__ import annotations

import dataclasses

import numpy as np

CELL, XFACE, YFACE, ZFACE = 0, 1, 2, 3

# amrex::BCType values
BC_INT_DIR, BC_FOEXTRAP, BC_EXT_DIR, BC_HOEXTRAP = 0, 2, 3, 4
# LinOp domain BCs
LO_PERIODIC, LO_NEUMANN, LO_DIRICHLET = 0, 1, 2
# PPM limiter types
PPM_NOLIMITER, PPM_VANLEER, PPM_WENOZ, PPM_WENO_JS, PPM_UPWIND, PPM_MINMOD = 0, 1, 2, 3, 4, 5


@dataclasses.dataclass
class Geometry:
    """amrex::Geometry subset: index domain, physical extent, periodicity."""

    n_cell: tuple
    prob_lo: tuple = (0.0, 0.0, 0.0)
    prob_hi: tuple = (1.0, 1.0, 1.0)
    periodic: tuple = (1, 1, 1)

    @property
    def dlo(self):
        return (0, 0, 0)

    @property
    def dhi(self):
        return tuple(int(n) - 1 for n in self.n_cell)

    @property
    def dx(self):
        return tuple((self.prob_hi[d] - self.prob_lo[d]) / self.n_cell[d] for d in range(3))

    def cell_centers(self, d):
        n = self.n_cell[d]
        return self.prob_lo[d] + (np.arange(n) + 0.5) * self.dx[d]


class BoxArray:
    """Disjoint boxes (cell-centred, inclusive lo/hi) tiling a Geometry's domain."""

    def __init__(self, geom: Geometry, boxes):
        self.geom = geom
        self.boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self.nbox = self.boxes.shape[0]
        ncell = int(np.prod([int(n) for n in geom.n_cell]))
        covered = int(np.sum(np.prod(self.boxes[:, 3:] - self.boxes[:, :3] + 1, axis=1)))
        if covered != ncell:
            raise ValueError("boxes do not tile the domain")

    # -- FAB shapes -------------------------------------------------------
    def fab_lo_n(self, itype, ngrow):
        lo = self.boxes[:, :3].astype(np.int64) - ngrow
        n = self.boxes[:, 3:].astype(np.int64) - self.boxes[:, :3] + 1 + 2 * ngrow
        if itype in (XFACE, YFACE, ZFACE):
            n[:, itype - 1] += 1
        return lo, n

    def mf_size(self, itype, ncomp, ngrow):
        _, n = self.fab_lo_n(itype, ngrow)
        return int(np.sum(np.prod(n, axis=1))) * ncomp

    def fab_offsets(self, itype, ncomp, ngrow):
        _, n = self.fab_lo_n(itype, ngrow)
        sz = np.prod(n, axis=1) * ncomp
        off = np.zeros(self.nbox + 1, dtype=np.int64)
        off[1:] = np.cumsum(sz)
        return off

    def fab_view(self, flat, b, itype, ncomp, ngrow):
        """numpy view (ncomp, nz, ny, nx) of FAB b inside a flat MultiFab array."""
        lo, n = self.fab_lo_n(itype, ngrow)
        off = self.fab_offsets(itype, ncomp, ngrow)
        nx, ny, nz = (int(v) for v in n[b])
        return flat[off[b]:off[b + 1]].reshape(ncomp, nz, ny, nx)

    # -- scatter / gather ---------------------------------------------------
    def scatter(self, field, itype, ngrow, out=None):
        """Copy a domain-wide array into the valid region of every FAB.

        field: (ncomp, NZ, NY, NX) with N = n_cell (+1 along the nodal
        direction for face data).  Ghost entries are left at zero (or as in
        ``out``); fill them with FillBoundary and the physical BC fill."""
        field = np.asarray(field, dtype=np.float64)
        if field.ndim == 3:
            field = field[None]
        ncomp = field.shape[0]
        if out is None:
            out = np.zeros(self.mf_size(itype, ncomp, ngrow), dtype=np.float64)
        for b in range(self.nbox):
            v = self.fab_view(out, b, itype, ncomp, ngrow)
            lo = self.boxes[b, :3]
            hi = self.boxes[b, 3:] + (np.arange(3) == (itype - 1)).astype(np.int32)
            g = ngrow
            v[:, g:g + hi[2] - lo[2] + 1, g:g + hi[1] - lo[1] + 1, g:g + hi[0] - lo[0] + 1] = field[
                :, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1
            ]
        return out

    def gather(self, flat, itype, ncomp, ngrow):
        """Domain-wide array of the valid data of every FAB (faces shared by
        two boxes are taken from the later box; they hold equal values)."""
        flat = np.asarray(flat)
        shp = [int(n) for n in self.geom.n_cell]
        if itype in (XFACE, YFACE, ZFACE):
            shp[itype - 1] += 1
        out = np.zeros((ncomp, shp[2], shp[1], shp[0]), dtype=flat.dtype)
        for b in range(self.nbox):
            v = self.fab_view(flat, b, itype, ncomp, ngrow)
            lo = self.boxes[b, :3]
            hi = self.boxes[b, 3:] + (np.arange(3) == (itype - 1)).astype(np.int32)
            g = ngrow
            out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = v[
                :, g:g + hi[2] - lo[2] + 1, g:g + hi[1] - lo[1] + 1, g:g + hi[0] - lo[0] + 1
            ]
        return out


class LevelBoxArray(BoxArray):
    """The BoxArray of a refined AMR level: disjoint boxes that cover part of
    the domain (no tiling check)."""

    def __init__(self, geom: Geometry, boxes):
        self.geom = geom
        self.boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self.nbox = self.boxes.shape[0]


def chop_uniform(n_cell, max_grid_size):
    """amr.max_grid_size chopping: equal-as-possible boxes no larger than
    max_grid_size along each axis."""
    segs = []
    for d in range(3):
        L = int(n_cell[d])
        m = int(max_grid_size[d] if np.ndim(max_grid_size) else max_grid_size)
        nseg = -(-L // m)
        base, extra = divmod(L, nseg)
        lens = [base + (1 if s < extra else 0) for s in range(nseg)]
        segs.append(lens)
    return _tensor_boxes(segs)


def chop_random(n_cell, choices, rng, granularity=8):
    """Uneven boxes: along each axis, segment lengths drawn from ``choices``
    (the last one shortened to fit, kept a multiple of ``granularity``)."""
    segs = []
    for d in range(3):
        L = int(n_cell[d])
        lens = []
        rem = L
        while rem > max(choices):
            ok = [int(c) for c in choices if rem - int(c) >= granularity]
            if not ok:
                break
            c = ok[int(rng.integers(len(ok)))]
            lens.append(c)
            rem -= c
        if rem > 0:
            lens.append(rem)
        segs.append(lens)
    return _tensor_boxes(segs)


def chop_bricks(n_cell, choices, rng, nbox=None):
    """Ragged brick-wall layout: z slabs with thicknesses drawn from
    ``choices``; every slab has its own y rows and every row its own x
    pieces, all drawn from ``choices`` (so boxes of neighbouring rows and
    slabs do not line up).  Extents must be multiples of min(choices), and
    every choice a multiple of it.  With ``nbox``, x pieces are then split in
    two or merged pairwise (lengths staying in ``choices``) until there are
    exactly ``nbox`` boxes."""
    ch = sorted(int(c) for c in choices)
    g = ch[0]
    if any(c % g for c in ch) or any(int(n) % g for n in n_cell):
        raise ValueError("extents and choices must be multiples of the smallest choice")

    def compose(L):
        parts, rem = [], L
        while rem > 0:
            ok = [c for c in ch if c <= rem]
            c = ok[int(rng.integers(len(ok)))]
            parts.append(c)
            rem -= c
        return parts

    rows = []  # [z0, dz, y0, dy, [x lengths]]
    z0 = 0
    for dz in compose(int(n_cell[2])):
        y0 = 0
        for dy in compose(int(n_cell[1])):
            rows.append([z0, dz, y0, dy, compose(int(n_cell[0]))])
            y0 += dy
        z0 += dz
    count = sum(len(r[4]) for r in rows)
    if nbox is not None and nbox > len(rows) * (int(n_cell[0]) // g):
        raise ValueError(f"{nbox} boxes need more than {len(rows)} rows of {n_cell[0]} cells")
    chs = set(ch)
    guard = 0
    while nbox is not None and count != nbox:
        guard += 1
        if guard > 100000:
            raise ValueError(f"cannot reach {nbox} boxes")
        r = rows[int(rng.integers(len(rows)))][4]
        q = int(rng.integers(len(r)))
        if count < nbox:
            cuts = [a for a in ch if a < r[q] and r[q] - a in chs]
            if cuts:
                a = cuts[int(rng.integers(len(cuts)))]
                r[q:q + 1] = [a, r[q] - a]
                count += 1
        elif q + 1 < len(r) and r[q] + r[q + 1] in chs:
            r[q:q + 2] = [r[q] + r[q + 1]]
            count -= 1
    boxes = []
    for z0, dz, y0, dy, xs in rows:
        x0 = 0
        for dx in xs:
            boxes.append([x0, y0, z0, x0 + dx - 1, y0 + dy - 1, z0 + dz - 1])
            x0 += dx
    return boxes


def _tensor_boxes(segs):
    starts = []
    for lens in segs:
        s = np.concatenate([[0], np.cumsum(lens)])
        starts.append([(int(s[i]), int(s[i + 1]) - 1) for i in range(len(lens))])
    boxes = []
    for (z0, z1) in starts[2]:
        for (y0, y1) in starts[1]:
            for (x0, x1) in starts[0]:
                boxes.append([x0, y0, z0, x1, y1, z1])
    return np.asarray(boxes, dtype=np.int32)


def bcrec_array(bcs):
    """list of ((lo0,lo1,lo2),(hi0,hi1,hi2)) -> int32 array (ncomp, 6)."""
    a = np.zeros((len(bcs), 6), dtype=np.int32)
    for n, (lo, hi) in enumerate(bcs):
        a[n, :3] = lo
        a[n, 3:] = hi
    return a
