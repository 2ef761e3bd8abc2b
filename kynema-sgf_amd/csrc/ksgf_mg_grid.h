// Multigrid level layouts and the per-cell bodies of the grid-level
// multigrid kernels (red-black half sweeps, residual + restriction).
//
// The bodies are plain functions of one cell so that they compile both into
// the HIP kernels and into host C++: a CPU test runs them cell by cell
// against a straightforward natural-layout implementation, bit for bit,
// before the kernels run on a GPU.
#pragma once

#include "../../include/ksgf.h"
#include "ksgf_hd.h"

namespace ksgf {

// One multigrid level: cell counts, the 7-point coefficients, the LinOp
// BCs and one of two layouts.
//
// Natural layout: a domain-wide array with one ghost layer, Fortran order;
// cell (i, j, k) lives at LIdx().  Cell i = 0 of every row sits kXOff
// elements into the row and the row pitch sy is a multiple of 16, so rows
// start 64-byte aligned (FP32).
//
// Red-black packed layout (even n0): the cells of colour c = (i+j+k) & 1
// form one block of n0/2 * n1 * n2 values (red block first), Fortran order
// within the block over (i/2, j, k); cell (i, j, k) lives at PIdx().  A half
// sweep then streams exactly the cells it updates and the other colour's
// values, instead of every cache line of both colours; there are no ghost
// cells -- domain sides are handled in the stencil (periodic wrap, Neumann
// mirror, Dirichlet odd reflection).
struct LevelGeom {
    int n0, n1, n2;
    long long sy, sz, tot;  // natural strides; tot = elements to allocate (either layout)
    double cx, cy, cz, dg, idg;
    int lo[3], hi[3];  // LinOp BCs
    int ghost_mode;    // natural layout only: explicit ghost fills before every colour
    int packed;        // 1: red-black packed layout
    int prow, pplane, pcol;  // packed strides: row (n0/2), plane, colour block
    // device-controlled solves: the kernels of a V-cycle return at once when
    // *act reads 0 (the solve has converged); nullptr otherwise
    const int* act = nullptr;
};

constexpr int kXOff = 16;

// level arrays hold fewer than 2^31 elements (checked when a hierarchy is
// built), so offsets are formed in 32-bit arithmetic
KSGF_HD long long LIdx(const LevelGeom& L, int i, int j, int k)
{
    return (long long)((i + kXOff) + (int)L.sy * (j + 1) + (int)L.sz * (k + 1));
}

KSGF_HD long long PIdx(const LevelGeom& L, int i, int j, int k)
{
    return (long long)(((i + j + k) & 1) * L.pcol + k * L.pplane + j * L.prow + (i >> 1));
}

KSGF_HD long long AnyIdx(const LevelGeom& L, int i, int j, int k)
{
    return L.packed ? PIdx(L, i, j, k) : LIdx(L, i, j, k);
}

// offset of cell (i, j, k) in a level whose layout is known at compile time
template <bool PACKED>
KSGF_HD int LvIdx(const LevelGeom& L, int i, int j, int k)
{
    if (PACKED) return ((i + j + k) & 1) * L.pcol + k * L.pplane + j * L.prow + (i >> 1);
    return (i + kXOff) + (int)L.sy * (j + 1) + (int)L.sz * (k + 1);
}

inline void level_layout(LevelGeom& L, bool packed)
{
    L.sy = ((long long)L.n0 + kXOff + 1 + 15) / 16 * 16;
    L.sz = L.sy * (L.n1 + 2);
    L.packed = packed ? 1 : 0;
    if (packed) {
        L.prow = L.n0 / 2;
        L.pplane = L.prow * L.n1;
        L.pcol = L.pplane * L.n2;
        L.tot = 2LL * L.pcol;
    } else {
        L.prow = L.pplane = L.pcol = 0;
        L.tot = L.sz * (L.n2 + 2);
    }
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup w runs on
// XCD w mod 8), and each XCD has its own L2.  The remap gives every XCD one
// contiguous range of logical tiles, so the rows and planes a stencil shares
// between neighbouring tiles are fetched into one L2 instead of three.
constexpr int kXcds = 8;

KSGF_HD int xcd_remap(int bid, int nb)
{
    const int per = nb / kXcds, rem = nb % kXcds;
    const int x = bid % kXcds, pos = bid / kXcds;
    return (x < rem) ? x * (per + 1) + pos : rem * (per + 1) + (x - rem) * per + pos;
}

// ---------------------------------------------------------------------------
// row tiles: 64 positions along x by 4 rows, kz planes, 256 threads
// ---------------------------------------------------------------------------
constexpr int kRowTX = 64, kRowTY = 4;

struct RowTiles {
    int ntx, nty, nkc, kz;
};

inline RowTiles row_tiles(int n0, int n1, int n2, int kz)
{
    RowTiles t;
    t.kz = kz;
    t.ntx = (n0 + kRowTX - 1) / kRowTX;
    t.nty = (n1 + kRowTY - 1) / kRowTY;
    t.nkc = (n2 + kz - 1) / kz;
    return t;
}

KSGF_HD void row_tile(const RowTiles& t, int lid, int tid, int& x, int& y, int& k0, int& k1, int n2)
{
    const int tx = lid % t.ntx, q = lid / t.ntx, ty = q % t.nty, kc = q / t.nty;
    x = tx * kRowTX + (tid & (kRowTX - 1));
    y = ty * kRowTY + tid / kRowTX;
    k0 = kc * t.kz;
    k1 = (k0 + t.kz < n2) ? k0 + t.kz : n2;
}

// ---------------------------------------------------------------------------
// packed red-black half sweep
// ---------------------------------------------------------------------------
// what the old iterate is: the array as it stands; a zero initial guess (the
// red half writes -rhs/diag, the black half mirrors 0 at walls); or the
// array plus the piecewise-constant prolongation of the coarse correction,
// formed on the fly (the red half reads prolonged black neighbours, the
// black half mirrors its prolonged own value at walls)
// kPkProlong: first sweep after the coarse correction, prolonging it on the
// fly (the red half adds it to every neighbour, the black half to itself at
// walls); kPkProlongWalls: red half of that sweep after pk_prolong_color has
// already added the correction to the black cells (so only a wall-mirrored
// self needs it)
enum { kPkPlain = 0, kPkZero = 1, kPkProlong = 2, kPkProlongWalls = 3 };

template <typename T>
struct PkArgs {
    T* x;          // packed level array (both colours)
    const T* b;    // packed right-hand side
    const T* xc;   // coarse correction (prolongation modes)
    LevelGeom L, C;
    int s0, s1, s2;  // log2 of the coarsening ratio to C
};

// wall factor of a side: a Neumann wall mirrors the cell (+1), a Dirichlet
// wall takes its negative (-1); multiplying by it is exact
template <typename T>
KSGF_HD T pk_wall(int bc)
{
    return bc == KSGF_LO_NEUMANN ? (T)1 : (T)-1;
}

// relax the colour-`color` cell at packed position (h, j, k):
// x = (cx*(W+E) + cy*(S+N) + cz*(B+T) - b) / diag, the oracle's operand order
template <typename T, int MODE, bool CP>
KSGF_HD void pk_relax(const PkArgs<T>& a, int color, int h, int j, int k)
{
    const LevelGeom& L = a.L;
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    const int par = (j + k + color) & 1;
    const int i = 2 * h + par;
    const int c = color * L.pcol + k * L.pplane + j * L.prow + h;
    if (MODE == kPkZero && color == 0) {
        a.x[c] = ((T)0 - a.b[c]) * idg;
        return;
    }
    const bool px = L.lo[0] == KSGF_LO_PERIODIC, py = L.lo[1] == KSGF_LO_PERIODIC, pz = L.lo[2] == KSGF_LO_PERIODIC;
    const bool in_w = i > 0 || px, in_e = i < L.n0 - 1 || px, in_s = j > 0 || py, in_n = j < L.n1 - 1 || py;
    const bool in_b = k > 0 || pz, in_t = k < L.n2 - 1 || pz;
    // the cell's own old value is needed only where a wall mirrors it
    T self = (T)0;
    if (!(in_w && in_e && in_s && in_n && in_b && in_t)) {
        if (MODE == kPkZero)
            self = (T)0;
        else if (MODE == kPkProlong || MODE == kPkProlongWalls)
            self = a.x[c] + a.xc[LvIdx<CP>(a.C, i >> a.s0, j >> a.s1, k >> a.s2)];
        else
            self = a.x[c];
    }
    // neighbours: the other colour; prolonged when this is the red half of a
    // prolongation sweep (the black cells have not been relaxed since)
    const bool pro = (MODE == kPkProlong) && color == 0;
    const int oth = (1 - color) * L.pcol;
    auto nb = [&](int ii, int jj, int kk) {
        T v = a.x[oth + kk * L.pplane + jj * L.prow + (ii >> 1)];
        if (pro) v = v + a.xc[LvIdx<CP>(a.C, ii >> a.s0, jj >> a.s1, kk >> a.s2)];
        return v;
    };
    const T w = in_w ? nb(i > 0 ? i - 1 : L.n0 - 1, j, k) : pk_wall<T>(L.lo[0]) * self;
    const T e = in_e ? nb(i < L.n0 - 1 ? i + 1 : 0, j, k) : pk_wall<T>(L.hi[0]) * self;
    const T s = in_s ? nb(i, j > 0 ? j - 1 : L.n1 - 1, k) : pk_wall<T>(L.lo[1]) * self;
    const T n = in_n ? nb(i, j < L.n1 - 1 ? j + 1 : 0, k) : pk_wall<T>(L.hi[1]) * self;
    const T bo = in_b ? nb(i, j, k > 0 ? k - 1 : L.n2 - 1) : pk_wall<T>(L.lo[2]) * self;
    const T t = in_t ? nb(i, j, k < L.n2 - 1 ? k + 1 : 0) : pk_wall<T>(L.hi[2]) * self;
    a.x[c] = ((cx * (w + e) + cy * (s + n) + cz * (bo + t)) - a.b[c]) * idg;
}

// four consecutive packed values, one 16-byte (FP32) or two (FP64) loads;
// the arrays start 256-byte aligned and a quad starts at a multiple of 4
template <typename T>
struct alignas(4 * sizeof(T)) PkQuad {
    T v[4];
};
// two consecutive packed values of one colour (an aligned pair of a quad's
// worth of cells: four consecutive cells of a row hold two of each colour)
template <typename T>
struct alignas(2 * sizeof(T)) PkPair {
    T v[2];
};

// pk_relax for the four colour-`color` cells at packed positions 4 hq ..
// 4 hq + 3 of row (j, k) (row length prow a multiple of 4), plain, zero
// initial guess or wall prolongation modes: the same value per cell, in the
// same operand order.
// The four cells share the row's parity, so their x neighbours are the
// other colour's quad at the same positions plus one value beyond it (before
// it on even rows, after it on odd ones); y and z neighbours are quads of
// the neighbouring rows and planes.  Only the first and the last quad of a
// row can touch an x side.
template <typename T, int MODE, bool CP>
KSGF_HD void pk_relax4(const PkArgs<T>& a, int color, int hq, int j, int k)
{
    static_assert(MODE != kPkProlong, "the on-the-fly prolongation has no four-cell form");
    const LevelGeom& L = a.L;
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    const int h0 = 4 * hq;
    const int row = k * L.pplane + j * L.prow;
    const int c = color * L.pcol + row + h0;
    const PkQuad<T> bq = *(const PkQuad<T>*)(a.b + c);
    PkQuad<T> out;
    if (MODE == kPkZero && color == 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m) out.v[m] = ((T)0 - bq.v[m]) * idg;
        *(PkQuad<T>*)(a.x + c) = out;
        return;
    }
    const int par = (j + k + color) & 1;
    const bool px = L.lo[0] == KSGF_LO_PERIODIC, py = L.lo[1] == KSGF_LO_PERIODIC, pz = L.lo[2] == KSGF_LO_PERIODIC;
    // x sides: cell 0 of the first quad of an even row is i = 0, cell 3 of
    // the last quad of an odd row is i = n0 - 1
    const bool in_w = !(h0 == 0 && par == 0) || px, in_e = !(h0 + 4 == L.prow && par == 1) || px;
    const bool in_s = j > 0 || py, in_n = j < L.n1 - 1 || py;
    const bool in_b = k > 0 || pz, in_t = k < L.n2 - 1 || pz;
    PkQuad<T> self{};
    if (MODE != kPkZero && !(in_w && in_e && in_s && in_n && in_b && in_t)) {
        self = *(const PkQuad<T>*)(a.x + c);
        if (MODE == kPkProlongWalls)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                self.v[m] = self.v[m] + a.xc[LvIdx<CP>(a.C, (2 * (h0 + m) + par) >> a.s0, j >> a.s1, k >> a.s2)];
    }
    const T* ob = a.x + (1 - color) * L.pcol;
    const PkQuad<T> o = *(const PkQuad<T>*)(ob + row + h0);
    // the x neighbour beyond the quad (wrapped at the row ends)
    const int hx = par == 0 ? (h0 > 0 ? h0 - 1 : L.prow - 1) : (h0 + 4 < L.prow ? h0 + 4 : 0);
    const T xb = ob[row + hx];
    PkQuad<T> w, e;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        w.v[m] = par == 0 ? (m == 0 ? xb : o.v[m - 1]) : o.v[m];
        e.v[m] = par == 0 ? o.v[m] : (m == 3 ? xb : o.v[m + 1]);
    }
    if (!in_w) w.v[0] = pk_wall<T>(L.lo[0]) * self.v[0];
    if (!in_e) e.v[3] = pk_wall<T>(L.hi[0]) * self.v[3];
    auto quad = [&](bool in, int jj, int kk, int bc) {
        PkQuad<T> q;
        if (in) {
            q = *(const PkQuad<T>*)(ob + kk * L.pplane + jj * L.prow + h0);
        } else {
#pragma unroll
            for (int m = 0; m < 4; ++m) q.v[m] = pk_wall<T>(bc) * self.v[m];
        }
        return q;
    };
    const PkQuad<T> s = quad(in_s, j > 0 ? j - 1 : L.n1 - 1, k, L.lo[1]);
    const PkQuad<T> n = quad(in_n, j < L.n1 - 1 ? j + 1 : 0, k, L.hi[1]);
    const PkQuad<T> bo = quad(in_b, j, k > 0 ? k - 1 : L.n2 - 1, L.lo[2]);
    const PkQuad<T> t = quad(in_t, j, k < L.n2 - 1 ? k + 1 : 0, L.hi[2]);
#pragma unroll
    for (int m = 0; m < 4; ++m)
        out.v[m] = ((cx * (w.v[m] + e.v[m]) + cy * (s.v[m] + n.v[m]) + cz * (bo.v[m] + t.v[m])) - bq.v[m]) * idg;
    *(PkQuad<T>*)(a.x + c) = out;
}

// x += coarse correction of the parent, for the colour-`color` cell at
// packed position (h, j, k) (the same sum the on-the-fly prolongation forms)
template <typename T, bool CP>
KSGF_HD void pk_prolong_color(const PkArgs<T>& a, int color, int h, int j, int k)
{
    const LevelGeom& L = a.L;
    const int i = 2 * h + ((j + k + color) & 1);
    const int c = color * L.pcol + k * L.pplane + j * L.prow + h;
    a.x[c] = a.x[c] + a.xc[LvIdx<CP>(a.C, i >> a.s0, j >> a.s1, k >> a.s2)];
}

// pk_prolong_color for the four cells at packed positions 4 hq .. 4 hq + 3
template <typename T, bool CP>
KSGF_HD void pk_prolong_color4(const PkArgs<T>& a, int color, int hq, int j, int k)
{
    const LevelGeom& L = a.L;
    const int h0 = 4 * hq, par = (j + k + color) & 1;
    const int c = color * L.pcol + k * L.pplane + j * L.prow + h0;
    PkQuad<T> x = *(const PkQuad<T>*)(a.x + c);
#pragma unroll
    for (int m = 0; m < 4; ++m)
        x.v[m] = x.v[m] + a.xc[LvIdx<CP>(a.C, (2 * (h0 + m) + par) >> a.s0, j >> a.s1, k >> a.s2)];
    *(PkQuad<T>*)(a.x + c) = x;
}

// residual of the fine children of coarse cell (I, J, K) of a packed fine
// level, averaged: sum of (b - (stencil(x) - diag*x)) times w, the oracle's
// summation order (i fastest); walls mirror the cell, periodic sides wrap
template <typename T>
KSGF_HD T pk_restrict_cell(const T* x, const T* b, const LevelGeom& F, int I, int J, int K, int r0, int r1, int r2,
                           T w)
{
    const T cx = (T)F.cx, cy = (T)F.cy, cz = (T)F.cz, dg = (T)F.dg;
    const bool px = F.lo[0] == KSGF_LO_PERIODIC, py = F.lo[1] == KSGF_LO_PERIODIC, pz = F.lo[2] == KSGF_LO_PERIODIC;
    T s = (T)0;
    for (int dk = 0; dk < r2; ++dk)
        for (int dj = 0; dj < r1; ++dj)
            for (int di = 0; di < r0; ++di) {
                const int i = I * r0 + di, j = J * r1 + dj, k = K * r2 + dk;
                const int col = (i + j + k) & 1, hh = i >> 1;
                const int row = k * F.pplane + j * F.prow;
                const int c = col * F.pcol + row + hh;
                const int o = (1 - col) * F.pcol + row;  // other colour, same row
                const T self = x[c];
                const T vw = (i > 0 || px) ? x[o + ((i > 0 ? i - 1 : F.n0 - 1) >> 1)] : pk_wall<T>(F.lo[0]) * self;
                const T ve = (i < F.n0 - 1 || px) ? x[o + ((i < F.n0 - 1 ? i + 1 : 0) >> 1)] : pk_wall<T>(F.hi[0]) * self;
                const T vs = (j > 0 || py) ? x[o + hh + (j > 0 ? -F.prow : (F.n1 - 1) * F.prow)] : pk_wall<T>(F.lo[1]) * self;
                const T vn = (j < F.n1 - 1 || py) ? x[o + hh + (j < F.n1 - 1 ? F.prow : -(F.n1 - 1) * F.prow)]
                                                  : pk_wall<T>(F.hi[1]) * self;
                const T vb = (k > 0 || pz) ? x[o + hh + (k > 0 ? -F.pplane : (F.n2 - 1) * F.pplane)]
                                           : pk_wall<T>(F.lo[2]) * self;
                const T vt = (k < F.n2 - 1 || pz) ? x[o + hh + (k < F.n2 - 1 ? F.pplane : -(F.n2 - 1) * F.pplane)]
                                                  : pk_wall<T>(F.hi[2]) * self;
                const T st = cx * (vw + ve) + cy * (vs + vn) + cz * (vb + vt);
                const T v = b[c] - (st - dg * self);
                s += v;
            }
    return s * w;
}

// pk_restrict_cell for the four coarse cells I0 .. I0 + 3 of row (J, K)
// when the x ratio is 1 (semicoarsening in y / z; n0 a multiple of 4): the
// children of the four are the same four-cell runs of the fine rows, read as
// colour pairs, with single loads for the x neighbours beyond a run; the
// same per-child values, summed in the same order
template <typename T>
KSGF_HD void pk_restrict_quad(const T* x, const T* b, const LevelGeom& F, int I0, int J, int K, int r1, int r2, T w,
                              T* out)
{
    const T cx = (T)F.cx, cy = (T)F.cy, cz = (T)F.cz, dg = (T)F.dg;
    const bool px = F.lo[0] == KSGF_LO_PERIODIC, py = F.lo[1] == KSGF_LO_PERIODIC, pz = F.lo[2] == KSGF_LO_PERIODIC;
    const int h0 = I0 >> 1;
    // the four cells i = I0 .. I0 + 3 of row (j, k) of array a
    auto run = [&](const T* a, int j, int k, T* v) {
        const int rp = (I0 + j + k) & 1, row = k * F.pplane + j * F.prow + h0;
        const PkPair<T> p = *(const PkPair<T>*)(a + rp * F.pcol + row);
        const PkPair<T> q = *(const PkPair<T>*)(a + (1 - rp) * F.pcol + row);
        v[0] = p.v[0];
        v[1] = q.v[0];
        v[2] = p.v[1];
        v[3] = q.v[1];
    };
    const bool in_w = I0 > 0 || px, in_e = I0 + 4 < F.n0 || px;
    const int iw = I0 > 0 ? I0 - 1 : F.n0 - 1, ie = I0 + 4 < F.n0 ? I0 + 4 : 0;
    T s[4] = {(T)0, (T)0, (T)0, (T)0};
    for (int dk = 0; dk < r2; ++dk)
        for (int dj = 0; dj < r1; ++dj) {
            const int j = J * r1 + dj, k = K * r2 + dk;
            T self[4], bb[4], vs[4], vn[4], vb[4], vt[4];
            run(x, j, k, self);
            run(b, j, k, bb);
            const T W = in_w ? x[PIdx(F, iw, j, k)] : pk_wall<T>(F.lo[0]) * self[0];
            const T E = in_e ? x[PIdx(F, ie, j, k)] : pk_wall<T>(F.hi[0]) * self[3];
            auto side = [&](bool in, int jj, int kk, int bc, T* v) {
                if (in) {
                    run(x, jj, kk, v);
                } else {
                    for (int m = 0; m < 4; ++m) v[m] = pk_wall<T>(bc) * self[m];
                }
            };
            side(j > 0 || py, j > 0 ? j - 1 : F.n1 - 1, k, F.lo[1], vs);
            side(j < F.n1 - 1 || py, j < F.n1 - 1 ? j + 1 : 0, k, F.hi[1], vn);
            side(k > 0 || pz, j, k > 0 ? k - 1 : F.n2 - 1, F.lo[2], vb);
            side(k < F.n2 - 1 || pz, j, k < F.n2 - 1 ? k + 1 : 0, F.hi[2], vt);
            for (int m = 0; m < 4; ++m) {
                const T vw = m == 0 ? W : self[m - 1];
                const T ve = m == 3 ? E : self[m + 1];
                const T st = cx * (vw + ve) + cy * (vs[m] + vn[m]) + cz * (vb[m] + vt[m]);
                const T v = bb[m] - (st - dg * self[m]);
                s[m] += v;
            }
        }
    for (int m = 0; m < 4; ++m) out[m] = s[m] * w;
}

// residual of the fine children of coarse cell (I, J, K) of a natural fine
// level; `st` evaluates the fine stencil with boundary handling
template <typename T, typename Stencil>
KSGF_HD T restrict_cell(const T* x, const T* b, const LevelGeom& F, int I, int J, int K, int r0, int r1, int r2, T w,
                        Stencil st)
{
    const T dg = (T)F.dg;
    T s = (T)0;
    for (int dk = 0; dk < r2; ++dk)
        for (int dj = 0; dj < r1; ++dj)
            for (int di = 0; di < r0; ++di) {
                const int i = I * r0 + di, j = J * r1 + dj, k = K * r2 + dk;
                const long long c = LIdx(F, i, j, k);
                const T v = b[c] - (st(c, i, j, k, x[c]) - dg * x[c]);
                s += v;
            }
    return s * w;
}

}  // namespace ksgf
