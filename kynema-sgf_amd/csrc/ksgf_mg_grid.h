// Multigrid level layout and the per-thread bodies of the grid-level
// multigrid kernels (fused red-black Gauss-Seidel sweep, residual +
// restriction, FP64 outer residual).
//
// The bodies are written as plain functions of (workgroup, thread) so that
// they compile both into the HIP kernels and into host C++: a CPU emulation
// test runs them thread by thread (barriers become loop boundaries) against
// a straightforward half-sweep implementation, checking every index and
// every bit before the kernels run on a GPU.
#pragma once

#include "../../include/ksgf.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define KSGF_HD __host__ __device__ __forceinline__
#else
#define KSGF_HD inline
#endif

namespace ksgf {

// One multigrid level: a domain-wide array with one ghost layer, Fortran
// order.  Cell (i, j, k) lives at LIdx(); cell i = 0 of every row sits kXOff
// elements into the row and the row pitch sy is a multiple of 16, so rows
// start 64-byte aligned (FP32) and an even i starts an aligned pair.
struct LevelGeom {
    int n0, n1, n2;
    long long sy, sz, tot;
    double cx, cy, cz, dg, idg;
    int lo[3], hi[3];  // LinOp BCs
    int ghost_mode;    // 1: explicit ghost fills before every colour
};

constexpr int kXOff = 16;

KSGF_HD long long LIdx(const LevelGeom& L, int i, int j, int k)
{
    return (long long)(i + kXOff) + L.sy * (j + 1) + L.sz * (k + 1);
}

inline void level_layout(LevelGeom& L)
{
    L.sy = ((long long)L.n0 + kXOff + 1 + 15) / 16 * 16;
    L.sz = L.sy * (L.n1 + 2);
    L.tot = L.sz * (L.n2 + 2);
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
// fused red-black sweep
// ---------------------------------------------------------------------------
// A workgroup owns a kSwTX x kSwTY column of the level and kz planes of it.
// It marches upwards: at step k it relaxes the red cells of plane k on the
// tile and a one-cell ring around it (the ring values are recomputed by every
// tile that needs them, so no tile waits for another), keeps the last three
// red planes in LDS, and relaxes the black cells of plane k-1 against them.
// The old iterate is staged plane by plane into LDS with a two-cell halo
// (four planes: k-1, k, k+1 in use, k+2 arriving) and the right-hand side
// with a one-cell halo (three planes), each loaded one step ahead into
// registers so the loads overlap the relaxation of the current plane.
// The sweep is out of place (x_in -> x_out): the ring reads old black values
// that a neighbouring tile may already have replaced in its own output.
constexpr int kSwTX = 64, kSwTY = 8;
constexpr int kSwRX = kSwTX + 2, kSwRY = kSwTY + 2;  // red planes and rhs: one-cell halo
constexpr int kSwPlane = kSwRX * kSwRY;
constexpr int kSwXX = kSwTX + 4, kSwXY = kSwTY + 4;  // old iterate: two-cell halo
constexpr int kSwXPlane = kSwXX * kSwXY;
constexpr int kSwThreads = 256;
constexpr int kSwRedPerRow = kSwRX / 2;
constexpr int kSwRed = kSwRY * kSwRedPerRow;
constexpr int kSwXPer = (kSwXPlane + kSwThreads - 1) / kSwThreads;
constexpr int kSwBPer = (kSwPlane + kSwThreads - 1) / kSwThreads;
static_assert(kSwTX / 2 * kSwTY == kSwThreads, "one black cell per thread");

// what x_in means: the previous iterate; a zero initial guess (x_in unused);
// or the previous iterate plus the piecewise-constant prolongation of the
// coarse correction (x_in + xc(parent), formed on the fly)
enum { kSwRead = 0, kSwZero = 1, kSwProlong = 2 };

template <typename T>
struct SweepArgs {
    const T* xin;
    T* xout;
    const T* b;
    const T* xc;
    LevelGeom L, C;
    int s0, s1, s2;  // log2 of the coarsening ratio to C
    int kz;          // planes per workgroup
    int ntx, nty, nkc;
};

template <typename T>
inline int sweep_blocks(SweepArgs<T>& a, int kz)
{
    a.kz = kz;
    a.ntx = (a.L.n0 + kSwTX - 1) / kSwTX;
    a.nty = (a.L.n1 + kSwTY - 1) / kSwTY;
    a.nkc = (a.L.n2 + kz - 1) / kz;
    return a.ntx * a.nty * a.nkc;
}

KSGF_HD void sweep_tile(int lid, int ntx, int nty, int& tx, int& ty, int& kc)
{
    tx = lid % ntx;
    const int r = lid / ntx;
    ty = r % nty;
    kc = r / nty;
}

// in-domain index of coordinate g (which lies at most `halo` cells outside
// [0, n)), wrapped across a periodic side; -1 where no value exists
KSGF_HD int sw_wrap(int g, int n, bool periodic, int halo)
{
    if (g >= 0 && g < n) return g;
    if (!periodic || g < -halo || g >= n + halo) return -1;
    return (g < 0) ? g + n : g - n;
}

// Per-thread data of a sweep workgroup that does not change from plane to
// plane: in-plane element offsets of the staged loads, and the LDS positions
// and neighbour flags of the cells the thread relaxes (for both parities of
// the plane index, which decides where the red cells of a row sit).
// Computed once per workgroup, so a staged load costs one add per plane.
struct SwThread {
    int xoff[kSwXPer], xcoff[kSwXPer];  // fine / coarse in-plane offsets of the old-iterate positions
    int boff[kSwBPer];                  // in-plane offsets of the rhs positions
    unsigned xmask, bmask;              // bit u: position u holds a value
    int rq[2][2], rx[2][2];             // red phase [round][plane parity]: one-/two-cell-halo LDS index, -1: none
    unsigned rflag[2][2];               // bit 0..3: W, E, S, N neighbour inside the domain or across a periodic side
    int bq[2], bx[2], bred[2];          // black phase [plane parity]: black cell (one-/two-cell halo), red cell
    unsigned bflag[2];
    int at0;                            // in-plane offset of the first cell of the thread's pair
};

template <typename T>
KSGF_HD void sw_thread_init(const SweepArgs<T>& a, SwThread& th, int tid, int tx, int ty)
{
    const LevelGeom& L = a.L;
    const bool px = L.lo[0] == KSGF_LO_PERIODIC, py = L.lo[1] == KSGF_LO_PERIODIC;
    const int sy = (int)L.sy, csy = (int)a.C.sy;
    th.xmask = 0;
    for (int u = 0; u < kSwXPer; ++u) {
        const int p = tid + u * kSwThreads;
        th.xoff[u] = 0;
        th.xcoff[u] = 0;
        if (p >= kSwXPlane) continue;
        const int xj = p / kSwXX, xi = p - xj * kSwXX;
        const int wi = sw_wrap(tx * kSwTX + xi - 2, L.n0, px, 2), wj = sw_wrap(ty * kSwTY + xj - 2, L.n1, py, 2);
        if (wi < 0 || wj < 0) continue;
        th.xmask |= 1u << u;
        th.xoff[u] = wi + kXOff + sy * (wj + 1);
        th.xcoff[u] = (wi >> a.s0) + kXOff + csy * ((wj >> a.s1) + 1);
    }
    th.bmask = 0;
    for (int u = 0; u < kSwBPer; ++u) {
        const int p = tid + u * kSwThreads;
        th.boff[u] = 0;
        if (p >= kSwPlane) continue;
        const int rj = p / kSwRX, ri = p - rj * kSwRX;
        const int wi = sw_wrap(tx * kSwTX + ri - 1, L.n0, px, 1), wj = sw_wrap(ty * kSwTY + rj - 1, L.n1, py, 1);
        if (wi < 0 || wj < 0) continue;
        th.bmask |= 1u << u;
        th.boff[u] = wi + kXOff + sy * (wj + 1);
    }
    for (int r = 0; r < 2; ++r)
        for (int par = 0; par < 2; ++par) {
            const int q = tid + r * kSwThreads;
            th.rq[r][par] = -1;
            th.rx[r][par] = 0;
            th.rflag[r][par] = 0;
            if (q >= kSwRed) continue;
            const int rj = q / kSwRedPerRow, m = q - rj * kSwRedPerRow;
            const int gj = ty * kSwTY + rj - 1;
            const int ri = 2 * m + ((1 + gj + par) & 1);
            const int gi = tx * kSwTX + ri - 1;
            if (sw_wrap(gi, L.n0, px, 1) < 0 || sw_wrap(gj, L.n1, py, 1) < 0) continue;
            th.rq[r][par] = rj * kSwRX + ri;
            th.rx[r][par] = (rj + 1) * kSwXX + ri + 1;
            th.rflag[r][par] = (unsigned)(gi > 0 || px) | (unsigned)(gi < L.n0 - 1 || px) << 1 |
                               (unsigned)(gj > 0 || py) << 2 | (unsigned)(gj < L.n1 - 1 || py) << 3;
        }
    const int row = tid / (kSwTX / 2), m = tid - row * (kSwTX / 2);
    const int gj = ty * kSwTY + row, gi0 = tx * kSwTX + 2 * m;
    th.at0 = gi0 + kXOff + sy * (gj + 1);
    for (int par = 0; par < 2; ++par) {
        th.bq[par] = -1;
        th.bx[par] = 0;
        th.bred[par] = 0;
        th.bflag[par] = 0;
        if (gj >= L.n1 || gi0 >= L.n0) continue;
        const bool red_first = ((gi0 + gj + par) & 1) == 0;
        const int gb = red_first ? gi0 + 1 : gi0;
        const int lb = gb - tx * kSwTX + 1;
        th.bq[par] = (row + 1) * kSwRX + lb;
        th.bx[par] = (row + 2) * kSwXX + lb + 1;
        th.bred[par] = (row + 1) * kSwRX + ((red_first ? gi0 : gi0 + 1) - tx * kSwTX + 1);
        th.bflag[par] = (unsigned)(gb > 0 || px) | (unsigned)(gb < L.n0 - 1 || px) << 1 |
                        (unsigned)(gj > 0 || py) << 2 | (unsigned)(gj < L.n1 - 1 || py) << 3 |
                        (unsigned)red_first << 4;
    }
}

// loads of plane k of the old iterate (two-cell halo) into registers
template <typename T, int MODE>
KSGF_HD void sw_load_x(const SweepArgs<T>& a, const SwThread& th, T* v, int k)
{
    const LevelGeom& L = a.L;
    const int kk = sw_wrap(k, L.n2, L.lo[2] == KSGF_LO_PERIODIC, 2);
    for (int u = 0; u < kSwXPer; ++u) v[u] = (T)0;
    if (MODE == kSwZero || kk < 0) return;
    const T* base = a.xin + L.sz * (kk + 1);
    const T* cbase = (MODE == kSwProlong) ? a.xc + a.C.sz * ((kk >> a.s2) + 1) : nullptr;
    for (int u = 0; u < kSwXPer; ++u) {
        if (!((th.xmask >> u) & 1u)) continue;
        if (MODE == kSwProlong)
            v[u] = base[th.xoff[u]] + cbase[th.xcoff[u]];
        else
            v[u] = base[th.xoff[u]];
    }
}

template <typename T>
KSGF_HD void sw_store_x(T* X, const T* v, int tid)
{
    for (int u = 0; u < kSwXPer; ++u) {
        const int p = tid + u * kSwThreads;
        if (p < kSwXPlane) X[p] = v[u];
    }
}

// loads of plane k of the right-hand side (one-cell halo) into registers
template <typename T>
KSGF_HD void sw_load_b(const SweepArgs<T>& a, const SwThread& th, T* v, int k)
{
    const LevelGeom& L = a.L;
    const int kk = sw_wrap(k, L.n2, L.lo[2] == KSGF_LO_PERIODIC, 1);
    for (int u = 0; u < kSwBPer; ++u) v[u] = (T)0;
    if (kk < 0) return;
    const T* base = a.b + L.sz * (kk + 1);
    for (int u = 0; u < kSwBPer; ++u)
        if ((th.bmask >> u) & 1u) v[u] = base[th.boff[u]];
}

template <typename T>
KSGF_HD void sw_store_b(T* B, const T* v, int tid)
{
    for (int u = 0; u < kSwBPer; ++u) {
        const int p = tid + u * kSwThreads;
        if (p < kSwPlane) B[p] = v[u];
    }
}

// wall factor of a side: a Neumann wall mirrors the cell (+1), a Dirichlet
// wall takes its negative (-1); multiplying by it is exact
template <typename T>
KSGF_HD T sw_wall(int bc)
{
    return bc == KSGF_LO_NEUMANN ? (T)1 : (T)-1;
}

// red half of step k: every red position of the tile + ring in plane k that
// is inside the domain, or outside it across a periodic side (then the value
// of the wrapped cell), relaxed against the staged old planes xm, x0, xp
// (k-1, k, k+1) and rhs plane b0, goes into the red plane:
// (cx*(W+E) + cy*(S+N) + cz*(B+T) - b) / diag, the oracle's operand order
template <typename T>
KSGF_HD void sw_phase_red(const LevelGeom& L, const SwThread& th, const T* xm, const T* x0, const T* xp, const T* b0,
                          T* red, int k)
{
    const bool pz = L.lo[2] == KSGF_LO_PERIODIC;
    if (sw_wrap(k, L.n2, pz, 1) < 0) return;
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    const T fw = sw_wall<T>(L.lo[0]), fe = sw_wall<T>(L.hi[0]), fs = sw_wall<T>(L.lo[1]), fn = sw_wall<T>(L.hi[1]);
    const T fb = sw_wall<T>(L.lo[2]), ft = sw_wall<T>(L.hi[2]);
    const bool zb = (k > 0 || pz), zt = (k < L.n2 - 1 || pz);
    const int par = k & 1;
    for (int r = 0; r < 2; ++r) {
        const int q = th.rq[r][par];
        if (q < 0) continue;
        const int xq = th.rx[r][par];
        const unsigned f = th.rflag[r][par];
        const T self = x0[xq];
        const T w = (f & 1u) ? x0[xq - 1] : fw * self;
        const T e = (f & 2u) ? x0[xq + 1] : fe * self;
        const T s = (f & 4u) ? x0[xq - kSwXX] : fs * self;
        const T n = (f & 8u) ? x0[xq + kSwXX] : fn * self;
        const T bo = zb ? xm[xq] : fb * self;
        const T t = zt ? xp[xq] : ft * self;
        red[q] = ((cx * (w + e) + cy * (s + n) + cz * (bo + t)) - b0[q]) * idg;
    }
}

// black half of plane kb (red planes kb-1, kb, kb+1 in pm, p0, pp; staged
// old plane x0 and rhs plane b0 of kb): each thread relaxes one black cell
// and returns the aligned pair (red, black in x order) that holds it; false
// when the pair lies outside the level
template <typename T>
KSGF_HD bool sw_phase_black(const LevelGeom& L, const SwThread& th, const T* pm, const T* p0, const T* pp, const T* x0,
                            const T* b0, int kb, long long& at, T& v0, T& v1)
{
    const int par = kb & 1;
    const int rb = th.bq[par];
    if (rb < 0) return false;
    const bool pz = L.lo[2] == KSGF_LO_PERIODIC;
    const unsigned f = th.bflag[par];
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    const T self = x0[th.bx[par]];
    const T w = (f & 1u) ? p0[rb - 1] : sw_wall<T>(L.lo[0]) * self;
    const T e = (f & 2u) ? p0[rb + 1] : sw_wall<T>(L.hi[0]) * self;
    const T s = (f & 4u) ? p0[rb - kSwRX] : sw_wall<T>(L.lo[1]) * self;
    const T n = (f & 8u) ? p0[rb + kSwRX] : sw_wall<T>(L.hi[1]) * self;
    const T bo = (kb > 0 || pz) ? pm[rb] : sw_wall<T>(L.lo[2]) * self;
    const T t = (kb < L.n2 - 1 || pz) ? pp[rb] : sw_wall<T>(L.hi[2]) * self;
    const T blk = ((cx * (w + e) + cy * (s + n) + cz * (bo + t)) - b0[rb]) * idg;
    const T red = p0[th.bred[par]];
    at = (long long)th.at0 + L.sz * (kb + 1);
    const bool red_first = (f >> 4) & 1u;
    v0 = red_first ? red : blk;
    v1 = red_first ? blk : red;
    return true;
}

// ---------------------------------------------------------------------------
// row tiles: 64 cells in x by 4 rows in y, kz planes, 256 threads
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

// residual of the fine children of coarse cell (I, J, K), averaged; `st`
// evaluates the fine stencil with boundary handling
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
