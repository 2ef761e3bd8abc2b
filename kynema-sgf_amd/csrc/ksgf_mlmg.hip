// Hydro::MacProjector + amrex::MLMG (MLPoisson) equivalent on one GPU.
//
// The projector owns its multigrid hierarchy laid out domain-wide: one
// contiguous array per level with a single ghost layer, Fortran order.  On a
// single GPU the BoxArray is only a storage decomposition, and red-black
// Gauss-Seidel with a 7-point stencil, 8-cell averaging and piecewise
// constant prolongation are all independent of it, so the domain-wide layout
// gives the same iterates as a per-box hierarchy while every sweep is one
// coalesced launch.  The MAC divergence is gathered from the box FABs into
// level 0 and grad(phi) is scattered back onto every valid face.
//
// Options and semantics follow MLMGOptions.{H,cpp} (maxiter 200, nu1 = nu2
// = 2, final smooth 8, bottom smooth 0, BiCGStab bottom with rtol eps*1e12,
// atol eps*1e4, 200 iterations, no abort on non-convergence) and the solve
// converges when |r|_inf <= max(atol, rtol*|rhs|_inf).
//
// Kernel organisation of one V-cycle on the large levels:
//   down:  zero + first red half sweep fused; remaining half sweeps;
//          residual fused with the 8-cell (or semicoarsened 2/4-cell)
//          average restriction (no residual array is written);
//   up:    piecewise-constant prolongation fused with the first red half
//          sweep; remaining half sweeps.
// Domain boundaries are handled in the stencil itself (periodic wrap,
// Neumann mirror, Dirichlet odd reflection), so no ghost-fill launches are
// needed; a level with an odd periodic extent (where wrap-around partners
// share a colour) falls back to explicit ghost fills.  All levels with at
// most kCoarseCells cells are handled by a single workgroup that runs the
// rest of the V-cycle, including the BiCGStab bottom solve, in one launch.
//
// Mixed precision (KSGF_MG_MIXED): each outer iteration forms the FP64
// residual of the FP64 solution (fused with adding the previous correction),
// runs one V-cycle on it with FP32 smoothing and grid transfers (the bottom
// BiCGStab stays FP64), and adds the FP32 correction to the FP64 solution.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <cstdlib>

#include <hip/hip_ext.h>

#include "ksgf_internal.h"
#include "ksgf_mg_grid.h"

namespace ksgf {

// a multigrid iteration whose residual exceeds this multiple of |rhs|_inf
// (or is not finite) is diverging and stops
constexpr double kDivergeFactor = 1.0e20;

constexpr long long kCoarseCells = 16384;
constexpr int kCoarseThreads = 1024;

template <typename T>
__device__ __forceinline__ T bc_ghost(int t, T near_v, T far_v)
{
    return (t == KSGF_LO_PERIODIC) ? far_v : (t == KSGF_LO_NEUMANN ? near_v : -near_v);
}

// value of the low / high neighbour of cell c along one direction; `self`
// is the cell's own value as a Neumann / Dirichlet boundary mirrors it
template <typename T>
__device__ __forceinline__ T nb_lo(const T* __restrict__ a, long long c, int pos, int n, long long s, int bc, int gm,
                                   T self)
{
    if (gm || pos > 0) return a[c - s];
    return bc_ghost<T>(bc, self, a[c + (long long)(n - 1) * s]);
}
template <typename T>
__device__ __forceinline__ T nb_hi(const T* __restrict__ a, long long c, int pos, int n, long long s, int bc, int gm,
                                   T self)
{
    if (gm || pos < n - 1) return a[c + s];
    return bc_ghost<T>(bc, self, a[c - (long long)(n - 1) * s]);
}

// cx*(W+E) + cy*(S+N) + cz*(B+T), the same operand order as the oracle
template <typename T>
__device__ __forceinline__ T stencil(const T* __restrict__ a, long long c, int i, int j, int k, const LevelGeom& L, T cx,
                                     T cy, T cz, T self)
{
    const int gm = L.ghost_mode;
    T w = nb_lo<T>(a, c, i, L.n0, 1, L.lo[0], gm, self), e = nb_hi<T>(a, c, i, L.n0, 1, L.hi[0], gm, self);
    T s = nb_lo<T>(a, c, j, L.n1, L.sy, L.lo[1], gm, self), n = nb_hi<T>(a, c, j, L.n1, L.sy, L.hi[1], gm, self);
    T b = nb_lo<T>(a, c, k, L.n2, L.sz, L.lo[2], gm, self), t = nb_hi<T>(a, c, k, L.n2, L.sz, L.hi[2], gm, self);
    return cx * (w + e) + cy * (s + n) + cz * (b + t);
}

// ---------------------------------------------------------------------------
// ghost layer (only for ghost_mode levels, the coarse part and phi)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void ghost_point(T* __restrict__ a, const LevelGeom& L, int t)
{
    const int nx = L.n1 * L.n2, ny = L.n0 * L.n2, nz = L.n0 * L.n1;
    if (t < 2 * nx) {
        const int side = t / nx, r = t - side * nx;
        const int j = r % L.n1, k = r / L.n1;
        T vl = a[LIdx(L, 0, j, k)], vr = a[LIdx(L, L.n0 - 1, j, k)];
        if (side == 0)
            a[LIdx(L, -1, j, k)] = bc_ghost<T>(L.lo[0], vl, vr);
        else
            a[LIdx(L, L.n0, j, k)] = bc_ghost<T>(L.hi[0], vr, vl);
        return;
    }
    t -= 2 * nx;
    if (t < 2 * ny) {
        const int side = t / ny, r = t - side * ny;
        const int i = r % L.n0, k = r / L.n0;
        T vl = a[LIdx(L, i, 0, k)], vr = a[LIdx(L, i, L.n1 - 1, k)];
        if (side == 0)
            a[LIdx(L, i, -1, k)] = bc_ghost<T>(L.lo[1], vl, vr);
        else
            a[LIdx(L, i, L.n1, k)] = bc_ghost<T>(L.hi[1], vr, vl);
        return;
    }
    t -= 2 * ny;
    if (t < 2 * nz) {
        const int side = t / nz, r = t - side * nz;
        const int i = r % L.n0, j = r / L.n0;
        T vl = a[LIdx(L, i, j, 0)], vr = a[LIdx(L, i, j, L.n2 - 1)];
        if (side == 0)
            a[LIdx(L, i, j, -1)] = bc_ghost<T>(L.lo[2], vl, vr);
        else
            a[LIdx(L, i, j, L.n2)] = bc_ghost<T>(L.hi[2], vr, vl);
    }
}

__host__ __device__ __forceinline__ long long ghost_count(const LevelGeom& L)
{
    return 2 * ((long long)L.n1 * L.n2 + (long long)L.n0 * L.n2 + (long long)L.n0 * L.n1);
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_ghosts(T* __restrict__ a, LevelGeom L)
{
    if (L.act && *L.act == 0) return;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < (int)ghost_count(L)) ghost_point<T>(a, L, t);
}

// ---------------------------------------------------------------------------
// grid-level smoothing / transfer kernels
// ---------------------------------------------------------------------------
// one red (color 0) or black (color 1) Gauss-Seidel half sweep.  With xc,
// this is the black half sweep right after the fused prolongation + red
// sweep: black cells still hold their pre-prolongation value, so the
// value a wall mirrors is x + xc(parent).
template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth(T* __restrict__ x, const T* __restrict__ b, LevelGeom L, int color,
                                                   const T* __restrict__ xc, LevelGeom C, int r0, int r1, int r2)
{
    if (L.act && *L.act == 0) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h0 = (L.n0 + 1) / 2;
    const long long tot = (long long)h0 * L.n1 * L.n2;
    if (t >= tot) return;
    const unsigned tt = (unsigned)t, r = tt / (unsigned)h0;
    int ih = (int)(tt - r * (unsigned)h0);
    int j = (int)(r % (unsigned)L.n1), k = (int)(r / (unsigned)L.n1);
    int i = 2 * ih + ((j + k + color) & 1);
    if (i >= L.n0) return;
    long long c = LIdx(L, i, j, k);
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    T self = x[c];
    if (xc) self = self + xc[LIdx(C, i / r0, j / r1, k / r2)];
    x[c] = (stencil<T>(x, c, i, j, k, L, cx, cy, cz, self) - b[c]) * idg;
}

// zero initial guess + the first red half sweep: red = (0 - b)/diag, black = 0
template <typename T>
__global__ void __launch_bounds__(256) k_mg_smooth_first(T* __restrict__ x, const T* __restrict__ b, LevelGeom L)
{
    if (L.act && *L.act == 0) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)L.n0 * L.n1 * L.n2;
    if (t >= tot) return;
    const unsigned tt = (unsigned)t, q = tt / (unsigned)L.n0;
    int i = (int)(tt - q * (unsigned)L.n0);
    int j = (int)(q % (unsigned)L.n1), k = (int)(q / (unsigned)L.n1);
    long long c = LIdx(L, i, j, k);
    x[c] = (((i + j + k) & 1) == 0) ? ((T)0 - b[c]) * (T)L.idg : (T)0;
}

// piecewise-constant prolongation fused with the first red half sweep: red
// cells are relaxed against their prolonged black neighbours (x + xc(parent),
// formed on the fly).  Black cells are not written: the black half sweep
// that follows overwrites them without reading their own value, so their
// prolonged value is never needed (and writing it here would race with the
// red cells reading the same neighbours).
template <typename T>
__device__ __forceinline__ T prolonged(const T* __restrict__ x, const T* __restrict__ xc, const LevelGeom& F,
                                       const LevelGeom& C, int i, int j, int k, int r0, int r1, int r2)
{
    return x[LIdx(F, i, j, k)] + xc[LIdx(C, i / r0, j / r1, k / r2)];
}

template <typename T>
__device__ __forceinline__ T pnb(const T* __restrict__ x, const T* __restrict__ xc, const LevelGeom& F,
                                 const LevelGeom& C, int i, int j, int k, int d, int s, int r0, int r1, int r2)
{
    int p[3] = {i, j, k};
    const int n[3] = {F.n0, F.n1, F.n2};
    int q = p[d] + s;
    if (q >= 0 && q < n[d]) {
        p[d] = q;
        return prolonged<T>(x, xc, F, C, p[0], p[1], p[2], r0, r1, r2);
    }
    const int bc = (s < 0) ? F.lo[d] : F.hi[d];
    if (bc == KSGF_LO_PERIODIC) {
        p[d] = (s < 0) ? n[d] - 1 : 0;
        return prolonged<T>(x, xc, F, C, p[0], p[1], p[2], r0, r1, r2);
    }
    T self = prolonged<T>(x, xc, F, C, i, j, k, r0, r1, r2);
    return (bc == KSGF_LO_NEUMANN) ? self : -self;
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong_red(T* __restrict__ x, const T* __restrict__ xc,
                                                        const T* __restrict__ b, LevelGeom F, LevelGeom C, int r0,
                                                        int r1, int r2)
{
    if (F.act && *F.act == 0) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)F.n0 * F.n1 * F.n2;
    if (t >= tot) return;
    const unsigned tt = (unsigned)t, q = tt / (unsigned)F.n0;
    int i = (int)(tt - q * (unsigned)F.n0);
    int j = (int)(q % (unsigned)F.n1), k = (int)(q / (unsigned)F.n1);
    long long c = LIdx(F, i, j, k);
    if (((i + j + k) & 1) != 0) return;
    const T cx = (T)F.cx, cy = (T)F.cy, cz = (T)F.cz, idg = (T)F.idg;
    T w = pnb<T>(x, xc, F, C, i, j, k, 0, -1, r0, r1, r2), e = pnb<T>(x, xc, F, C, i, j, k, 0, 1, r0, r1, r2);
    T s = pnb<T>(x, xc, F, C, i, j, k, 1, -1, r0, r1, r2), n = pnb<T>(x, xc, F, C, i, j, k, 1, 1, r0, r1, r2);
    T bo = pnb<T>(x, xc, F, C, i, j, k, 2, -1, r0, r1, r2), to = pnb<T>(x, xc, F, C, i, j, k, 2, 1, r0, r1, r2);
    x[c] = ((cx * (w + e) + cy * (s + n) + cz * (bo + to)) - b[c]) * idg;
}

template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong_add(T* __restrict__ f, const T* __restrict__ c, LevelGeom F,
                                                        LevelGeom C, int r0, int r1, int r2)
{
    if (F.act && *F.act == 0) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)F.n0 * F.n1 * F.n2;
    if (t >= tot) return;
    const unsigned tt = (unsigned)t, q = tt / (unsigned)F.n0;
    int i = (int)(tt - q * (unsigned)F.n0);
    int j = (int)(q % (unsigned)F.n1), k = (int)(q / (unsigned)F.n1);
    f[LIdx(F, i, j, k)] += c[LIdx(C, i / r0, j / r1, k / r2)];
}

// one red (color 0) or black (color 1) half sweep of a packed level over
// row tiles of packed positions (i/2, j), kz planes per workgroup
template <typename T, int MODE, bool CP>
__global__ void __launch_bounds__(256) k_pk_half(PkArgs<T> a, int color, RowTiles t)
{
    if (a.L.act && *a.L.act == 0) return;
    int h, j, k0, k1;
    row_tile(t, xcd_remap((int)blockIdx.x, (int)gridDim.x), (int)threadIdx.x, h, j, k0, k1, a.L.n2);
    if (h >= a.L.prow || j >= a.L.n1) return;
    for (int k = k0; k < k1; ++k) pk_relax<T, MODE, CP>(a, color, h, j, k);
}

// the same half sweep (plain, zero-guess or wall-prolongation mode) and the
// colour prolongation four cells per thread: one thread per quad of a packed
// row (rows of a multiple of 4 cells), 256 consecutive (quad, row) pairs of
// one plane chunk per workgroup, kz planes
struct QuadTiles {
    int nq, nb, kz;  // quads per row, workgroups per plane chunk, planes per workgroup
};
inline QuadTiles quad_tiles(const LevelGeom& L, int kz)
{
    QuadTiles t;
    t.nq = L.prow / 4;
    t.nb = (int)(((long long)t.nq * L.n1 + 255) / 256);
    t.kz = kz;
    return t;
}
__device__ __forceinline__ bool quad_tile(const QuadTiles& t, int n1, int n2, int& hq, int& j, int& k0, int& k1)
{
    const int lid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int kc = lid / t.nb, wb = lid - kc * t.nb;
    const unsigned idx = (unsigned)wb * 256u + threadIdx.x;
    const unsigned jj = idx / (unsigned)t.nq;
    hq = (int)(idx - jj * (unsigned)t.nq);
    j = (int)jj;
    k0 = kc * t.kz;
    k1 = min(k0 + t.kz, n2);
    return j < n1;
}
template <typename T, int MODE, bool CP>
__global__ void __launch_bounds__(256) k_pk_half4(PkArgs<T> a, int color, QuadTiles t)
{
    if (a.L.act && *a.L.act == 0) return;
    int hq, j, k0, k1;
    if (!quad_tile(t, a.L.n1, a.L.n2, hq, j, k0, k1)) return;
    for (int k = k0; k < k1; ++k) pk_relax4<T, MODE, CP>(a, color, hq, j, k);
}
template <typename T, bool CP>
__global__ void __launch_bounds__(256) k_pk_prolong_color4(PkArgs<T> a, int color, QuadTiles t)
{
    if (a.L.act && *a.L.act == 0) return;
    int hq, j, k0, k1;
    if (!quad_tile(t, a.L.n1, a.L.n2, hq, j, k0, k1)) return;
    for (int k = k0; k < k1; ++k) pk_prolong_color4<T, CP>(a, color, hq, j, k);
}

// x += coarse correction on the colour-`color` cells of a packed level
template <typename T, bool CP>
__global__ void __launch_bounds__(256) k_pk_prolong_color(PkArgs<T> a, int color, RowTiles t)
{
    if (a.L.act && *a.L.act == 0) return;
    int h, j, k0, k1;
    row_tile(t, xcd_remap((int)blockIdx.x, (int)gridDim.x), (int)threadIdx.x, h, j, k0, k1, a.L.n2);
    if (h >= a.L.prow || j >= a.L.n1) return;
    for (int k = k0; k < k1; ++k) pk_prolong_color<T, CP>(a, color, h, j, k);
}

// x += prolongation of the coarse correction on a packed level (only when
// there is no post-smoothing to fuse it into)
template <typename T>
__global__ void __launch_bounds__(256) k_pk_prolong_add(T* __restrict__ x, const T* __restrict__ xc, LevelGeom F,
                                                        LevelGeom C, int s0, int s1, int s2, RowTiles t)
{
    if (F.act && *F.act == 0) return;
    int i, j, k0, k1;
    row_tile(t, xcd_remap((int)blockIdx.x, (int)gridDim.x), (int)threadIdx.x, i, j, k0, k1, F.n2);
    if (i >= F.n0 || j >= F.n1) return;
    for (int k = k0; k < k1; ++k) x[PIdx(F, i, j, k)] += xc[AnyIdx(C, i >> s0, j >> s1, k >> s2)];
}

// residual + restriction over row tiles of the coarse level; a packed fine
// level may restrict into a packed or a natural coarse level, a natural one
// only into a natural one.  cred (packed coarse level, optional): the coarse
// correction, whose red cells get the red half of a zero-initial-guess
// sweep, -res / diag, as pk_relax's zero mode forms it (so the coarse
// level's first sweep starts with its black half)
template <typename T>
__global__ void __launch_bounds__(256) k_resid_restrict_rt(const T* __restrict__ x, const T* __restrict__ b,
                                                           T* __restrict__ cres, T* __restrict__ cred, LevelGeom F,
                                                           LevelGeom C, RowTiles t, int r0, int r1, int r2, T w)
{
    if (F.act && *F.act == 0) return;
    int I, J, K0, K1;
    row_tile(t, xcd_remap((int)blockIdx.x, (int)gridDim.x), (int)threadIdx.x, I, J, K0, K1, C.n2);
    if (I >= C.n0 || J >= C.n1) return;
    const T cx = (T)F.cx, cy = (T)F.cy, cz = (T)F.cz;
    for (int K = K0; K < K1; ++K) {
        if (F.packed) {
            const long long o = AnyIdx(C, I, J, K);
            const T v = pk_restrict_cell<T>(x, b, F, I, J, K, r0, r1, r2, w);
            cres[o] = v;
            if (cred && ((I + J + K) & 1) == 0) cred[o] = ((T)0 - v) * (T)C.idg;
        } else {
            auto st = [&](long long c, int i, int j, int k, T self) {
                return stencil<T>(x, c, i, j, k, F, cx, cy, cz, self);
            };
            cres[LIdx(C, I, J, K)] = restrict_cell<T>(x, b, F, I, J, K, r0, r1, r2, w, st);
        }
    }
}

// the same over quads of coarse cells when the x ratio is 1 (n0 a multiple
// of 4, packed fine level): four coarse cells I0 .. I0 + 3 per thread
template <typename T>
__global__ void __launch_bounds__(256) k_resid_restrict4(const T* __restrict__ x, const T* __restrict__ b,
                                                         T* __restrict__ cres, T* __restrict__ cred, LevelGeom F,
                                                         LevelGeom C, RowTiles t, int r1, int r2, T w)
{
    if (F.act && *F.act == 0) return;
    int IQ, J, K0, K1;
    row_tile(t, xcd_remap((int)blockIdx.x, (int)gridDim.x), (int)threadIdx.x, IQ, J, K0, K1, C.n2);
    const int I0 = 4 * IQ;
    if (I0 >= C.n0 || J >= C.n1) return;
    for (int K = K0; K < K1; ++K) {
        T q[4];
        pk_restrict_quad<T>(x, b, F, I0, J, K, r1, r2, w, q);
        if (C.packed) {
            const int rp = (I0 + J + K) & 1, row = K * C.pplane + J * C.prow + (I0 >> 1);
            *(PkPair<T>*)(cres + rp * C.pcol + row) = PkPair<T>{{q[0], q[2]}};
            *(PkPair<T>*)(cres + (1 - rp) * C.pcol + row) = PkPair<T>{{q[1], q[3]}};
            if (cred) {
                const T ra = rp == 0 ? q[0] : q[1], rb = rp == 0 ? q[2] : q[3];
                const T idg = (T)C.idg;
                *(PkPair<T>*)(cred + row) = PkPair<T>{{((T)0 - ra) * idg, ((T)0 - rb) * idg}};
            }
        } else {
            *(PkQuad<T>*)(cres + LIdx(C, I0, J, K)) = PkQuad<T>{{q[0], q[1], q[2], q[3]}};
        }
    }
}

// block maximum of non-negative values into part[blockIdx.x]
__device__ __forceinline__ void block_max_part(double m, double* __restrict__ part)
{
    __shared__ double sh[4];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// FP64 outer residual over row tiles.  phi and rhs are FP64 in the natural
// layout P; the correction cor and the residual r are in the level-0 layout
// (packed when CP, else the natural layout P).  With a correction: phi_out =
// phi + cor and r = rhs - A(phi + cor) in one pass (ping-pong phi buffers);
// without: r = rhs - A phi.  r is written in precision R (FP32 feeds the
// mixed V-cycle directly); per-block maxima of |r| go to part.  Each thread
// marches up its column keeping phi + cor of planes k-1, k, k+1 in
// registers, so the z neighbours are not fetched again.
template <bool CP>
struct ResOffsets {
    int q, w, e, s, n;  // correction offsets: the cell and its x/y neighbours
};

template <bool CP>
__device__ __forceinline__ ResOffsets<CP> res_offsets(const LevelGeom& L, int c, int i, int j, int k, int dw, int de,
                                                      int ds, int dn, int iw, int ie)
{
    ResOffsets<CP> o;
    if (CP) {
        const int col = (i + j + k) & 1, h = i >> 1;
        const int row = k * L.pplane + j * L.prow;
        const int ob = (1 - col) * L.pcol + row;
        o.q = col * L.pcol + row + h;
        o.w = ob + (iw >> 1);
        o.e = ob + (ie >> 1);
        o.s = ob + h + ((j > 0) ? -L.prow : (L.n1 - 1) * L.prow);
        o.n = ob + h + ((j < L.n1 - 1) ? L.prow : -(L.n1 - 1) * L.prow);
    } else {
        o.q = c;
        o.w = c + dw;
        o.e = c + de;
        o.s = c + ds;
        o.n = c + dn;
    }
    return o;
}

// |v| for a residual norm; NaN counts as +inf (fmax alone would drop it)
__device__ __forceinline__ double abs_or_inf(double v)
{
    const double a = fabs(v);
    return a == a ? a : __builtin_huge_val();
}

template <typename C, typename R, bool CP>
__device__ __forceinline__ double update_residual_tile(const double* __restrict__ phi, const C* __restrict__ cor,
                                                       double* __restrict__ phi_out, const double* __restrict__ rhs,
                                                       R* __restrict__ r, const LevelGeom& P, const LevelGeom& L,
                                                       const RowTiles& tl, int lid)
{
    int i, j, k0, k1;
    row_tile(tl, lid, (int)threadIdx.x, i, j, k0, k1, P.n2);
    double m = 0.0;
    if (i < P.n0 && j < P.n1) {
        if (P.ghost_mode) {
            // explicit ghosts (odd periodic extents, natural layout): with a
            // correction only phi += cor (the caller fills the ghosts and calls
            // again without one); without, the plain residual from the ghosts
            for (int k = k0; k < k1; ++k) {
                const int c = (int)LIdx(P, i, j, k);
                double v;
                if (cor) {
                    phi_out[c] = phi[c] + (double)cor[c];
                    v = 0.0;
                } else {
                    v = rhs[c] - (stencil<double>(phi, c, i, j, k, P, P.cx, P.cy, P.cz, phi[c]) - P.dg * phi[c]);
                }
                r[c] = (R)v;
                m = fmax(m, abs_or_inf(v));
            }
        } else {
            const bool px = P.lo[0] == KSGF_LO_PERIODIC, py = P.lo[1] == KSGF_LO_PERIODIC,
                       pz = P.lo[2] == KSGF_LO_PERIODIC;
            const int sy = (int)P.sy;
            const int dw = (i > 0) ? -1 : P.n0 - 1, de = (i < P.n0 - 1) ? 1 : -(P.n0 - 1);
            const int ds = (j > 0) ? -sy : (P.n1 - 1) * sy, dn = (j < P.n1 - 1) ? sy : -(P.n1 - 1) * sy;
            const int iw = (i > 0) ? i - 1 : P.n0 - 1, ie = (i < P.n0 - 1) ? i + 1 : 0;
            const bool in_w = i > 0 || px, in_e = i < P.n0 - 1 || px, in_s = j > 0 || py, in_n = j < P.n1 - 1 || py;
            // phi + cor of this column at plane kk (wrapped); only called in range
            auto colv = [&](int kk) {
                const int c = (int)LIdx(P, i, j, kk);
                double x = phi[c];
                if (cor) x = x + (double)cor[CP ? (int)PIdx(L, i, j, kk) : c];
                return x;
            };
            double vm = 0.0, v0 = colv(k0), vp = 0.0;
            if (k0 > 0 || pz) vm = colv(k0 > 0 ? k0 - 1 : P.n2 - 1);
            for (int k = k0; k < k1; ++k) {
                if (k < P.n2 - 1 || pz) vp = colv(k < P.n2 - 1 ? k + 1 : 0);
                const int c = (int)LIdx(P, i, j, k);
                const ResOffsets<CP> o = res_offsets<CP>(L, c, i, j, k, dw, de, ds, dn, iw, ie);
                auto val = [&](int pc, int qc) {
                    double x = phi[pc];
                    if (cor) x = x + (double)cor[qc];
                    return x;
                };
                const double self = v0;
                auto wall = [&](int bc) { return bc == KSGF_LO_NEUMANN ? self : -self; };
                const double w = in_w ? val(c + dw, o.w) : wall(P.lo[0]);
                const double e = in_e ? val(c + de, o.e) : wall(P.hi[0]);
                const double s = in_s ? val(c + ds, o.s) : wall(P.lo[1]);
                const double n = in_n ? val(c + dn, o.n) : wall(P.hi[1]);
                const double bo = (k > 0 || pz) ? vm : wall(P.lo[2]);
                const double t = (k < P.n2 - 1 || pz) ? vp : wall(P.hi[2]);
                const double st = P.cx * (w + e) + P.cy * (s + n) + P.cz * (bo + t);
                const double v = rhs[c] - (st - P.dg * self);
                if (cor) phi_out[c] = self;
                r[o.q] = (R)v;
                m = fmax(m, abs_or_inf(v));
                vm = v0;
                v0 = vp;
            }
        }
    }
    return m;
}

// The same residual four cells per thread (x extent a multiple of 4, no
// explicit ghosts): a thread marches the column of cells 4 iq .. 4 iq + 3 of
// row j.  phi, rhs and phi_out move as quads; a packed correction and
// residual as two pairs per quad (cells 0 and 2 have the row's colour, 1 and
// 3 the other, at the same two packed positions); the x neighbours beyond
// the quad are single loads.  Every cell's value, operand order and the
// block maxima are those of update_residual_tile.
template <typename C, typename R, bool CP>
__device__ __forceinline__ double update_residual_quads(const double* __restrict__ phi, const C* __restrict__ cor,
                                                        double* __restrict__ phi_out, const double* __restrict__ rhs,
                                                        R* __restrict__ r, const LevelGeom& P, const LevelGeom& L,
                                                        const RowTiles& tl, int lid)
{
    int iq, j, k0, k1;
    row_tile(tl, lid, (int)threadIdx.x, iq, j, k0, k1, P.n2);
    const int i0 = 4 * iq;
    double m = 0.0;
    if (i0 >= P.n0 || j >= P.n1) return m;
    const bool px = P.lo[0] == KSGF_LO_PERIODIC, py = P.lo[1] == KSGF_LO_PERIODIC, pz = P.lo[2] == KSGF_LO_PERIODIC;
    const bool in_w = i0 > 0 || px, in_e = i0 + 4 < P.n0 || px, in_s = j > 0 || py, in_n = j < P.n1 - 1 || py;
    const int js = j > 0 ? j - 1 : P.n1 - 1, jn = j < P.n1 - 1 ? j + 1 : 0;
    const int iw = i0 > 0 ? i0 - 1 : P.n0 - 1, ie = i0 + 4 < P.n0 ? i0 + 4 : 0;
    // phi + cor of cells i0 .. i0 + 3 of row (jj, kk)
    auto quad_at = [&](int jj, int kk) {
        const int c = (int)LIdx(P, i0, jj, kk);
        PkQuad<double> x = *(const PkQuad<double>*)(phi + c);
        if (cor) {
            if (CP) {
                const int rp = (i0 + jj + kk) & 1, row = kk * L.pplane + jj * L.prow + (i0 >> 1);
                const PkPair<C> a = *(const PkPair<C>*)(cor + rp * L.pcol + row);
                const PkPair<C> b = *(const PkPair<C>*)(cor + (1 - rp) * L.pcol + row);
                x.v[0] = x.v[0] + (double)a.v[0];
                x.v[1] = x.v[1] + (double)b.v[0];
                x.v[2] = x.v[2] + (double)a.v[1];
                x.v[3] = x.v[3] + (double)b.v[1];
            } else {
                const PkQuad<C> q = *(const PkQuad<C>*)(cor + c);
#pragma unroll
                for (int e = 0; e < 4; ++e) x.v[e] = x.v[e] + (double)q.v[e];
            }
        }
        return x;
    };
    // phi + cor of the single cell (ii, jj, kk)
    auto one_at = [&](int ii, int jj, int kk) {
        const int c = (int)LIdx(P, ii, jj, kk);
        double x = phi[c];
        if (cor) x = x + (double)cor[CP ? (int)PIdx(L, ii, jj, kk) : c];
        return x;
    };
    PkQuad<double> vm{}, v0 = quad_at(j, k0), vp{};
    if (k0 > 0 || pz) vm = quad_at(j, k0 > 0 ? k0 - 1 : P.n2 - 1);
    for (int k = k0; k < k1; ++k) {
        if (k < P.n2 - 1 || pz) vp = quad_at(j, k < P.n2 - 1 ? k + 1 : 0);
        const int c = (int)LIdx(P, i0, j, k);
        const PkQuad<double> self = v0;
        auto wall = [&](int bc, int e) { return bc == KSGF_LO_NEUMANN ? self.v[e] : -self.v[e]; };
        const double wb = in_w ? one_at(iw, j, k) : wall(P.lo[0], 0);
        const double eb = in_e ? one_at(ie, j, k) : wall(P.hi[0], 3);
        PkQuad<double> sq, nq;
        if (in_s) sq = quad_at(js, k);
        if (in_n) nq = quad_at(jn, k);
        const PkQuad<double> rq = *(const PkQuad<double>*)(rhs + c);
        PkQuad<double> v;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double w = e == 0 ? wb : self.v[e - 1];
            const double ea = e == 3 ? eb : self.v[e + 1];
            const double sv = in_s ? sq.v[e] : wall(P.lo[1], e);
            const double nv = in_n ? nq.v[e] : wall(P.hi[1], e);
            const double bo = (k > 0 || pz) ? vm.v[e] : wall(P.lo[2], e);
            const double t = (k < P.n2 - 1 || pz) ? vp.v[e] : wall(P.hi[2], e);
            const double st = P.cx * (w + ea) + P.cy * (sv + nv) + P.cz * (bo + t);
            v.v[e] = rq.v[e] - (st - P.dg * self.v[e]);
            m = fmax(m, abs_or_inf(v.v[e]));
        }
        if (cor) *(PkQuad<double>*)(phi_out + c) = self;
        if (CP) {
            const int rp = (i0 + j + k) & 1, row = k * L.pplane + j * L.prow + (i0 >> 1);
            *(PkPair<R>*)(r + rp * L.pcol + row) = PkPair<R>{{(R)v.v[0], (R)v.v[2]}};
            *(PkPair<R>*)(r + (1 - rp) * L.pcol + row) = PkPair<R>{{(R)v.v[1], (R)v.v[3]}};
        } else {
            *(PkQuad<R>*)(r + c) = PkQuad<R>{{(R)v.v[0], (R)v.v[1], (R)v.v[2], (R)v.v[3]}};
        }
        vm = v0;
        v0 = vp;
    }
    return m;
}

template <typename C, typename R, bool CP>
__global__ void __launch_bounds__(256) k_update_residual4(const double* __restrict__ phi, const C* __restrict__ cor,
                                                          double* __restrict__ phi_out, const double* __restrict__ rhs,
                                                          R* __restrict__ r, LevelGeom P, LevelGeom L, RowTiles tl,
                                                          double* __restrict__ part)
{
    if (P.act && *P.act == 0) return;
    const double m = update_residual_quads<C, R, CP>(phi, cor, phi_out, rhs, r, P, L, tl,
                                                     xcd_remap((int)blockIdx.x, (int)gridDim.x));
    block_max_part(m, part);
}

template <typename C, typename R, bool CP>
__global__ void __launch_bounds__(256) k_update_residual(const double* __restrict__ phi, const C* __restrict__ cor,
                                                         double* __restrict__ phi_out, const double* __restrict__ rhs,
                                                         R* __restrict__ r, LevelGeom P, LevelGeom L, RowTiles tl,
                                                         double* __restrict__ part)
{
    if (P.act && *P.act == 0) return;
    const double m = update_residual_tile<C, R, CP>(phi, cor, phi_out, rhs, r, P, L, tl,
                                                    xcd_remap((int)blockIdx.x, (int)gridDim.x));
    block_max_part(m, part);
}

// maximum of the n partials into *norm (as the bits of a non-negative double);
// kMaxPartsThreads threads with four independent loads in flight each (the
// partials are finite or +inf, so the maximum does not depend on the order)
constexpr int kMaxPartsThreads = 1024;
__global__ void __launch_bounds__(kMaxPartsThreads) k_max_parts(const double* __restrict__ part, int n,
                                                                unsigned long long* __restrict__ norm,
                                                                const int* __restrict__ act)
{
    if (act && *act == 0) return;
    constexpr int B = kMaxPartsThreads;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    int t = (int)threadIdx.x;
    for (; t + 3 * B < n; t += 4 * B) {
        m0 = fmax(m0, part[t]);
        m1 = fmax(m1, part[t + B]);
        m2 = fmax(m2, part[t + 2 * B]);
        m3 = fmax(m3, part[t + 3 * B]);
    }
    for (; t < n; t += B) m0 = fmax(m0, part[t]);
    double m = fmax(fmax(m0, m1), fmax(m2, m3));
    __shared__ double sh[B / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int w = 1; w < B / 64; ++w) r = fmax(r, sh[w]);
        *norm = (unsigned long long)__double_as_longlong(r);
    }
}


template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ a, long long n, T v)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

template <typename A, typename B>
__global__ void __launch_bounds__(256) k_convert(A* __restrict__ a, const B* __restrict__ b, long long n)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = (A)b[t];
}


// deterministic sum of interior cells: per-block partial sums
constexpr int kSumBlocks = 512;
__global__ void __launch_bounds__(256) k_partial_sum(const double* __restrict__ a, LevelGeom L, double* __restrict__ part)
{
    __shared__ double sh[256];
    const long long tot = (long long)L.n0 * L.n1 * L.n2;
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
        const unsigned tt = (unsigned)t, q = tt / (unsigned)L.n0;
        int i = (int)(tt - q * (unsigned)L.n0);
        int j = (int)(q % (unsigned)L.n1), k = (int)(q / (unsigned)L.n1);
        s += a[LIdx(L, i, j, k)];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_sub_mean(double* __restrict__ a, LevelGeom L, const double* __restrict__ part,
                                                  int npart)
{
    // every block sums the partials in the same fixed order (256 threads)
    __shared__ double sh[4];
    __shared__ double mean;
    double s = 0.0;
    for (int p = (int)threadIdx.x; p < npart; p += (int)blockDim.x) s += part[p];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) mean = ((sh[0] + sh[1]) + (sh[2] + sh[3])) / ((double)L.n0 * L.n1 * L.n2);
    __syncthreads();
    const long long tot = (long long)L.n0 * L.n1 * L.n2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
        const unsigned tt = (unsigned)t, q = tt / (unsigned)L.n0;
        int i = (int)(tt - q * (unsigned)L.n0);
        int j = (int)(q % (unsigned)L.n1), k = (int)(q / (unsigned)L.n1);
        a[LIdx(L, i, j, k)] -= mean;
    }
}

// ---------------------------------------------------------------------------
// single-workgroup part: the coarse end of the V-cycle and the bottom solve
// ---------------------------------------------------------------------------
struct BottomWork {
    double *r, *rh, *p, *v, *s, *t;
    double *x, *b;  // FP64 bottom unknown / rhs (FP32 V-cycles convert in and out)
    int* stats;     // [0] BiCGStab iterations (accumulated), [1] failures
    int lds;        // byte offset of the eight vectors in the LDS pool, or -1
};

template <typename T>
struct LevelPtrs {
    LevelGeom G;
    T* cor;
    T* res;
    int r0, r1, r2;  // ratio to the next coarser level
    int lds;         // byte offset of cor (then res) in the LDS pool, or -1
};

// LDS pool of the single-workgroup coarse cycle.  The levels below its
// first level live only inside the kernel, so from the bottom up as many of
// them as fit (and the bottom solver's vectors) are kept in LDS instead of
// HBM: every pass of the coarse cycle ends in a workgroup barrier, and LDS
// latency instead of L2 latency is what makes those passes cheap.  The
// first level joins them when it fits too: its residual (written by the
// grid-level restriction) is copied in at the start and its correction
// (read by the grid-level prolongation) copied out at the end.
constexpr int kLdsPoolBytes = 152 * 1024;

__host__ __device__ inline int lds_align(long long bytes)
{
    return (int)((bytes + 15) / 16 * 16);
}

// Geometry of a level held in the pool: LIdx's index function with rows of
// n0 + 2 elements instead of the HBM arrays' 64-byte aligned pitch (LIdx
// puts cell i = -1 at element kXOff - 1 of a row, so the slice starts
// kXOff - 1 elements early and holds (n0+2)(n1+2)(n2+2) cells after that)
__host__ __device__ inline LevelGeom lds_geom(const LevelGeom& G)
{
    LevelGeom g = G;
    g.sy = G.n0 + 2;
    g.sz = g.sy * (G.n1 + 2);
    g.tot = (kXOff - 1) + g.sz * (G.n2 + 2);
    return g;
}

// block-wide sum / maximum: a wavefront butterfly, then the per-wave
// results in fixed order (deterministic, two barriers)
__device__ double blk_sum(double v, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r += sh[q];
    __syncthreads();
    return r;
}
__device__ double blk_max(double v, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = fmax(r, sh[q]);
    __syncthreads();
    return r;
}

template <typename T>
__device__ void blk_ghosts(T* a, const LevelGeom& L)
{
    const int n = (int)ghost_count(L);
    for (int t = (int)threadIdx.x; t < n; t += (int)blockDim.x) ghost_point<T>(a, L, t);
    __syncthreads();
}

#define BLK_FOR(L, t_, c)                                                                              \
    for (int t_ = (int)threadIdx.x, ncell_ = (L).n0 * (L).n1 * (L).n2; t_ < ncell_; t_ += (int)blockDim.x)  \
        for (long long c = LIdx((L), t_ % (L).n0, (t_ / (L).n0) % (L).n1, (t_ / (L).n0) / (L).n1), once_ = 0; \
             once_ < 1; ++once_)

template <typename T>
__device__ void blk_smooth(T* x, const T* b, const LevelGeom& L, int nu)
{
    const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz, idg = (T)L.idg;
    for (int sweep = 0; sweep < nu; ++sweep)
        for (int color = 0; color < 2; ++color) {
            blk_ghosts<T>(x, L);
            BLK_FOR(L, t_, c)
            {
                const int q = t_ / L.n0;
                const int i = t_ % L.n0, j = q % L.n1, k = q / L.n1;
                if (((i + j + k + color) & 1) == 0)
                    x[c] = (cx * (x[c - 1] + x[c + 1]) + cy * (x[c - L.sy] + x[c + L.sy]) +
                            cz * (x[c - L.sz] + x[c + L.sz]) - b[c]) *
                           idg;
            }
            __syncthreads();
        }
}

__device__ void blk_apply(double* x, double* y, const LevelGeom& L)
{
    blk_ghosts<double>(x, L);
    BLK_FOR(L, t_, c)
    {
        y[c] = (L.cx * (x[c - 1] + x[c + 1]) + L.cy * (x[c - L.sy] + x[c + L.sy]) + L.cz * (x[c - L.sz] + x[c + L.sz])) -
               L.dg * x[c];
    }
    __syncthreads();
}

// unpreconditioned BiCGStab (van der Vorst 1992) on the bottom level, then
// nub (converged) or nuf (not converged) smoothing sweeps
__device__ void bottom_solve(double* x, double* b, const LevelGeom& L, const BottomWork& w, int singular, double rtol,
                             double atol, int maxiter, int nub, int nuf, double* sh)
{
    const long long ncell = (long long)L.n0 * L.n1 * L.n2;
    if (singular) {
        double s = 0.0;
        BLK_FOR(L, t_, c) { s += b[c]; }
        double mean = blk_sum(s, sh) / (double)ncell;
        BLK_FOR(L, t_, c) { b[c] -= mean; }
        __syncthreads();
    }
    for (long long c = threadIdx.x; c < L.tot; c += blockDim.x) {
        x[c] = 0.0;
        w.r[c] = b[c];
        w.rh[c] = b[c];
        w.p[c] = 0.0;
        w.v[c] = 0.0;
        w.s[c] = 0.0;
        w.t[c] = 0.0;
    }
    __syncthreads();
    double m = 0.0;
    BLK_FOR(L, t_, c) { m = fmax(m, fabs(w.r[c])); }
    const double rnorm0 = blk_max(m, sh);
    const double tol = fmax(rtol * rnorm0, atol);
    int ret = 1, it = 0;
    if (rnorm0 == 0.0 || rnorm0 <= tol) {
        ret = 0;
    } else {
        double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
        for (it = 1; it <= maxiter; ++it) {
            double acc = 0.0;
            BLK_FOR(L, t_, c) { acc += w.rh[c] * w.r[c]; }
            double rho = blk_sum(acc, sh);
            if (rho == 0.0) { ret = 2; break; }
            if (it == 1) {
                BLK_FOR(L, t_, c) { w.p[c] = w.r[c]; }
            } else {
                double beta = (rho / rho_prev) * (alpha / omega);
                BLK_FOR(L, t_, c) { w.p[c] = w.r[c] + beta * (w.p[c] - omega * w.v[c]); }
            }
            __syncthreads();
            blk_apply(w.p, w.v, L);
            acc = 0.0;
            BLK_FOR(L, t_, c) { acc += w.rh[c] * w.v[c]; }
            double rv = blk_sum(acc, sh);
            if (rv == 0.0) { ret = 2; break; }
            alpha = rho / rv;
            m = 0.0;
            BLK_FOR(L, t_, c)
            {
                w.s[c] = w.r[c] - alpha * w.v[c];
                m = fmax(m, fabs(w.s[c]));
            }
            double sn = blk_max(m, sh);
            if (sn <= tol) {
                BLK_FOR(L, t_, c) { x[c] += alpha * w.p[c]; }
                __syncthreads();
                ret = 0;
                break;
            }
            blk_apply(w.s, w.t, L);
            acc = 0.0;
            BLK_FOR(L, t_, c) { acc += w.t[c] * w.t[c]; }
            double tt = blk_sum(acc, sh);
            if (tt == 0.0) { ret = 2; break; }
            acc = 0.0;
            BLK_FOR(L, t_, c) { acc += w.t[c] * w.s[c]; }
            omega = blk_sum(acc, sh) / tt;
            m = 0.0;
            BLK_FOR(L, t_, c)
            {
                x[c] += alpha * w.p[c] + omega * w.s[c];
                w.r[c] = w.s[c] - omega * w.t[c];
                m = fmax(m, fabs(w.r[c]));
            }
            double rn = blk_max(m, sh);
            if (rn <= tol) { ret = 0; break; }
            if (omega == 0.0) { ret = 2; break; }
            rho_prev = rho;
        }
        if (it > maxiter) it = maxiter;
    }
    __syncthreads();
    blk_smooth<double>(x, b, L, (ret == 0) ? nub : nuf);
    if (threadIdx.x == 0) {
        w.stats[0] += it;
        if (ret != 0) w.stats[1] += 1;
    }
    __syncthreads();
}

// The same BiCGStab on a bottom level of at most 64 K cells (K = 1, 2) run by one
// wavefront with no workgroup barrier: lane l holds cells l, l + 64, ...
// (slot s = cell / 64) of every vector in registers, neighbours come by lane
// shuffles (the ghost rules of bc_ghost applied to the lane's own value), and
// the sums and maxima are wave butterflies.  With K = 1 (a bottom of at most
// 64 cells) the sums take blk_sum's order, so the iterates are those of
// bottom_solve bit for bit.  Returns (through the pointers) bottom_solve's
// status and iteration count; x is written back to memory.
template <int K>
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

template <int K>
__device__ __forceinline__ double wave_get(const double (&v)[K], int cell)
{
    // every lane takes part in every shuffle (uniform slot loop)
    double r = 0.0;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const double t = __shfl(v[s], cell & 63, 64);
        if ((cell >> 6) == s) r = t;
    }
    return r;
}

template <int K>
__device__ __forceinline__ void wave_apply(const double (&x)[K], double (&y)[K], const LevelGeom& L, int lane)
{
    const int n0 = L.n0, n1 = L.n1, n2 = L.n2, ncell = n0 * n1 * n2;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int t = s * 64 + lane;
        const int tt = (t < ncell) ? t : 0;
        const int i = tt % n0, j = (tt / n0) % n1, k = tt / (n0 * n1);
        const double self = x[s];
        auto nb = [&](int ii, int jj, int kk, int d, int side) {
            const int pos = (d == 0) ? ii : (d == 1) ? jj : kk;
            const int n = (d == 0) ? n0 : (d == 1) ? n1 : n2;
            int c[3] = {ii, jj, kk};
            const bool out = pos < 0 || pos >= n;
            if (out) c[d] = (pos < 0) ? n - 1 : 0;
            const double v = wave_get<K>(x, c[0] + n0 * (c[1] + n1 * c[2]));
            if (!out) return v;
            return bc_ghost<double>(side ? L.hi[d] : L.lo[d], self, v);
        };
        const double w = nb(i - 1, j, k, 0, 0), e = nb(i + 1, j, k, 0, 1);
        const double so = nb(i, j - 1, k, 1, 0), no = nb(i, j + 1, k, 1, 1);
        const double bo = nb(i, j, k - 1, 2, 0), to = nb(i, j, k + 1, 2, 1);
        y[s] = (L.cx * (w + e) + L.cy * (so + no) + L.cz * (bo + to)) - L.dg * self;
    }
}

template <int K>
__device__ void wave_bottom_solve(double* xm, double* bm, const LevelGeom& L, int singular, double rtol, double atol,
                                  int maxiter, int* ret_out, int* it_out)
{
    const int lane = (int)threadIdx.x;
    const int n0 = L.n0, n1 = L.n1, ncell = L.n0 * L.n1 * L.n2;
    bool valid[K];
    double b[K], x[K], r[K], rh[K], p[K], v[K], sv[K], t[K];
    auto idx = [&](int s) {
        const int cc = s * 64 + lane;
        return LIdx(L, cc % n0, (cc / n0) % n1, cc / (n0 * n1));
    };
#pragma unroll
    for (int s = 0; s < K; ++s) {
        valid[s] = s * 64 + lane < ncell;
        b[s] = valid[s] ? bm[idx(s)] : 0.0;
    }
    if (singular) {
        double acc = 0.0;
#pragma unroll
        for (int s = 0; s < K; ++s) acc += b[s];
        const double mean = wave_sum<K>(acc) / (double)ncell;
#pragma unroll
        for (int s = 0; s < K; ++s)
            if (valid[s]) {
                b[s] -= mean;
                bm[idx(s)] = b[s];
            }
    }
#pragma unroll
    for (int s = 0; s < K; ++s) {
        x[s] = 0.0;
        r[s] = b[s];
        rh[s] = b[s];
        p[s] = v[s] = sv[s] = t[s] = 0.0;
    }
    auto vmax = [&](const double (&a)[K]) {
        double m = 0.0;
#pragma unroll
        for (int s = 0; s < K; ++s) m = fmax(m, fabs(a[s]));
        return wave_max(m);
    };
    auto vdot = [&](const double (&a)[K], const double (&c)[K]) {
        double acc = 0.0;
#pragma unroll
        for (int s = 0; s < K; ++s) acc += a[s] * c[s];
        return wave_sum<K>(acc);
    };
    const double rnorm0 = vmax(r);
    const double tol = fmax(rtol * rnorm0, atol);
    int ret = 1, it = 0;
    if (rnorm0 == 0.0 || rnorm0 <= tol) {
        ret = 0;
    } else {
        double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
        for (it = 1; it <= maxiter; ++it) {
            const double rho = vdot(rh, r);
            if (rho == 0.0) { ret = 2; break; }
            if (it == 1) {
#pragma unroll
                for (int s = 0; s < K; ++s) p[s] = r[s];
            } else {
                const double beta = (rho / rho_prev) * (alpha / omega);
#pragma unroll
                for (int s = 0; s < K; ++s) p[s] = r[s] + beta * (p[s] - omega * v[s]);
            }
            wave_apply<K>(p, v, L, lane);
#pragma unroll
            for (int s = 0; s < K; ++s)
                if (!valid[s]) v[s] = 0.0;
            const double rv = vdot(rh, v);
            if (rv == 0.0) { ret = 2; break; }
            alpha = rho / rv;
#pragma unroll
            for (int s = 0; s < K; ++s) sv[s] = r[s] - alpha * v[s];
            if (vmax(sv) <= tol) {
#pragma unroll
                for (int s = 0; s < K; ++s) x[s] += alpha * p[s];
                ret = 0;
                break;
            }
            wave_apply<K>(sv, t, L, lane);
#pragma unroll
            for (int s = 0; s < K; ++s)
                if (!valid[s]) t[s] = 0.0;
            const double tt = vdot(t, t);
            if (tt == 0.0) { ret = 2; break; }
            omega = vdot(t, sv) / tt;
#pragma unroll
            for (int s = 0; s < K; ++s) {
                x[s] += alpha * p[s] + omega * sv[s];
                r[s] = sv[s] - omega * t[s];
            }
            if (vmax(r) <= tol) { ret = 0; break; }
            if (omega == 0.0) { ret = 2; break; }
            rho_prev = rho;
        }
        if (it > maxiter) it = maxiter;
    }
#pragma unroll
    for (int s = 0; s < K; ++s)
        if (valid[s]) xm[idx(s)] = x[s];
    *ret_out = ret;
    *it_out = it;
}

// bottom_solve with the one-wavefront BiCGStab where the bottom level has at
// most 128 cells: wave 0 solves while the other waves wait at the barrier
__device__ void bottom_solve_any(double* x, double* b, const LevelGeom& L, const BottomWork& w, int singular,
                                 double rtol, double atol, int maxiter, int nub, int nuf, double* sh)
{
    const int ncell = L.n0 * L.n1 * L.n2;
    if (ncell > 128) {
        bottom_solve(x, b, L, w, singular, rtol, atol, maxiter, nub, nuf, sh);
        return;
    }
    // the ghost layer keeps its old values; the final smoothing refills it
    for (long long c = threadIdx.x; c < L.tot; c += blockDim.x) x[c] = 0.0;
    __syncthreads();
    if (threadIdx.x < 64) {
        int ret = 0, it = 0;
        if (ncell <= 64)
            wave_bottom_solve<1>(x, b, L, singular, rtol, atol, maxiter, &ret, &it);
        else
            wave_bottom_solve<2>(x, b, L, singular, rtol, atol, maxiter, &ret, &it);
        if (threadIdx.x == 0) {
            sh[0] = (double)ret;
            sh[1] = (double)it;
        }
    }
    __syncthreads();
    const int ret = (int)sh[0], it = (int)sh[1];
    __syncthreads();
    blk_smooth<double>(x, b, L, (ret == 0) ? nub : nuf);
    if (threadIdx.x == 0) {
        w.stats[0] += it;
        if (ret != 0) w.stats[1] += 1;
    }
    __syncthreads();
}

// levels lc .. nlev-1 of a V-cycle: zero, pre-smooth, residual, restrict
// down to the bottom, bottom solve (FP64), prolong, post-smooth back up
template <typename T>
__device__ __forceinline__ T* lv_cor(const LevelPtrs<T>* lp, int l, char* pool)
{
    return lp[l].lds >= 0 ? (T*)(pool + lp[l].lds) : lp[l].cor;
}
template <typename T>
__device__ __forceinline__ T* lv_res(const LevelPtrs<T>* lp, int l, char* pool)
{
    return lp[l].lds >= 0 ? (T*)(pool + lp[l].lds + lds_align((long long)lds_geom(lp[l].G).tot * sizeof(T)))
                          : lp[l].res;
}
template <typename T>
__device__ __forceinline__ LevelGeom lv_geom(const LevelPtrs<T>* lp, int l)
{
    return lp[l].lds >= 0 ? lds_geom(lp[l].G) : lp[l].G;
}
// copy the valid cells of level l between its HBM array (geometry Gs or Gd)
// and its pool slice (the other geometry), one (i, j) column per thread and
// all of a column's loads issued before its stores; ends in a workgroup
// barrier
constexpr int kStageBatch = 8;
template <typename T>
__device__ void lv_stage(T* dst, const LevelGeom& Gd, const T* src, const LevelGeom& Gs)
{
    const int ncol = Gd.n0 * Gd.n1, szd = (int)Gd.sz, szs = (int)Gs.sz;
    for (int t = (int)threadIdx.x; t < ncol; t += (int)blockDim.x) {
        const int j = t / Gd.n0, i = t - j * Gd.n0;
        const int cd = (int)LIdx(Gd, i, j, 0), cs = (int)LIdx(Gs, i, j, 0);
        for (int k0 = 0; k0 < Gd.n2; k0 += kStageBatch) {
            T v[kStageBatch];
#pragma unroll
            for (int m = 0; m < kStageBatch; ++m) v[m] = src[cs + min(k0 + m, Gd.n2 - 1) * szs];
#pragma unroll
            for (int m = 0; m < kStageBatch; ++m)
                if (k0 + m < Gd.n2) dst[cd + (k0 + m) * szd] = v[m];
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Passes of a level held in the pool (lds_geom layout) with the domain sides
// inside the stencil instead of a ghost pass before every half: a periodic
// side wraps, a Neumann side reads the cell itself, a Dirichlet side the
// cell itself negated -- the values bc_ghost puts in the ghost layer, and
// exactly those: wherever a half sweep reads a wrapped neighbour, that
// neighbour has the other colour (no periodic direction of odd extent), so
// it is not written during the half.  Each thread takes (i, j) columns and
// marches along k, so the x and y neighbour offsets are formed once per
// column; the loads of kPoolBatch cells of a column (half as many in FP64)
// are issued before any of their stores (a half sweep reads the other colour, its own cells and b;
// the restriction and prolongation write no array they read).  The pointers
// are formed from the pool itself, so the accesses are LDS instructions.
// ---------------------------------------------------------------------------
constexpr int kPoolBatch = 4;

__device__ __forceinline__ bool pool_sweepable(const LevelGeom& G)
{
    const int n[3] = {G.n0, G.n1, G.n2};
    for (int d = 0; d < 3; ++d)
        if ((G.lo[d] == KSGF_LO_PERIODIC || G.hi[d] == KSGF_LO_PERIODIC) && (n[d] & 1)) return false;
    return true;
}

// the domain sides of a pool level: the offset a boundary cell's low / high
// neighbour is at (the periodic image, or the cell itself), and whether that
// value is negated (Dirichlet)
struct PoolSides {
    int n[3], s[3], wlo[3], whi[3];
    unsigned nlo, nhi;  // bit d: negate at the low / high side of direction d
};
__device__ __forceinline__ PoolSides pool_sides(const LevelGeom& F)
{
    PoolSides p;
    p.n[0] = F.n0;
    p.n[1] = F.n1;
    p.n[2] = F.n2;
    p.s[0] = 1;
    p.s[1] = (int)F.sy;
    p.s[2] = (int)F.sz;
    p.nlo = p.nhi = 0u;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        p.wlo[d] = F.lo[d] == KSGF_LO_PERIODIC ? (p.n[d] - 1) * p.s[d] : 0;
        p.whi[d] = F.hi[d] == KSGF_LO_PERIODIC ? -(p.n[d] - 1) * p.s[d] : 0;
        if (F.lo[d] != KSGF_LO_PERIODIC && F.lo[d] != KSGF_LO_NEUMANN) p.nlo |= 1u << d;
        if (F.hi[d] != KSGF_LO_PERIODIC && F.hi[d] != KSGF_LO_NEUMANN) p.nhi |= 1u << d;
    }
    return p;
}

// one (i, j) column: offset of cell k = 0, the relative offsets of the W, E,
// S, N neighbours, and which of those are negated (bits 0..3)
struct PoolCol {
    int c0, o[4];
    unsigned neg;
};
__device__ __forceinline__ PoolCol pool_col(const LevelGeom& F, const PoolSides& p, int i, int j)
{
    PoolCol q;
    q.c0 = (int)LIdx(F, i, j, 0);
    const bool wl = i == 0, wh = i == p.n[0] - 1, sl = j == 0, sh = j == p.n[1] - 1;
    q.o[0] = wl ? p.wlo[0] : -1;
    q.o[1] = wh ? p.whi[0] : 1;
    q.o[2] = sl ? p.wlo[1] : -p.s[1];
    q.o[3] = sh ? p.whi[1] : p.s[1];
    q.neg = ((wl && (p.nlo & 1u)) ? 1u : 0u) | ((wh && (p.nhi & 1u)) ? 2u : 0u) |
            ((sl && (p.nlo & 2u)) ? 4u : 0u) | ((sh && (p.nhi & 2u)) ? 8u : 0u);
    return q;
}

// the six neighbours of cell k of a column, sides applied, summed as
// cx*(W+E) + cy*(S+N) + cz*(B+T) (the oracle's operand order)
template <typename T>
struct PoolNb6 {
    T v[6];
    unsigned neg;
};
template <typename T>
__device__ __forceinline__ PoolNb6<T> pool_load_nb(const T* x, const PoolCol& q, const PoolSides& p, int c, int k)
{
    PoolNb6<T> r;
    const bool bl = k == 0, bh = k == p.n[2] - 1;
    r.v[0] = x[c + q.o[0]];
    r.v[1] = x[c + q.o[1]];
    r.v[2] = x[c + q.o[2]];
    r.v[3] = x[c + q.o[3]];
    r.v[4] = x[c + (bl ? p.wlo[2] : -p.s[2])];
    r.v[5] = x[c + (bh ? p.whi[2] : p.s[2])];
    r.neg = q.neg | ((bl && (p.nlo & 4u)) ? 16u : 0u) | ((bh && (p.nhi & 4u)) ? 32u : 0u);
    return r;
}
template <typename T>
__device__ __forceinline__ T pool_sum6(const PoolNb6<T>& r, T cx, T cy, T cz)
{
    T u[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) u[s] = ((r.neg >> s) & 1u) ? -r.v[s] : r.v[s];
    return cx * (u[0] + u[1]) + cy * (u[2] + u[3]) + cz * (u[4] + u[5]);
}

// one half sweep of colour col
template <typename T>
__device__ __forceinline__ void pool_half(T* x, const T* b, const LevelGeom& F, int col)
{
    constexpr int NB = sizeof(T) == sizeof(float) ? kPoolBatch : kPoolBatch / 2;
    const T cx = (T)F.cx, cy = (T)F.cy, cz = (T)F.cz, idg = (T)F.idg;
    const PoolSides ps = pool_sides(F);
    const int ncol = F.n0 * F.n1, sz = ps.s[2];
    for (int t = (int)threadIdx.x; t < ncol; t += (int)blockDim.x) {
        const int j = t / F.n0, i = t - j * F.n0;
        const PoolCol pc = pool_col(F, ps, i, j);
        for (int kb = (i + j + col) & 1; kb < F.n2; kb += 2 * NB) {
            PoolNb6<T> nb[NB];
            T bv[NB];
#pragma unroll
            for (int m = 0; m < NB; ++m) {
                // past the column's end: load the batch's first cell again
                const int k = (kb + 2 * m < F.n2) ? kb + 2 * m : kb, c = pc.c0 + k * sz;
                nb[m] = pool_load_nb<T>(x, pc, ps, c, k);
                bv[m] = b[c];
            }
#pragma unroll
            for (int m = 0; m < NB; ++m) {
                const int k = kb + 2 * m;
                if (k < F.n2) x[pc.c0 + k * sz] = (pool_sum6<T>(nb[m], cx, cy, cz) - bv[m]) * idg;
            }
        }
    }
}

// the first half of a sweep from a zero initial guess: colour-0 cells
// (0 - b)/diag (their neighbours are all zero), colour-1 cells 0
template <typename T>
__device__ __forceinline__ void pool_half_zero(T* x, const T* b, const LevelGeom& F)
{
    const T idg = (T)F.idg;
    const int ncol = F.n0 * F.n1, sz = (int)F.sz;
    for (int t = (int)threadIdx.x; t < ncol; t += (int)blockDim.x) {
        const int j = t / F.n0, i = t - j * F.n0;
        const int c0 = (int)LIdx(F, i, j, 0);
        for (int k = 0; k < F.n2; ++k) {
            const int c = c0 + k * sz;
            x[c] = ((i + j + k) & 1) == 0 ? ((T)0 - b[c]) * idg : (T)0;
        }
    }
}

// nu red-black sweeps; zero_first: from a zero initial guess
template <typename T>
__device__ __forceinline__ void pool_smooth(T* x, const T* b, const LevelGeom& F, int nu, bool zero_first)
{
    for (int s = 0; s < nu; ++s) {
        if (zero_first && s == 0)
            pool_half_zero<T>(x, b, F);
        else
            pool_half<T>(x, b, F, 0);
        __syncthreads();
        pool_half<T>(x, b, F, 1);
        __syncthreads();
    }
}

// x = 0 on the valid cells
template <typename T>
__device__ __forceinline__ void pool_zero(T* x, const LevelGeom& F)
{
    const int ncol = F.n0 * F.n1, sz = (int)F.sz;
    for (int t = (int)threadIdx.x; t < ncol; t += (int)blockDim.x) {
        const int j = t / F.n0, i = t - j * F.n0;
        const int c0 = (int)LIdx(F, i, j, 0);
        for (int k = 0; k < F.n2; ++k) x[c0 + k * sz] = (T)0;
    }
    __syncthreads();
}

// cres = the average over each coarse cell's children of b - A x (children
// in (dk, dj, di) order, the oracle's); ratios of 1 or 2 per direction.
// Each thread takes a coarse column and its up to four fine columns.
template <typename T>
__device__ __forceinline__ void pool_restrict(const T* x, const T* b, const LevelGeom& F, T* cres,
                                              const LevelGeom& C, int r0, int r1, int r2)
{
    constexpr int CB = sizeof(T) == sizeof(float) ? 4 : 2;  // children whose loads are in flight together
    const T wgt = (T)(1.0 / (double)(r0 * r1 * r2));
    const T cx = (T)F.cx, cy = (T)F.cy, cz = (T)F.cz, dg = (T)F.dg;
    const PoolSides ps = pool_sides(F);
    const int ncol = C.n0 * C.n1, fsz = ps.s[2], csz = (int)C.sz;
    for (int t = (int)threadIdx.x; t < ncol; t += (int)blockDim.x) {
        const int J = t / C.n0, I = t - J * C.n0;
        PoolCol pc[4];
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) {
            const int di = ch & 1, dj = ch >> 1;
            pc[ch] = pool_col(F, ps, I * r0 + (di < r0 ? di : 0), J * r1 + (dj < r1 ? dj : 0));
        }
        const int cc0 = (int)LIdx(C, I, J, 0);
        for (int K = 0; K < C.n2; ++K) {
            T s = (T)0;
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                if (dk >= r2) break;
                const int k = K * r2 + dk;
#pragma unroll
                for (int g = 0; g < 4; g += CB) {
                    PoolNb6<T> nb[CB];
                    T bv[CB], xv[CB];
#pragma unroll
                    for (int m = 0; m < CB; ++m) {
                        const int c = pc[g + m].c0 + k * fsz;
                        nb[m] = pool_load_nb<T>(x, pc[g + m], ps, c, k);
                        bv[m] = b[c];
                        xv[m] = x[c];
                    }
#pragma unroll
                    for (int m = 0; m < CB; ++m) {
                        const int di = (g + m) & 1, dj = (g + m) >> 1;
                        if (di < r0 && dj < r1) s += bv[m] - (pool_sum6<T>(nb[m], cx, cy, cz) - dg * xv[m]);
                    }
                }
            }
            cres[cc0 + K * csz] = s * wgt;
        }
    }
    __syncthreads();
}

// x += the coarse correction of each cell's parent (ratios of 1 or 2)
template <typename T>
__device__ __forceinline__ void pool_prolong(T* x, const LevelGeom& F, const T* xc, const LevelGeom& C, int r0, int r1,
                                             int r2)
{
    const int ncol = F.n0 * F.n1, sz = (int)F.sz, csz = (int)C.sz, s2 = r2 >> 1;
    for (int t = (int)threadIdx.x; t < ncol; t += (int)blockDim.x) {
        const int j = t / F.n0, i = t - j * F.n0;
        const int c0 = (int)LIdx(F, i, j, 0), p0 = (int)LIdx(C, i >> (r0 >> 1), j >> (r1 >> 1), 0);
        for (int kb = 0; kb < F.n2; kb += kStageBatch) {
            T a[kStageBatch], p[kStageBatch];
#pragma unroll
            for (int m = 0; m < kStageBatch; ++m) {
                const int k = min(kb + m, F.n2 - 1);
                a[m] = x[c0 + k * sz];
                p[m] = xc[p0 + (k >> s2) * csz];
            }
#pragma unroll
            for (int m = 0; m < kStageBatch; ++m)
                if (kb + m < F.n2) x[c0 + (kb + m) * sz] = a[m] + p[m];
        }
    }
    __syncthreads();
}

// the correction / residual slice of a level known to be in the pool
template <typename T>
__device__ __forceinline__ T* pool_cor(const LevelPtrs<T>* lp, int l, char* pool)
{
    return (T*)(pool + lp[l].lds);
}
template <typename T>
__device__ __forceinline__ T* pool_res(const LevelPtrs<T>* lp, int l, char* pool)
{
    return (T*)(pool + lp[l].lds + lds_align((long long)lds_geom(lp[l].G).tot * sizeof(T)));
}

template <typename T>
__device__ void coarse_vcycle_body(const LevelPtrs<T>* __restrict__ lp, int lc, int nlev, const BottomWork& wg,
                                   int singular, double rtol, double atol, int maxiter, int nub, int nuf, int nu1,
                                   int nu2, double* sh, char* pool)
{
    BottomWork w = wg;
    if (w.lds >= 0) {
        const int vb = lds_align((long long)lv_geom(lp, nlev - 1).tot * sizeof(double));
        double** v[8] = {&w.r, &w.rh, &w.p, &w.v, &w.s, &w.t, &w.x, &w.b};
        for (int q = 0; q < 8; ++q) *v[q] = (double*)(pool + w.lds + q * vb);
    }
    if (lp[lc].lds >= 0) lv_stage<T>(pool_res(lp, lc, pool), lds_geom(lp[lc].G), lp[lc].res, lp[lc].G);
    for (int l = lc; l < nlev - 1; ++l) {
        const LevelGeom F = lv_geom(lp, l);
        const LevelGeom C = lv_geom(lp, l + 1);
        const int r0 = lp[l].r0, r1 = lp[l].r1, r2 = lp[l].r2;
        if (lp[l].lds >= 0 && pool_sweepable(F) && r0 <= 2 && r1 <= 2 && r2 <= 2) {
            // pool levels are the coarsest ones, so level l + 1 is in the pool too
            T* x = pool_cor(lp, l, pool);
            const T* b = pool_res(lp, l, pool);
            if (nu1 > 0)
                pool_smooth<T>(x, b, F, nu1, true);
            else
                pool_zero<T>(x, F);
            pool_restrict<T>(x, b, F, pool_res(lp, l + 1, pool), C, r0, r1, r2);
            continue;
        }
        T* x = lv_cor(lp, l, pool);
        const T* b = lv_res(lp, l, pool);
        for (long long c = threadIdx.x; c < F.tot; c += blockDim.x) x[c] = (T)0;
        __syncthreads();
        blk_smooth<T>(x, b, F, nu1);
        blk_ghosts<T>(x, F);
        const T wgt = (T)(1.0 / (double)(r0 * r1 * r2));
        const T cx = (T)F.cx, cy = (T)F.cy, cz = (T)F.cz, dg = (T)F.dg;
        T* cres = lv_res(lp, l + 1, pool);
        BLK_FOR(C, t_, cc)
        {
            const int q = t_ / C.n0;
            const int I = t_ % C.n0, J = q % C.n1, K = q / C.n1;
            T s = (T)0;
            for (int dk = 0; dk < r2; ++dk)
                for (int dj = 0; dj < r1; ++dj)
                    for (int di = 0; di < r0; ++di) {
                        long long c = LIdx(F, I * r0 + di, J * r1 + dj, K * r2 + dk);
                        T v = b[c] - ((cx * (x[c - 1] + x[c + 1]) + cy * (x[c - F.sy] + x[c + F.sy]) +
                                       cz * (x[c - F.sz] + x[c + F.sz])) -
                                      dg * x[c]);
                        s += v;
                    }
            cres[cc] = s * wgt;
        }
        __syncthreads();
    }
    // bottom
    {
        const LevelGeom Lb = lv_geom(lp, nlev - 1);
        T* xb = lv_cor(lp, nlev - 1, pool);
        T* bb = lv_res(lp, nlev - 1, pool);
        double* x64 = (sizeof(T) == sizeof(double)) ? (double*)xb : w.x;
        double* b64 = (sizeof(T) == sizeof(double)) ? (double*)bb : w.b;
        if (sizeof(T) != sizeof(double)) {
            for (long long c = threadIdx.x; c < Lb.tot; c += blockDim.x) b64[c] = (double)bb[c];
            __syncthreads();
        }
        bottom_solve_any(x64, b64, Lb, w, singular, rtol, atol, maxiter, nub, nuf, sh);
        if (sizeof(T) != sizeof(double)) {
            for (long long c = threadIdx.x; c < Lb.tot; c += blockDim.x) xb[c] = (T)x64[c];
            __syncthreads();
        }
    }
    for (int l = nlev - 2; l >= lc; --l) {
        const LevelGeom F = lv_geom(lp, l);
        const LevelGeom C = lv_geom(lp, l + 1);
        const int r0 = lp[l].r0, r1 = lp[l].r1, r2 = lp[l].r2;
        if (lp[l].lds >= 0 && pool_sweepable(F) && r0 <= 2 && r1 <= 2 && r2 <= 2) {
            T* x = pool_cor(lp, l, pool);
            pool_prolong<T>(x, F, pool_cor(lp, l + 1, pool), C, r0, r1, r2);
            pool_smooth<T>(x, pool_res(lp, l, pool), F, nu2, false);
            continue;
        }
        T* x = lv_cor(lp, l, pool);
        const T* xc = lv_cor(lp, l + 1, pool);
        BLK_FOR(F, t_, c)
        {
            const int q = t_ / F.n0;
            const int i = t_ % F.n0, j = q % F.n1, k = q / F.n1;
            x[c] += xc[LIdx(C, i / r0, j / r1, k / r2)];
        }
        __syncthreads();
        blk_smooth<T>(x, lv_res(lp, l, pool), F, nu2);
    }
    if (lp[lc].lds >= 0) lv_stage<T>(lp[lc].cor, lp[lc].G, pool_cor(lp, lc, pool), lds_geom(lp[lc].G));
}

template <typename T>
__global__ void __launch_bounds__(kCoarseThreads) k_coarse_vcycle(const LevelPtrs<T>* __restrict__ lp, int lc, int nlev,
                                                                  BottomWork w, int singular, double rtol, double atol,
                                                                  int maxiter, int nub, int nuf, int nu1, int nu2,
                                                                  const int* __restrict__ act)
{
    if (act && *act == 0) return;
    __shared__ double sh[kCoarseThreads / 64];
    __shared__ __align__(16) char pool[kLdsPoolBytes];
    coarse_vcycle_body<T>(lp, lc, nlev, w, singular, rtol, atol, maxiter, nub, nuf, nu1, nu2, sh, pool);
}

// ---------------------------------------------------------------------------
// MAC-specific gathers / scatters between box FABs and the level-0 array
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mac_rhs(CMFDev u, CMFDev v, CMFDev w, double* __restrict__ rhs, LevelGeom L,
                                                 RegDev r, const int* __restrict__ box, int dlo0, int dlo1, int dlo2,
                                                 double dxi, double dyi, double dzi)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab fu = fab_at(u, box, b), fv = fab_at(v, box, b), fw = fab_at(w, box, b);
    rhs[LIdx(L, i - dlo0, j - dlo1, k - dlo2)] = (fu(i + 1, j, k) - fu(i, j, k)) * dxi +
                                                 (fv(i, j + 1, k) - fv(i, j, k)) * dyi +
                                                 (fw(i, j, k + 1) - fw(i, j, k)) * dzi;
}

template <int D>
__global__ void __launch_bounds__(256) k_mac_correct(MFDev u, const double* __restrict__ phi, LevelGeom L, RegDev r,
                                                     const int* __restrict__ box, int dlo0, int dlo1, int dlo2,
                                                     double bd)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab fu = fab_at(u, box, b);
    long long c = LIdx(L, i - dlo0, j - dlo1, k - dlo2);
    const long long sd = (D == 0) ? 1 : (D == 1 ? L.sy : L.sz);
    double pr = phi[c], pl = phi[c - sd];
    fu(i, j, k) = fu(i, j, k) - bd * (pr - pl);
}

// the three corrections in one pass over the cells grown by one on the high
// sides: position (i, j, k) updates each face array whose face it is (an x
// face for j, k inside the box, ...), so phi is read once for all three;
// the per-face arithmetic is k_mac_correct's
__global__ void __launch_bounds__(256) k_mac_correct3(MFDev u, MFDev v, MFDev w, const double* __restrict__ phi,
                                                      LevelGeom L, RegDev r, const int* __restrict__ box, int dlo0,
                                                      int dlo1, int dlo2, double bdx, double bdy, double bdz)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const int* bx = box + 6 * b;
    const bool ix = i <= bx[3], jy = j <= bx[4], kz = k <= bx[5];
    const long long c = LIdx(L, i - dlo0, j - dlo1, k - dlo2);
    const double pr = phi[c];
    if (jy && kz) {
        Fab fu = fab_at(u, box, b);
        fu(i, j, k) = fu(i, j, k) - bdx * (pr - phi[c - 1]);
    }
    if (ix && kz) {
        Fab fv = fab_at(v, box, b);
        fv(i, j, k) = fv(i, j, k) - bdy * (pr - phi[c - L.sy]);
    }
    if (ix && jy) {
        Fab fw = fab_at(w, box, b);
        fw(i, j, k) = fw(i, j, k) - bdz * (pr - phi[c - L.sz]);
    }
}

// one component of a cell MultiFab into the level-0 array (times scale), and back
__global__ void __launch_bounds__(256) k_mf_gather(CMFDev m, int comp, double scale, double* __restrict__ out,
                                                   LevelGeom L, RegDev r, const int* __restrict__ box, int dlo0,
                                                   int dlo1, int dlo2)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab f = fab_at(m, box, b);
    out[LIdx(L, i - dlo0, j - dlo1, k - dlo2)] = scale * f(i, j, k, comp);
}
__global__ void __launch_bounds__(256) k_mf_scatter(MFDev m, int comp, const double* __restrict__ a, LevelGeom L,
                                                    RegDev r, const int* __restrict__ box, int dlo0, int dlo1, int dlo2)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab f = fab_at(m, box, b);
    f(i, j, k, comp) = a[LIdx(L, i - dlo0, j - dlo1, k - dlo2)];
}
// y = alpha x - div(beta grad x) from a level-0 array whose ghosts hold the BCs
__global__ void __launch_bounds__(256) k_abec_apply(const double* __restrict__ x, MFDev y, int comp, LevelGeom L,
                                                    RegDev r, const int* __restrict__ box, int dlo0, int dlo1, int dlo2)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const long long c = LIdx(L, i - dlo0, j - dlo1, k - dlo2);
    const double off = L.cx * (x[c - 1] + x[c + 1]) + L.cy * (x[c - L.sy] + x[c + L.sy]) +
                       L.cz * (x[c - L.sz] + x[c + L.sz]);
    Fab f = fab_at(y, box, b);
    f(i, j, k, comp) = -(off - L.dg * x[c]);
}

__global__ void __launch_bounds__(256) k_pack(const double* __restrict__ a, double* __restrict__ out, LevelGeom L,
                                              int unpack)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)L.n0 * L.n1 * L.n2;
    if (t >= tot) return;
    const unsigned tt = (unsigned)t, q = tt / (unsigned)L.n0;
    int i = (int)(tt - q * (unsigned)L.n0);
    int j = (int)(q % (unsigned)L.n1), k = (int)(q / (unsigned)L.n1);
    long long c = LIdx(L, i, j, k);
    if (unpack)
        out[c] = a[t];
    else
        out[t] = a[c];
}

// state of a device-controlled solve: the convergence test of every
// iteration runs on the device, and the kernels of later iterations return
// at once (LevelGeom::act points at `active`)
struct MgCtl {
    int active;    // 1 while iterating
    int iters;     // iterations done
    int diverged;  // the residual exceeded kDivergeFactor |rhs| (or is not finite)
    int pad;
    double rhsnorm, target, r;
};

// the first residual's norm (|rhs|_inf) sets the target max(atol, rtol |rhs|)
__global__ void k_ctl_init(MgCtl* __restrict__ c, const unsigned long long* __restrict__ norm, double rtol,
                           double atol, int maxiter)
{
    if (threadIdx.x != 0) return;
    const double rn = __longlong_as_double((long long)*norm);
    c->rhsnorm = rn;
    c->r = rn;
    c->iters = 0;
    c->diverged = 0;
    c->target = fmax(atol, rtol * rn);
    c->active = (rn <= c->target || maxiter < 1) ? 0 : 1;
}

// after an iteration: stop at convergence, divergence or maxiter (the host
// loop's tests, in its order)
__global__ void k_ctl_step(MgCtl* __restrict__ c, const unsigned long long* __restrict__ norm, int maxiter)
{
    if (threadIdx.x != 0 || !c->active) return;
    const double r = __longlong_as_double((long long)*norm);
    c->r = r;
    c->iters += 1;
    if (r <= c->target)
        c->active = 0;
    else if (!(r <= kDivergeFactor * c->rhsnorm)) {
        c->diverged = 1;
        c->active = 0;
    } else if (c->iters >= maxiter)
        c->active = 0;
}

// the outer update writes phi out of place: after m iterations of the ncap
// enqueued ones the iterate sits in the buffer the host names phi when m and
// ncap have the same parity, else in phi2; bring it into phi
__global__ void __launch_bounds__(256) k_ctl_select(const MgCtl* __restrict__ c, int ncap, double* __restrict__ home,
                                                    const double* __restrict__ other, long long n)
{
    if (((c->iters - ncap) & 1) == 0) return;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        home[t] = other[t];
}

}  // namespace ksgf

using namespace ksgf;

struct ksgf_macproj {
    const ksgf_boxarray* ba = nullptr;
    int lobc[3], hibc[3];
    double beta = 1.0;
    double alpha = 0.0;  // diagonal shift: div(beta grad phi) - alpha phi (ksgf_abeclap_*)
    ksgf_mlmg_options opt;
    int singular = 1;
    int nlev = 0;
    int lc = 0;  // first level handled by the single-workgroup coarse kernel
    std::vector<LevelGeom> G;
    std::vector<int> ratio;               // 3 per level (to the next)
    LevelGeom P{};                        // natural level-0 geometry of phi and rhs (FP64)
    std::vector<double*> cor, res;        // FP64 per level (FP64 mode)
    std::vector<float*> cor32, res32;     // FP32 per level (mixed mode)
    std::vector<int> kz_sweep;            // autotuned planes per workgroup of the packed half sweep
    double* rpart = nullptr;              // per-block maxima of the outer residual
    int kz_res = 16;                      // autotuned planes per thread of the outer residual
    std::vector<int> kz_rst;              // autotuned coarse planes per thread of the restriction from each level
    LevelPtrs<double>* dlp64 = nullptr;   // device level tables for the coarse kernel
    LevelPtrs<float>* dlp32 = nullptr;
    double* phi = nullptr;
    double* phi2 = nullptr;  // ping-pong partner of phi
    double* rhs = nullptr;
    double* part = nullptr;
    unsigned long long* dnorm = nullptr;
    unsigned long long* hnorm = nullptr;  // pinned
    BottomWork bw{};
    int* hstats = nullptr;  // pinned
    // face coefficients (set_beta_faces): the variable-coefficient solver
    // replaces the constant-beta hierarchy until set_beta is called again
    ksgf::VbSolver* vb = nullptr;
    int use_vb = 0;
    // coefficients last written into the level tables (set_beta skips a refill
    // that would change nothing)
    double coef_beta = -1.0, coef_alpha = -1.0;
    // device-controlled solves (project_begin / finish / correct)
    ksgf::MgCtl* ctl = nullptr;   // device
    ksgf::MgCtl* hctl = nullptr;  // pinned
    int ncap = 0;                 // V-cycles the last begin enqueued
    double* phi_a = nullptr;      // the buffer a device-controlled solve leaves the iterate in
    // the projection's solve replayed from a hipGraph (mg_solve_graphed): the
    // device-controlled iterations captured once per key (a few kept: a
    // caller may alternate tolerances), the iteration count of the last
    // solve sizing the next capture
    struct SolveGraph {
        hipGraphExec_t exec = nullptr;
        double rtol, atol, beta, alpha;
        int ncyc, singular;
        long long used;  // replay stamp (the least recently used is replaced)
    };
    std::vector<SolveGraph> sgraphs;
    long long sg_clock = 0;
    hipStream_t cap_st = nullptr;  // capture stream (nothing runs on it)
    int iters_hint = 0;
};

namespace ksgf {

static int mp_alloc(void** p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        set_error("hipMalloc multigrid level");
        return KSGF_ERR_NOMEM;
    }
    if (hipMemset(*p, 0, bytes) != hipSuccess) {
        set_error("hipMemset multigrid level");
        return KSGF_ERR_HIP;
    }
    return KSGF_OK;
}

// stencil coefficients of div(beta grad) - alpha on a level of cell size h
// (alpha, a zeroth-order term, is the same on every level)
static void fill_coefs(LevelGeom& L, const double h[3], double beta, double alpha = 0.0)
{
    L.cx = beta / (h[0] * h[0]);
    L.cy = beta / (h[1] * h[1]);
    L.cz = beta / (h[2] * h[2]);
    L.dg = 2.0 * (L.cx + L.cy + L.cz) + alpha;
    L.idg = 1.0 / L.dg;
}

static std::vector<std::array<double, 3>> level_h(const ksgf_macproj* mp)
{
    std::vector<std::array<double, 3>> h(mp->nlev);
    h[0] = {mp->ba->dx[0], mp->ba->dx[1], mp->ba->dx[2]};
    for (int l = 1; l < mp->nlev; ++l)
        for (int d = 0; d < 3; ++d) h[l][d] = h[l - 1][d] * mp->ratio[3 * (l - 1) + d];
    return h;
}

// LDS placement of the coarse cycle: the bottom solver's vectors first, then
// levels from the bottom up while they fit; never level lc, whose arrays the
// grid kernels read and write
static std::vector<int> lds_layout(ksgf_macproj* mp, size_t word, int* bw_lds)
{
    std::vector<int> off(mp->nlev, -1);
    *bw_lds = -1;
    const int nb = mp->nlev - 1;
    auto need = [&](int l) { return 2LL * lds_align(lds_geom(mp->G[l]).tot * (long long)word); };
    // the bottom solver's eight FP64 vectors take the bottom level's
    // geometry: the pool's when the bottom level is in the pool as well
    const long long bw_pool = 8LL * lds_align(lds_geom(mp->G[nb]).tot * (long long)sizeof(double));
    if (bw_pool + need(nb) <= kLdsPoolBytes) {
        *bw_lds = 0;
        long long used = bw_pool;
        for (int l = nb; l >= mp->lc; --l) {
            if (used + need(l) > kLdsPoolBytes) break;
            off[l] = (int)used;
            used += need(l);
        }
        return off;
    }
    if (8LL * lds_align(mp->G[nb].tot * (long long)sizeof(double)) <= kLdsPoolBytes) *bw_lds = 0;
    return off;
}

static int upload_level_tables(ksgf_macproj* mp)
{
    if (mp->opt.precision == KSGF_MG_MIXED) {
        const std::vector<int> off = lds_layout(mp, sizeof(float), &mp->bw.lds);
        std::vector<LevelPtrs<float>> t(mp->nlev);
        for (int l = 0; l < mp->nlev; ++l)
            t[l] = {mp->G[l], mp->cor32[l], mp->res32[l], mp->ratio[3 * l], mp->ratio[3 * l + 1], mp->ratio[3 * l + 2],
                    off[l]};
        KSGF_HIP_CHECK(hipMemcpy(mp->dlp32, t.data(), sizeof(LevelPtrs<float>) * t.size(), hipMemcpyHostToDevice));
    } else {
        const std::vector<int> off = lds_layout(mp, sizeof(double), &mp->bw.lds);
        std::vector<LevelPtrs<double>> t(mp->nlev);
        for (int l = 0; l < mp->nlev; ++l)
            t[l] = {mp->G[l], mp->cor[l], mp->res[l], mp->ratio[3 * l], mp->ratio[3 * l + 1], mp->ratio[3 * l + 2],
                    off[l]};
        KSGF_HIP_CHECK(hipMemcpy(mp->dlp64, t.data(), sizeof(LevelPtrs<double>) * t.size(), hipMemcpyHostToDevice));
    }
    return KSGF_OK;
}

template <typename T>
static void launch_ghosts(T* a, const LevelGeom& L, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_ghosts<T>, dim3(grid_for(ghost_count(L), 256)), dim3(256), 0, st, a, L);
}

// ---------------------------------------------------------------------------
// live kernel timing for the bench: HIP events on the launch stream around
// every finest-level packed half sweep, with its algorithmic bytes (the
// half of the cells it relaxes: read their rhs, read the other colour,
// write them -- 3 words per relaxed cell, 1.5 per level cell; a red half on
// a zero initial guess reads no other colour; the red half of a
// prolongation sweep also reads the coarse correction once)
// ---------------------------------------------------------------------------
struct KernelProf {
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    double bytes = 0.0;
    long long launches = 0;
    std::string name = "k_pk_half(level 0)";
};
static KernelProf g_prof;

static hipEvent_t prof_event()
{
    if (g_prof.used == g_prof.ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        g_prof.ev.push_back(e);
    }
    return g_prof.ev[g_prof.used++];
}

// half sweeps number `first` .. 2*nu-1 of nu red-black sweeps (half sweep
// 2s is red, 2s+1 black) on a natural-layout level (odd extents, explicit
// ghosts)
template <typename T>
static void launch_half_sweeps(T* x, const T* b, const LevelGeom& L, int first, int nu, hipStream_t st,
                               const T* xc = nullptr, const LevelGeom* C = nullptr, const int* r = nullptr)
{
    long long n = (long long)((L.n0 + 1) / 2) * L.n1 * L.n2;
    for (int h = first; h < 2 * nu; ++h) {
        const T* xch = (xc && h == first) ? xc : nullptr;
        const LevelGeom Cg = C ? *C : L;
        const int r0 = r ? r[0] : 1, r1 = r ? r[1] : 1, r2 = r ? r[2] : 1;
        int color = h & 1;
        if (L.ghost_mode) launch_ghosts<T>(x, L, st);
        hipLaunchKernelGGL(k_mg_smooth<T>, dim3(grid_for(n, 256)), dim3(256), 0, st, x, b, L, color, xch, Cg, r0, r1,
                           r2);
    }
}

template <typename T>
static void launch_fill(T* a, long long n, T v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n, 256)), dim3(256), 0, st, a, n, v);
}

static void launch_sub_mean(double* a, const LevelGeom& L, double* part, hipStream_t st)
{
    hipLaunchKernelGGL(k_partial_sum, dim3(kSumBlocks), dim3(256), 0, st, (const double*)a, L, part);
    hipLaunchKernelGGL(k_sub_mean, dim3(kSumBlocks), dim3(256), 0, st, a, L, (const double*)part, kSumBlocks);
}

// planes per workgroup of the row-tile kernels
constexpr int kRestrictKz = 4;
constexpr int kResidualKzMin = 1;  // candidates 1, 2, 4, 8, 16, 32, 64 planes per thread

template <typename T>
static PkArgs<T> pk_args(const ksgf_macproj* mp, int l, T* x, const T* b, const T* xc)
{
    PkArgs<T> a{};
    a.x = x;
    a.b = b;
    a.xc = xc;
    a.L = mp->G[l];
    a.C = (l + 1 < mp->nlev) ? mp->G[l + 1] : mp->G[l];
    const int* r = &mp->ratio[3 * l];
    a.s0 = (r[0] == 2);
    a.s1 = (r[1] == 2);
    a.s2 = (r[2] == 2);
    return a;
}

template <typename T>
// e0/e1 (optional): events the dispatch itself records at the kernel's
// start and end (hipExtLaunchKernel), so the interval is the kernel's own
// execution time without launch gaps
static void launch_pk_kernel(const PkArgs<T>& a, int color, int mode, int kz, hipStream_t st,
                             hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const dim3 blk(256);
    if (mode != kPkProlong && a.L.prow % 4 == 0) {
        const QuadTiles q = quad_tiles(a.L, kz);
        const dim3 gq(q.nb * ((a.L.n2 + kz - 1) / kz));
        if (mode == kPkZero)
            hipExtLaunchKernelGGL((k_pk_half4<T, kPkZero, false>), gq, blk, 0, st, e0, e1, 0, a, color, q);
        else if (mode == kPkProlongWalls && a.C.packed)
            hipExtLaunchKernelGGL((k_pk_half4<T, kPkProlongWalls, true>), gq, blk, 0, st, e0, e1, 0, a, color, q);
        else if (mode == kPkProlongWalls)
            hipExtLaunchKernelGGL((k_pk_half4<T, kPkProlongWalls, false>), gq, blk, 0, st, e0, e1, 0, a, color, q);
        else
            hipExtLaunchKernelGGL((k_pk_half4<T, kPkPlain, false>), gq, blk, 0, st, e0, e1, 0, a, color, q);
        return;
    }
    const RowTiles t = row_tiles(a.L.prow, a.L.n1, a.L.n2, kz);
    const dim3 g(t.ntx * t.nty * t.nkc);
    if (mode == kPkZero)
        hipExtLaunchKernelGGL((k_pk_half<T, kPkZero, false>), g, blk, 0, st, e0, e1, 0, a, color, t);
    else if (mode == kPkProlongWalls && a.C.packed)
        hipExtLaunchKernelGGL((k_pk_half<T, kPkProlongWalls, true>), g, blk, 0, st, e0, e1, 0, a, color, t);
    else if (mode == kPkProlongWalls)
        hipExtLaunchKernelGGL((k_pk_half<T, kPkProlongWalls, false>), g, blk, 0, st, e0, e1, 0, a, color, t);
    else if (mode == kPkProlong && a.C.packed)
        hipExtLaunchKernelGGL((k_pk_half<T, kPkProlong, true>), g, blk, 0, st, e0, e1, 0, a, color, t);
    else if (mode == kPkProlong)
        hipExtLaunchKernelGGL((k_pk_half<T, kPkProlong, false>), g, blk, 0, st, e0, e1, 0, a, color, t);
    else
        hipExtLaunchKernelGGL((k_pk_half<T, kPkPlain, false>), g, blk, 0, st, e0, e1, 0, a, color, t);
}

// one red-black sweep (red half, black half) of packed level l
template <typename T>
static void launch_pk_sweep(ksgf_macproj* mp, int l, int mode, T* x, const T* b, const T* xc, hipStream_t st,
                            int first_color = 0)
{
    const PkArgs<T> a = pk_args<T>(mp, l, x, b, xc);
    if (mode == kPkProlong) {
        // the coarse correction goes into the black cells in a short pass of
        // its own; the red half then relaxes against plain neighbours (a
        // prolongation inside the red half's six neighbour reads was latency
        // bound at twice a plain half's time), the black half is plain
        const int kz = mp->kz_sweep[l];
        if (a.L.prow % 4 == 0) {
            const QuadTiles q = quad_tiles(a.L, kz);
            const dim3 gq(q.nb * ((a.L.n2 + kz - 1) / kz));
            if (a.C.packed)
                hipLaunchKernelGGL((k_pk_prolong_color4<T, true>), gq, dim3(256), 0, st, a, 1, q);
            else
                hipLaunchKernelGGL((k_pk_prolong_color4<T, false>), gq, dim3(256), 0, st, a, 1, q);
        } else {
            const RowTiles t = row_tiles(a.L.prow, a.L.n1, a.L.n2, kz);
            if (a.C.packed)
                hipLaunchKernelGGL((k_pk_prolong_color<T, true>), dim3(t.ntx * t.nty * t.nkc), dim3(256), 0, st, a, 1,
                                   t);
            else
                hipLaunchKernelGGL((k_pk_prolong_color<T, false>), dim3(t.ntx * t.nty * t.nkc), dim3(256), 0, st, a,
                                   1, t);
        }
    }
    for (int color = first_color; color < 2; ++color) {
        const int m = (mode == kPkProlong) ? (color == 0 ? kPkProlongWalls : kPkPlain) : mode;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (l == 0 && g_prof.on) {
            e0 = prof_event();
            e1 = prof_event();
            if (!e0 || !e1) e0 = e1 = nullptr;
        }
        launch_pk_kernel<T>(a, color, m, mp->kz_sweep[l], st, e0, e1);
        if (e0 && e1) {
            const double half = 0.5 * (double)a.L.n0 * a.L.n1 * a.L.n2;
            const double words = (m == kPkZero && color == 0) ? 2.0 * half : 3.0 * half;
            g_prof.launches += 1;
            g_prof.bytes += sizeof(T) * words;
            const bool quad = (m != kPkProlong && a.L.prow % 4 == 0);
            g_prof.name = std::string(quad ? "k_pk_half4" : "k_pk_half") + (sizeof(T) == 8 ? "<double>" : "<float>") +
                          "(level 0)";
        }
    }
}

// Duration of reps launches issued back to back between two stream events.
// Levels below kTuneSmallCells cells are not timed: their kernels are bound
// by latency, not bandwidth, so they take the most parallel shape (one plane
// per thread); timing them back to back would measure the launch rate.
constexpr long long kTuneSmallCells = 1LL << 20;
template <typename L>
static float time_launches(L&& launch, int reps)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1e30f;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 1e30f;
    }
    (void)hipEventRecord(e0, nullptr);
    for (int rep = 0; rep < reps; ++rep) launch(rep);
    (void)hipEventRecord(e1, nullptr);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 1e30f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

// Autotuner for the planes per workgroup of the packed half sweep: every
// candidate is timed on the level's own arrays (red halves on the
// correction array, which every solve re-initialises before reading), the
// fastest is kept per (level shape, precision) for the life of the process.
static std::map<std::array<long long, 4>, int>& kz_cache()
{
    static std::map<std::array<long long, 4>, int> m;
    return m;
}

template <typename T>
static int tune_sweep(ksgf_macproj* mp, int l, T* x, const T* b)
{
    const LevelGeom& L = mp->G[l];
    // KSGF_SWEEP_PLANES=1|2|4|8 fixes the choice (reproducible launch shapes,
    // e.g. for hardware-counter runs that must match launches by grid size)
    if (const char* env = std::getenv("KSGF_SWEEP_PLANES")) {
        const int v = std::atoi(env);
        if (v == 1 || v == 2 || v == 4 || v == 8) return std::min(v, std::max(1, L.n2));
    }
    if ((long long)L.n0 * L.n1 * L.n2 < kTuneSmallCells) return 1;
    const std::array<long long, 4> key{L.n0, L.n1, L.n2, (long long)sizeof(T)};
    auto it = kz_cache().find(key);
    if (it != kz_cache().end()) return it->second;
    const int cand[5] = {1, 2, 4, 8, 16};
    int best = 1;
    float best_ms = 1e30f;
    const PkArgs<T> a = pk_args<T>(mp, l, x, b, nullptr);
    for (int c : cand) {
        if (c > L.n2) continue;
        launch_pk_kernel<T>(a, 0, kPkPlain, c, nullptr);
        const float ms = time_launches([&](int rep) { launch_pk_kernel<T>(a, rep & 1, kPkPlain, c, nullptr); }, 4);
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    kz_cache()[key] = best;
    return best;
}

// the FP64 outer residual without explicit ghosts at kz planes per thread,
// four cells per thread where the x extent allows; returns the number of
// workgroups (block maxima in mp->rpart)
template <typename C, typename R>
static int launch_ures(ksgf_macproj* mp, const C* cor, R* r, int kz, hipStream_t st)
{
    const LevelGeom& P = mp->P;
    const LevelGeom& L0 = mp->G[0];
    const bool quads = P.n0 % 4 == 0;
    const RowTiles t = row_tiles(quads ? P.n0 / 4 : P.n0, P.n1, P.n2, kz);
    const int nb = t.ntx * t.nty * t.nkc;
    const dim3 g(nb), blk(256);
    const double* phi = mp->phi;
    const double* rhs = mp->rhs;
    if (quads && L0.packed)
        hipLaunchKernelGGL((k_update_residual4<C, R, true>), g, blk, 0, st, phi, cor, mp->phi2, rhs, r, P, L0, t,
                           mp->rpart);
    else if (quads)
        hipLaunchKernelGGL((k_update_residual4<C, R, false>), g, blk, 0, st, phi, cor, mp->phi2, rhs, r, P, L0, t,
                           mp->rpart);
    else if (L0.packed)
        hipLaunchKernelGGL((k_update_residual<C, R, true>), g, blk, 0, st, phi, cor, mp->phi2, rhs, r, P, L0, t,
                           mp->rpart);
    else
        hipLaunchKernelGGL((k_update_residual<C, R, false>), g, blk, 0, st, phi, cor, mp->phi2, rhs, r, P, L0, t,
                           mp->rpart);
    return nb;
}

// Autotuner of the planes per thread of the FP64 outer residual (the
// z-march keeps a column's planes in registers: more planes fetch less halo,
// fewer planes give more threads in flight).  Each candidate is timed as the
// solve calls it, with a correction (phi2 = phi + cor, r = rhs - A phi2) and
// the reduction of its per-block maxima, on the projector's own arrays
// (phi2 and the level-0 residual and correction are scratch until a solve
// writes them); the fastest is cached per (shape, precision).
template <typename T>
static int tune_residual(ksgf_macproj* mp)
{
    static std::map<std::array<long long, 4>, int> cache;
    const LevelGeom& P = mp->P;
    const std::array<long long, 4> key{P.n0, P.n1, P.n2, (long long)sizeof(T)};
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    T* r = (sizeof(T) == sizeof(float)) ? (T*)mp->res32[0] : (T*)mp->res[0];
    const T* cor = (sizeof(T) == sizeof(float)) ? (const T*)mp->cor32[0] : (const T*)mp->cor[0];
    hipEvent_t e0, e1;
    int best = 16;
    if (hipEventCreate(&e0) != hipSuccess) return best;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return best;
    }
    float best_ms = 1e30f;
    for (int kz : {1, 2, 4, 8, 16, 32, 64}) {
        if (kz < kResidualKzMin || kz > std::max(P.n2, kResidualKzMin)) continue;
        auto launch = [&]() {
            const int nb = launch_ures<T, T>(mp, cor, r, kz, nullptr);
            hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(kMaxPartsThreads), 0, nullptr, (const double*)mp->rpart, nb,
                               mp->dnorm, (const int*)nullptr);
        };
        launch();
        (void)hipEventRecord(e0, nullptr);
        for (int rep = 0; rep < 3; ++rep) launch();
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
        if (ms < best_ms) {
            best_ms = ms;
            best = kz;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cache[key] = best;
    return best;
}

// Autotuner of the coarse planes per thread of the residual + restriction
// from level l (candidates 1, 2, 4, 8, 16), timed on the level's own correction
// and residual arrays into the next level's residual (all scratch at set-up),
// cached per (fine shape, coarse shape, precision)
template <typename T>
// cred: the coarse correction, when the restriction also writes the red
// half of the coarse level's zero-initial-guess sweep (packed coarse level)
static void launch_restrict(ksgf_macproj* mp, int l, const T* x, const T* b, T* out, int kz, hipStream_t st,
                            T* cred = nullptr)
{
    const LevelGeom& L = mp->G[l];
    const LevelGeom& C = mp->G[l + 1];
    const int* r = &mp->ratio[3 * l];
    const T w = (T)(1.0 / (double)(r[0] * r[1] * r[2]));
    if (!C.packed || !L.packed) cred = nullptr;
    if (r[0] == 1 && L.packed && L.n0 % 4 == 0) {
        const RowTiles t = row_tiles(C.n0 / 4, C.n1, C.n2, kz);
        hipLaunchKernelGGL(k_resid_restrict4<T>, dim3(t.ntx * t.nty * t.nkc), dim3(256), 0, st, x, b, out, cred, L, C,
                           t, r[1], r[2], w);
        return;
    }
    const RowTiles t = row_tiles(C.n0, C.n1, C.n2, kz);
    hipLaunchKernelGGL(k_resid_restrict_rt<T>, dim3(t.ntx * t.nty * t.nkc), dim3(256), 0, st, x, b, out, cred, L, C,
                       t, r[0], r[1], r[2], w);
}

// whether the restriction into level l writes the red half of l's first
// (zero-initial-guess) sweep: packed grid levels above the coarse kernel's
template <typename T>
static bool restrict_writes_red(const ksgf_macproj* mp, int l)
{
    return l >= 1 && l < mp->lc && mp->G[l].packed && mp->G[l - 1].packed && mp->opt.num_pre_smooth > 0;
}

template <typename T>
static int tune_restrict(ksgf_macproj* mp, int l, const T* x, const T* b, T* out)
{
    const LevelGeom& L = mp->G[l];
    const LevelGeom& C = mp->G[l + 1];
    static std::map<std::array<long long, 7>, int> cache;
    if ((long long)L.n0 * L.n1 * L.n2 < kTuneSmallCells) return 1;
    const std::array<long long, 7> key{L.n0, L.n1, L.n2, C.n0, C.n1, C.n2, (long long)sizeof(T)};
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int best = 1;
    float best_ms = 1e30f;
    for (int kz : {1, 2, 4, 8, 16}) {
        if (kz > C.n2) continue;
        launch_restrict<T>(mp, l, x, b, out, kz, nullptr);
        const float ms = time_launches([&](int) { launch_restrict<T>(mp, l, x, b, out, kz, nullptr); }, 3);
        if (ms < best_ms) {
            best_ms = ms;
            best = kz;
        }
    }
    cache[key] = best;
    return best;
}

// V-cycle in precision T from level l down (levels >= lc in one workgroup)
template <typename T>
static void vcycle(ksgf_macproj* mp, int l, std::vector<T*>& cor, std::vector<T*>& res, const LevelPtrs<T>* dlp,
                   hipStream_t st)
{
    if (l >= mp->lc) {
        hipLaunchKernelGGL(k_coarse_vcycle<T>, dim3(1), dim3(kCoarseThreads), 0, st, dlp, l, mp->nlev, mp->bw,
                           mp->singular, mp->opt.bottom_rtol, mp->opt.bottom_atol, mp->opt.bottom_maxiter,
                           mp->opt.num_bottom_smooth, mp->opt.num_final_smooth, mp->opt.num_pre_smooth,
                           mp->opt.num_post_smooth, mp->G[0].act);
        return;
    }
    const LevelGeom& L = mp->G[l];
    const LevelGeom& C = mp->G[l + 1];
    const int* r = &mp->ratio[3 * l];
    const long long nf = (long long)L.n0 * L.n1 * L.n2;
    const int nu1 = mp->opt.num_pre_smooth, nu2 = mp->opt.num_post_smooth;
    // down
    if (L.packed) {
        if (nu1 == 0) launch_fill<T>(cor[l], L.tot, (T)0, st);
        // the restriction into this level has formed the red half of the first sweep
        const int red_done = restrict_writes_red<T>(mp, l) ? 1 : 0;
        for (int s = 0; s < nu1; ++s)
            launch_pk_sweep<T>(mp, l, s == 0 ? kPkZero : kPkPlain, cor[l], res[l], nullptr, st, s == 0 ? red_done : 0);
    } else if (nu1 > 0 && !L.ghost_mode) {
        hipLaunchKernelGGL(k_mg_smooth_first<T>, dim3(grid_for(nf, 256)), dim3(256), 0, st, cor[l], (const T*)res[l], L);
        launch_half_sweeps<T>(cor[l], res[l], L, 1, nu1, st);
    } else {
        launch_fill<T>(cor[l], L.tot, (T)0, st);
        launch_half_sweeps<T>(cor[l], res[l], L, 0, nu1, st);
    }
    if (L.ghost_mode) launch_ghosts<T>(cor[l], L, st);
    launch_restrict<T>(mp, l, (const T*)cor[l], (const T*)res[l], res[l + 1], mp->kz_rst[l], st,
                       restrict_writes_red<T>(mp, l + 1) ? cor[l + 1] : nullptr);
    vcycle<T>(mp, l + 1, cor, res, dlp, st);
    // up
    if (L.packed) {
        if (nu2 == 0) {
            RowTiles t = row_tiles(L.n0, L.n1, L.n2, kRestrictKz);
            hipLaunchKernelGGL(k_pk_prolong_add<T>, dim3(t.ntx * t.nty * t.nkc), dim3(256), 0, st, cor[l],
                               (const T*)cor[l + 1], L, C, (int)(r[0] == 2), (int)(r[1] == 2), (int)(r[2] == 2), t);
        }
        for (int s = 0; s < nu2; ++s)
            launch_pk_sweep<T>(mp, l, s == 0 ? kPkProlong : kPkPlain, cor[l], res[l], s == 0 ? cor[l + 1] : nullptr,
                               st);
    } else if (nu2 > 0 && !L.ghost_mode) {
        hipLaunchKernelGGL(k_mg_prolong_red<T>, dim3(grid_for(nf, 256)), dim3(256), 0, st, cor[l], (const T*)cor[l + 1],
                           (const T*)res[l], L, C, r[0], r[1], r[2]);
        launch_half_sweeps<T>(cor[l], res[l], L, 1, nu2, st, (const T*)cor[l + 1], &C, r);
    } else {
        hipLaunchKernelGGL(k_mg_prolong_add<T>, dim3(grid_for(nf, 256)), dim3(256), 0, st, cor[l], (const T*)cor[l + 1],
                           L, C, r[0], r[1], r[2]);
        launch_half_sweeps<T>(cor[l], res[l], L, 0, nu2, st);
    }
}

static int read_norm(ksgf_macproj* mp, hipStream_t st, double* out)
{
    KSGF_HIP_CHECK(hipMemcpyAsync(mp->hnorm, mp->dnorm, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    unsigned long long bits = *mp->hnorm;
    double v;
    std::memcpy(&v, &bits, sizeof(double));
    *out = v;
    return KSGF_OK;
}

template <typename C, typename R>
static void launch_update_residual(ksgf_macproj* mp, const C* cor, R* r, hipStream_t st)
{
    const LevelGeom& P = mp->P;
    const LevelGeom& L0 = mp->G[0];
    const RowTiles t = row_tiles(P.n0, P.n1, P.n2, mp->kz_res);
    int nb = t.ntx * t.nty * t.nkc;
    if (P.ghost_mode) {
        // explicit ghosts: phi += cor first, then the residual from ghosts
        if (cor) {
            hipLaunchKernelGGL((k_update_residual<C, R, false>), dim3(nb), dim3(256), 0, st, (const double*)mp->phi, cor,
                               mp->phi2, (const double*)mp->rhs, r, P, L0, t, mp->rpart);
            std::swap(mp->phi, mp->phi2);
        }
        launch_ghosts<double>(mp->phi, P, st);
        hipLaunchKernelGGL((k_update_residual<C, R, false>), dim3(nb), dim3(256), 0, st, (const double*)mp->phi,
                           (const C*)nullptr, (double*)nullptr, (const double*)mp->rhs, r, P, L0, t, mp->rpart);
    } else {
        nb = launch_ures<C, R>(mp, cor, r, mp->kz_res, st);
        if (cor) std::swap(mp->phi, mp->phi2);
    }
    hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(kMaxPartsThreads), 0, st, (const double*)mp->rpart, nb, mp->dnorm,
                       mp->P.act);
}

// solve with the level-0 rhs in mp->rhs (ghost layout); phi in mp->phi
static int mg_solve(ksgf_macproj* mp, double rtol, double atol, ksgf_mlmg_info* info, hipStream_t st)
{
    const LevelGeom& L0 = mp->P;  // phi / rhs
    const bool mixed = (mp->opt.precision == KSGF_MG_MIXED);
    launch_fill<double>(mp->phi, L0.tot, 0.0, st);
    if (mp->singular) launch_sub_mean(mp->rhs, L0, mp->part, st);
    KSGF_HIP_CHECK(hipMemsetAsync(mp->bw.stats, 0, 2 * sizeof(int), st));
    // |rhs|_inf and the first residual (= rhs)
    if (mixed)
        launch_update_residual<float, float>(mp, nullptr, mp->res32[0], st);
    else
        launch_update_residual<double, double>(mp, nullptr, mp->res[0], st);
    KSGF_LAUNCH_CHECK();
    double rhsnorm;
    int e = read_norm(mp, st, &rhsnorm);
    if (e) return e;
    const double target = std::fmax(atol, rtol * rhsnorm);
    double r = rhsnorm;
    ksgf_mlmg_info inf{};
    inf.r0 = r;
    inf.nlevels = mp->nlev;
    inf.converged = (r <= target);
    bool diverged = false;
    if (!inf.converged) {
        for (int it = 1; it <= mp->opt.maxiter; ++it) {
            if (mixed) {
                vcycle<float>(mp, 0, mp->cor32, mp->res32, mp->dlp32, st);
                launch_update_residual<float, float>(mp, mp->cor32[0], mp->res32[0], st);
            } else {
                vcycle<double>(mp, 0, mp->cor, mp->res, mp->dlp64, st);
                launch_update_residual<double, double>(mp, mp->cor[0], mp->res[0], st);
            }
            KSGF_LAUNCH_CHECK();
            e = read_norm(mp, st, &r);
            if (e) return e;
            inf.iters = it;
            if (r <= target) {
                inf.converged = 1;
                break;
            }
            if (!(r <= kDivergeFactor * rhsnorm)) {
                diverged = true;
                break;
            }
        }
    }
    launch_ghosts<double>(mp->phi, L0, st);
    KSGF_HIP_CHECK(hipMemcpyAsync(mp->hstats, mp->bw.stats, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    inf.rfinal = r;
    inf.outer_iters = mixed ? inf.iters : 0;
    inf.bottom_iters = mp->hstats[0];
    inf.bottom_fail = mp->hstats[1];
    if (info) *info = inf;
    if (diverged) {
        set_error("MLMG: the residual " + std::to_string(r) + " exceeds 1e20 |rhs| = " + std::to_string(rhsnorm) +
                  " after " + std::to_string(inf.iters) + " iterations (diverging)");
        return KSGF_ERR_DIVERGED;
    }
    return inf.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
}

static void set_act(ksgf_macproj* mp, const int* act)
{
    for (auto& L : mp->G) L.act = act;
    mp->P.act = act;
}

// mg_solve's kernels with the convergence test on the device: the first
// residual, then ncyc iterations whose kernels return at once after the test
// has stopped the solve, then the iterate brought into the buffer the host
// names phi.  No host synchronisation and no memset: a stream capture can
// record it.
static int mg_solve_begin(ksgf_macproj* mp, double rtol, double atol, int ncyc, hipStream_t st)
{
    // an even number of enqueued iterations ends in the buffer it starts
    // from: phi_a, whatever replays or host-side solves ran in between
    if (mp->phi != mp->phi_a) std::swap(mp->phi, mp->phi2);
    ncyc += ncyc & 1;
    const LevelGeom& L0 = mp->P;
    const bool mixed = (mp->opt.precision == KSGF_MG_MIXED);
    set_act(mp, nullptr);
    launch_fill<double>(mp->phi, L0.tot, 0.0, st);
    if (mp->singular) launch_sub_mean(mp->rhs, L0, mp->part, st);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(256), 0, st, mp->bw.stats, 2LL, 0);
    if (mixed)
        launch_update_residual<float, float>(mp, nullptr, mp->res32[0], st);
    else
        launch_update_residual<double, double>(mp, nullptr, mp->res[0], st);
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, st, mp->ctl, (const unsigned long long*)mp->dnorm, rtol, atol,
                       mp->opt.maxiter);
    set_act(mp, &mp->ctl->active);
    for (int it = 1; it <= ncyc; ++it) {
        if (mixed) {
            vcycle<float>(mp, 0, mp->cor32, mp->res32, mp->dlp32, st);
            launch_update_residual<float, float>(mp, mp->cor32[0], mp->res32[0], st);
        } else {
            vcycle<double>(mp, 0, mp->cor, mp->res, mp->dlp64, st);
            launch_update_residual<double, double>(mp, mp->cor[0], mp->res[0], st);
        }
        hipLaunchKernelGGL(k_ctl_step, dim3(1), dim3(64), 0, st, mp->ctl, (const unsigned long long*)mp->dnorm,
                           mp->opt.maxiter);
    }
    set_act(mp, nullptr);
    if (ncyc > 0)
        hipLaunchKernelGGL(k_ctl_select, dim3(std::min<long long>(1024, grid_for(L0.tot, 256))), dim3(256), 0, st,
                           (const MgCtl*)mp->ctl, ncyc, mp->phi, (const double*)mp->phi2, L0.tot);
    KSGF_LAUNCH_CHECK();
    mp->ncap = ncyc;
    return KSGF_OK;
}

// read the device-controlled solve's state (one host synchronisation); an
// unfinished solve continues with host-side tests, as mg_solve iterates
static int mg_solve_finish(ksgf_macproj* mp, ksgf_mlmg_info* info, hipStream_t st)
{
    const bool mixed = (mp->opt.precision == KSGF_MG_MIXED);
    KSGF_HIP_CHECK(hipMemcpyAsync(mp->hctl, mp->ctl, sizeof(MgCtl), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    const MgCtl c = *mp->hctl;
    if (mp->phi != mp->phi_a) std::swap(mp->phi, mp->phi2);  // begin left the iterate in phi_a
    ksgf_mlmg_info inf{};
    inf.r0 = c.rhsnorm;
    inf.nlevels = mp->nlev;
    inf.iters = c.iters;
    double r = c.r;
    bool diverged = c.diverged != 0;
    inf.converged = (r <= c.target);
    if (c.active) {
        double* home = mp->phi;
        for (int it = c.iters + 1; it <= mp->opt.maxiter; ++it) {
            if (mixed) {
                vcycle<float>(mp, 0, mp->cor32, mp->res32, mp->dlp32, st);
                launch_update_residual<float, float>(mp, mp->cor32[0], mp->res32[0], st);
            } else {
                vcycle<double>(mp, 0, mp->cor, mp->res, mp->dlp64, st);
                launch_update_residual<double, double>(mp, mp->cor[0], mp->res[0], st);
            }
            KSGF_LAUNCH_CHECK();
            int e = read_norm(mp, st, &r);
            if (e) return e;
            inf.iters = it;
            if (r <= c.target) {
                inf.converged = 1;
                break;
            }
            if (!(r <= kDivergeFactor * c.rhsnorm)) {
                diverged = true;
                break;
            }
        }
        if (mp->phi != home) {  // keep the iterate where a captured correction reads it
            KSGF_HIP_CHECK(hipMemcpyAsync(home, mp->phi, sizeof(double) * (size_t)mp->P.tot, hipMemcpyDeviceToDevice,
                                          st));
            std::swap(mp->phi, mp->phi2);
        }
    }
    KSGF_HIP_CHECK(hipMemcpyAsync(mp->hstats, mp->bw.stats, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    inf.rfinal = r;
    inf.outer_iters = mixed ? inf.iters : 0;
    inf.bottom_iters = mp->hstats[0];
    inf.bottom_fail = mp->hstats[1];
    if (info) *info = inf;
    if (diverged) {
        set_error("MLMG: the residual " + std::to_string(r) + " exceeds 1e20 |rhs| = " + std::to_string(c.rhsnorm) +
                  " after " + std::to_string(inf.iters) + " iterations (diverging)");
        return KSGF_ERR_DIVERGED;
    }
    return inf.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
}

// mg_solve with its iterations replayed from a hipGraph: the first solve of
// a projector runs mg_solve (host-tested) and records its iteration count m;
// later solves replay mg_solve_begin's device-controlled iterations, m
// rounded down to even of them (captured once per count, tolerances and
// coefficients), and mg_solve_finish reads the state once and finishes an
// unconverged solve with host-side tests.  Iterates, residuals and statistics
// equal mg_solve's bit for bit; the host waits once per solve instead of once
// per V-cycle, and the V-cycles' launches leave no gaps.
// With kernel timing on (ksgf_prof_reset) every solve runs mg_solve, whose
// level-0 half sweeps each carry their own start / end events.
static int mg_solve_graphed(ksgf_macproj* mp, double rtol, double atol, ksgf_mlmg_info* info, hipStream_t st)
{
    ksgf_mlmg_info inf{};
    const int ncyc = std::min(mp->iters_hint & ~1, mp->opt.maxiter & ~1);
    if (ncyc < 2 || g_prof.on) {
        const int r = mg_solve(mp, rtol, atol, &inf, st);
        if (info) *info = inf;
        if (r >= 0) mp->iters_hint = inf.iters;
        return r;
    }
    constexpr size_t kMaxGraphs = 4;
    ksgf_macproj::SolveGraph* sg = nullptr;
    for (auto& g : mp->sgraphs)
        if (g.ncyc == ncyc && g.rtol == rtol && g.atol == atol && g.beta == mp->coef_beta &&
            g.alpha == mp->coef_alpha && g.singular == mp->singular)
            sg = &g;
    if (!sg) {
        if (!mp->cap_st && hipStreamCreateWithFlags(&mp->cap_st, hipStreamNonBlocking) != hipSuccess) {
            mp->cap_st = nullptr;
            set_error("MLMG: capture stream");
            return KSGF_ERR_HIP;
        }
        KSGF_HIP_CHECK(hipStreamBeginCapture(mp->cap_st, hipStreamCaptureModeThreadLocal));
        const int e = mg_solve_begin(mp, rtol, atol, ncyc, mp->cap_st);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(mp->cap_st, &g);
        if (e != KSGF_OK || ee != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            if (e == KSGF_OK) set_error("MLMG: solve capture");
            return e != KSGF_OK ? e : KSGF_ERR_HIP;
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            set_error("MLMG: solve graph instantiate");
            return KSGF_ERR_HIP;
        }
        if (mp->sgraphs.size() >= kMaxGraphs) {
            auto old = std::min_element(mp->sgraphs.begin(), mp->sgraphs.end(),
                                        [](const auto& x, const auto& y) { return x.used < y.used; });
            (void)hipGraphExecDestroy(old->exec);
            mp->sgraphs.erase(old);
        }
        mp->sgraphs.push_back({exec, rtol, atol, mp->coef_beta, mp->coef_alpha, ncyc, mp->singular, 0});
        sg = &mp->sgraphs.back();
    }
    sg->used = ++mp->sg_clock;
    KSGF_HIP_CHECK(hipGraphLaunch(sg->exec, st));
    mp->ncap = ncyc;
    const int r = mg_solve_finish(mp, &inf, st);
    launch_ghosts<double>(mp->phi, mp->P, st);  // as mg_solve leaves phi
    if (info) *info = inf;
    if (r >= 0) mp->iters_hint = inf.iters;
    return r;
}

}  // namespace ksgf

namespace ksgf {
// the projector's V-cycles (opts.maxiter of them, or fewer if the residual
// reaches zero) on rhs, phi out whether or not the tolerance was met: the
// coarse correction of a composite two-level solve
int macproj_vcycles(ksgf_macproj* mp, const double* rhs, double* phi, ksgf_mlmg_info* info, hipStream_t st)
{
    const LevelGeom& L0 = mp->P;
    long long n = (long long)L0.n0 * L0.n1 * L0.n2;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, st, rhs, mp->rhs, L0, 1);
    KSGF_LAUNCH_CHECK();
    int ret = mg_solve(mp, 0.0, 0.0, info, st);
    if (ret < 0 && ret != KSGF_ERR_NOT_CONVERGED) return ret;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, st, (const double*)mp->phi, phi, L0, 0);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}
}  // namespace ksgf

extern "C" {

int ksgf_prof_reset(void)
{
    g_prof.on = true;
    g_prof.used = 0;
    g_prof.bytes = 0.0;
    g_prof.launches = 0;
    return KSGF_OK;
}

int ksgf_prof_read(double* ms, long long* launches, double* bytes, char* name, int name_len)
{
    if (!ms || !launches || !bytes) return KSGF_ERR_INVALID;
    double tot = 0.0;
    for (size_t i = 0; i + 1 < g_prof.used; i += 2) {
        KSGF_HIP_CHECK(hipEventSynchronize(g_prof.ev[i + 1]));
        float t = 0.f;
        KSGF_HIP_CHECK(hipEventElapsedTime(&t, g_prof.ev[i], g_prof.ev[i + 1]));
        tot += t;
    }
    *ms = tot;
    *launches = g_prof.launches;
    *bytes = g_prof.bytes;
    if (name && name_len > 0) {
        std::strncpy(name, g_prof.name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    g_prof.on = false;
    return KSGF_OK;
}

void ksgf_mlmg_default_options(ksgf_mlmg_options* o)
{
    if (!o) return;
    o->maxiter = 200;
    o->num_pre_smooth = 2;
    o->num_post_smooth = 2;
    o->num_final_smooth = 8;
    o->num_bottom_smooth = 0;
    o->bottom_maxiter = 200;
    o->bottom_rtol = 2.220446049250313e-16 * 1.0e12;
    o->bottom_atol = 2.220446049250313e-16 * 1.0e4;
    o->do_semicoarsening = 0;
    o->max_coarsening_level = 30;
    o->precision = KSGF_MG_FP64;
}

int ksgf_macproj_destroy(ksgf_macproj* mp)
{
    if (!mp) return KSGF_OK;
    for (auto* p : mp->cor) if (p) (void)hipFree(p);
    for (auto* p : mp->res) if (p) (void)hipFree(p);
    for (auto* p : mp->cor32) if (p) (void)hipFree(p);
    for (auto* p : mp->res32) if (p) (void)hipFree(p);
    if (mp->rpart) (void)hipFree(mp->rpart);
    double* bws[8] = {mp->bw.r, mp->bw.rh, mp->bw.p, mp->bw.v, mp->bw.s, mp->bw.t, mp->bw.x, mp->bw.b};
    for (auto* p : bws) if (p) (void)hipFree(p);
    if (mp->bw.stats) (void)hipFree(mp->bw.stats);
    if (mp->dlp64) (void)hipFree(mp->dlp64);
    if (mp->dlp32) (void)hipFree(mp->dlp32);
    if (mp->phi) (void)hipFree(mp->phi);
    if (mp->phi2) (void)hipFree(mp->phi2);
    if (mp->rhs) (void)hipFree(mp->rhs);
    if (mp->part) (void)hipFree(mp->part);
    if (mp->dnorm) (void)hipFree(mp->dnorm);
    if (mp->hnorm) (void)hipHostFree(mp->hnorm);
    if (mp->hstats) (void)hipHostFree(mp->hstats);
    if (mp->ctl) (void)hipFree(mp->ctl);
    if (mp->hctl) (void)hipHostFree(mp->hctl);
    for (auto& g : mp->sgraphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (mp->cap_st) (void)hipStreamDestroy(mp->cap_st);
    vb_destroy(mp->vb);
    delete mp;
    return KSGF_OK;
}

int ksgf_macproj_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc, double beta,
                        const ksgf_mlmg_options* opts, ksgf_macproj** out)
{
    if (!ba || !lobc || !hibc || !out || !(beta > 0.0)) {
        set_error("ksgf_macproj_create: bad argument");
        return KSGF_ERR_INVALID;
    }
    ksgf_macproj* mp = new ksgf_macproj();
    mp->ba = ba;
    mp->beta = beta;
    if (opts)
        mp->opt = *opts;
    else
        ksgf_mlmg_default_options(&mp->opt);
    if (mp->opt.precision != KSGF_MG_FP64 && mp->opt.precision != KSGF_MG_MIXED) {
        delete mp;
        set_error("ksgf_macproj_create: bad precision");
        return KSGF_ERR_INVALID;
    }
    if (mp->opt.num_pre_smooth < 0 || mp->opt.num_post_smooth < 0 || mp->opt.maxiter < 0) {
        delete mp;
        set_error("ksgf_macproj_create: negative smoothing/iteration count");
        return KSGF_ERR_INVALID;
    }
    mp->singular = 1;
    for (int d = 0; d < 3; ++d) {
        mp->lobc[d] = lobc[d];
        mp->hibc[d] = hibc[d];
        for (int t : {lobc[d], hibc[d]})
            if (t != KSGF_LO_PERIODIC && t != KSGF_LO_NEUMANN && t != KSGF_LO_DIRICHLET) {
                delete mp;
                set_error("ksgf_macproj_create: bad LinOp BC");
                return KSGF_ERR_INVALID;
            }
        if ((lobc[d] == KSGF_LO_PERIODIC) != (ba->periodic[d] != 0) ||
            (hibc[d] == KSGF_LO_PERIODIC) != (ba->periodic[d] != 0)) {
            delete mp;
            set_error("ksgf_macproj_create: periodic LinOp BC must match the geometry");
            return KSGF_ERR_INVALID;
        }
        if (lobc[d] == KSGF_LO_DIRICHLET || hibc[d] == KSGF_LO_DIRICHLET) mp->singular = 0;
    }
    // hierarchy
    int cn[3];
    double ch[3];
    for (int d = 0; d < 3; ++d) {
        cn[d] = ba->dhi[d] - ba->dlo[d] + 1;
        ch[d] = ba->dx[d];
    }
    for (;;) {
        LevelGeom L{};
        L.n0 = cn[0];
        L.n1 = cn[1];
        L.n2 = cn[2];
        level_layout(L, false);
        L.ghost_mode = 0;
        for (int d = 0; d < 3; ++d) {
            L.lo[d] = lobc[d];
            L.hi[d] = hibc[d];
            if (lobc[d] == KSGF_LO_PERIODIC && (cn[d] % 2) == 1) L.ghost_mode = 1;
        }
        fill_coefs(L, ch, beta);
        mp->G.push_back(L);
        mp->nlev++;
        int r[3] = {1, 1, 1};
        bool stop = (mp->nlev >= 40 || mp->nlev > mp->opt.max_coarsening_level);
        if (!stop) {
            int can[3], any = 0;
            for (int d = 0; d < 3; ++d) can[d] = (cn[d] % 2 == 0 && cn[d] / 2 >= 2);
            if (!mp->opt.do_semicoarsening) {
                if (can[0] && can[1] && can[2]) {
                    r[0] = r[1] = r[2] = 2;
                    any = 1;
                }
            } else {
                // h_min over every direction, also one that cannot coarsen
                // any more: the others then stop as soon as they would exceed
                // 1.5x its cell size, the anisotropy a point smoother damps
                double hmin = 1e300;
                for (int d = 0; d < 3; ++d)
                    if (ch[d] < hmin) hmin = ch[d];
                for (int d = 0; d < 3; ++d)
                    if (can[d] && ch[d] <= 1.5 * hmin) {
                        r[d] = 2;
                        any = 1;
                    }
            }
            if (!any) stop = true;
        }
        for (int d = 0; d < 3; ++d) mp->ratio.push_back(r[d]);
        if (stop) break;
        for (int d = 0; d < 3; ++d) {
            cn[d] /= r[d];
            ch[d] *= r[d];
        }
    }
    // first level of the single-workgroup coarse part (the bottom at least)
    mp->lc = mp->nlev - 1;
    while (mp->lc > 0 && (long long)mp->G[mp->lc - 1].n0 * mp->G[mp->lc - 1].n1 * mp->G[mp->lc - 1].n2 <= kCoarseCells)
        mp->lc--;
    if (mp->G[0].tot >= (1LL << 31)) {
        ksgf_macproj_destroy(mp);
        set_error("ksgf_macproj_create: level 0 exceeds 2^31 elements");
        return KSGF_ERR_UNSUPPORTED;
    }
    // phi and rhs stay in the natural layout; grid levels without explicit
    // ghosts and with an even x extent take the red-black packed layout
    mp->P = mp->G[0];
    for (int l = 0; l < mp->lc; ++l)
        if (!mp->G[l].ghost_mode && mp->G[l].n0 % 2 == 0) level_layout(mp->G[l], true);
    const LevelGeom& Lb = mp->G[mp->nlev - 1];
    if ((long long)Lb.n0 * Lb.n1 * Lb.n2 > (1LL << 22)) {
        ksgf_macproj_destroy(mp);
        set_error("ksgf_macproj_create: coarsest level too large for the single-workgroup bottom solve");
        return KSGF_ERR_UNSUPPORTED;
    }
    int e;
    const bool mixed = (mp->opt.precision == KSGF_MG_MIXED);
    for (int l = 0; l < mp->nlev; ++l) {
        size_t n = (size_t)mp->G[l].tot;
        void *a, *b;
        const size_t w = mixed ? sizeof(float) : sizeof(double);
        if ((e = mp_alloc(&a, n * w)) || (e = mp_alloc(&b, n * w))) {
            ksgf_macproj_destroy(mp);
            return e;
        }
        if (mixed) {
            mp->cor32.push_back((float*)a);
            mp->res32.push_back((float*)b);
        } else {
            mp->cor.push_back((double*)a);
            mp->res.push_back((double*)b);
        }
    }
    void* p;
    const size_t n0 = (size_t)mp->P.tot;
    if ((e = mp_alloc(&p, n0 * sizeof(double)))) { ksgf_macproj_destroy(mp); return e; }
    mp->phi = (double*)p;
    mp->phi_a = mp->phi;
    if ((e = mp_alloc(&p, n0 * sizeof(double)))) { ksgf_macproj_destroy(mp); return e; }
    mp->phi2 = (double*)p;
    if ((e = mp_alloc(&p, n0 * sizeof(double)))) { ksgf_macproj_destroy(mp); return e; }
    mp->rhs = (double*)p;
    if ((e = mp_alloc(&p, kSumBlocks * sizeof(double)))) { ksgf_macproj_destroy(mp); return e; }
    mp->part = (double*)p;
    {
        // sized for the smallest candidate of the residual tuner
        const RowTiles t = row_tiles(mp->P.n0, mp->P.n1, mp->P.n2, kResidualKzMin);
        if ((e = mp_alloc(&p, (size_t)t.ntx * t.nty * t.nkc * sizeof(double)))) { ksgf_macproj_destroy(mp); return e; }
        mp->rpart = (double*)p;
    }
    double** bws[8] = {&mp->bw.r, &mp->bw.rh, &mp->bw.p, &mp->bw.v, &mp->bw.s, &mp->bw.t, &mp->bw.x, &mp->bw.b};
    for (auto pp : bws) {
        if ((e = mp_alloc(&p, (size_t)Lb.tot * sizeof(double)))) {
            ksgf_macproj_destroy(mp);
            return e;
        }
        *pp = (double*)p;
    }
    if ((e = mp_alloc(&p, 2 * sizeof(int)))) { ksgf_macproj_destroy(mp); return e; }
    mp->bw.stats = (int*)p;
    if ((e = mp_alloc(&p, sizeof(unsigned long long)))) { ksgf_macproj_destroy(mp); return e; }
    mp->dnorm = (unsigned long long*)p;
    if (mixed) {
        if ((e = mp_alloc(&p, sizeof(LevelPtrs<float>) * mp->nlev))) { ksgf_macproj_destroy(mp); return e; }
        mp->dlp32 = (LevelPtrs<float>*)p;
    } else {
        if ((e = mp_alloc(&p, sizeof(LevelPtrs<double>) * mp->nlev))) { ksgf_macproj_destroy(mp); return e; }
        mp->dlp64 = (LevelPtrs<double>*)p;
    }
    if ((e = mp_alloc(&p, sizeof(MgCtl)))) { ksgf_macproj_destroy(mp); return e; }
    mp->ctl = (MgCtl*)p;
    if (hipHostMalloc(&mp->hnorm, 8, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&mp->hstats, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&mp->hctl, sizeof(MgCtl), hipHostMallocDefault) != hipSuccess) {
        ksgf_macproj_destroy(mp);
        set_error("ksgf_macproj_create: pinned scalars");
        return KSGF_ERR_NOMEM;
    }
    if ((e = upload_level_tables(mp))) {
        ksgf_macproj_destroy(mp);
        return e;
    }
    mp->coef_beta = beta;
    mp->coef_alpha = mp->alpha;
    if (!mp->P.ghost_mode) mp->kz_res = mixed ? tune_residual<float>(mp) : tune_residual<double>(mp);
    mp->kz_sweep.assign(mp->nlev, 8);
    mp->kz_rst.assign(mp->nlev, 1);
    for (int l = 0; l < mp->lc; ++l) {
        if (mixed)
            mp->kz_rst[l] = tune_restrict<float>(mp, l, mp->cor32[l], mp->res32[l], mp->res32[l + 1]);
        else
            mp->kz_rst[l] = tune_restrict<double>(mp, l, mp->cor[l], mp->res[l], mp->res[l + 1]);
        if (!mp->G[l].packed) continue;
        mp->kz_sweep[l] = mixed ? tune_sweep<float>(mp, l, mp->cor32[l], mp->res32[l])
                                : tune_sweep<double>(mp, l, mp->cor[l], mp->res[l]);
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        ksgf_macproj_destroy(mp);
        set_error("ksgf_macproj_create: device synchronize");
        return KSGF_ERR_HIP;
    }
    *out = mp;
    return KSGF_OK;
}

int ksgf_macproj_set_beta(ksgf_macproj* mp, double beta)
{
    if (!mp || !(beta > 0.0)) {
        set_error("ksgf_macproj_set_beta: bad argument");
        return KSGF_ERR_INVALID;
    }
    mp->beta = beta;
    mp->use_vb = 0;
    if (beta == mp->coef_beta && mp->alpha == mp->coef_alpha) return KSGF_OK;  // the tables hold these already
    auto h = level_h(mp);
    for (int l = 0; l < mp->nlev; ++l) fill_coefs(mp->G[l], h[l].data(), beta, mp->alpha);
    fill_coefs(mp->P, h[0].data(), beta, mp->alpha);
    const int e = upload_level_tables(mp);
    if (e == KSGF_OK) {
        mp->coef_beta = beta;
        mp->coef_alpha = mp->alpha;
    }
    return e;
}

int ksgf_macproj_set_beta_faces(ksgf_macproj* mp, const double* beta_x, const double* beta_y, const double* beta_z,
                                int ngrow, ksgf_stream_t stream)
{
    if (!mp || !beta_x || !beta_y || !beta_z || ngrow < 0) {
        set_error("ksgf_macproj_set_beta_faces: bad argument");
        return KSGF_ERR_INVALID;
    }
    if (mp->alpha != 0.0) {
        set_error("ksgf_macproj_set_beta_faces: face coefficients are for the MAC projection (no diagonal term)");
        return KSGF_ERR_UNSUPPORTED;
    }
    if (!mp->vb) {
        int err = KSGF_OK;
        mp->vb = vb_create(mp->ba, mp->lobc, mp->hibc, mp->opt, &err);
        if (!mp->vb) return err;
    }
    const double* b3[3] = {beta_x, beta_y, beta_z};
    const int e = vb_set_faces(mp->vb, mp->ba, b3, ngrow, (hipStream_t)stream);
    if (e) return e;
    mp->use_vb = 1;
    return KSGF_OK;
}

int ksgf_macproj_project(ksgf_macproj* mp, double* u_mac, double* v_mac, double* w_mac, double rtol, double atol,
                         ksgf_mlmg_info* info, ksgf_stream_t stream)
{
    if (!mp || !u_mac || !v_mac || !w_mac) {
        set_error("ksgf_macproj_project: null argument");
        return KSGF_ERR_INVALID;
    }
    const ksgf_boxarray* ba = mp->ba;
    hipStream_t st = (hipStream_t)stream;
    CMFDev cu, cv, cw;
    MFDev mu[3];
    double* mac[3] = {u_mac, v_mac, w_mac};
    int e;
    if (mp->use_vb) return vb_project(mp->vb, ba, mac, 1, rtol, atol, info, st);
    if ((e = cmfdev(ba, u_mac, KSGF_XFACE, 1, 1, &cu)) || (e = cmfdev(ba, v_mac, KSGF_YFACE, 1, 1, &cv)) ||
        (e = cmfdev(ba, w_mac, KSGF_ZFACE, 1, 1, &cw)))
        return e;
    for (int d = 0; d < 3; ++d)
        if ((e = mfdev(ba, mac[d], KSGF_XFACE + d, 1, 1, &mu[d]))) return e;
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* rv = get_region(ba, z, z, &err);
    if (!rv) return err;
    const LevelGeom& L0 = mp->P;
    hipLaunchKernelGGL(k_mac_rhs, dim3(rv->nchunks), dim3(256), 0, st, cu, cv, cw, mp->rhs, L0,
                       regdev(rv, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1], ba->dlo[2], 1.0 / ba->dx[0],
                       1.0 / ba->dx[1], 1.0 / ba->dx[2]);
    KSGF_LAUNCH_CHECK();
    int ret = mg_solve_graphed(mp, rtol, atol, info, st);
    if (ret < 0) return ret;
    {
        const int glo[3] = {0, 0, 0}, ghi[3] = {1, 1, 1};
        const Region* rf = get_region(ba, glo, ghi, &err);
        if (!rf) return err;
        hipLaunchKernelGGL(k_mac_correct3, dim3(rf->nchunks), dim3(256), 0, st, mu[0], mu[1], mu[2],
                           (const double*)mp->phi, L0, regdev(rf, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1],
                           ba->dlo[2], mp->beta * (1.0 / ba->dx[0]), mp->beta * (1.0 / ba->dx[1]),
                           mp->beta * (1.0 / ba->dx[2]));
        KSGF_LAUNCH_CHECK();
    }
    return ret;
}

static int mac_rhs(ksgf_macproj* mp, const double* u_mac, const double* v_mac, const double* w_mac, hipStream_t st)
{
    const ksgf_boxarray* ba = mp->ba;
    CMFDev cu, cv, cw;
    int e;
    if ((e = cmfdev(ba, u_mac, KSGF_XFACE, 1, 1, &cu)) || (e = cmfdev(ba, v_mac, KSGF_YFACE, 1, 1, &cv)) ||
        (e = cmfdev(ba, w_mac, KSGF_ZFACE, 1, 1, &cw)))
        return e;
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* rv = get_region(ba, z, z, &err);
    if (!rv) return err;
    hipLaunchKernelGGL(k_mac_rhs, dim3(rv->nchunks), dim3(256), 0, st, cu, cv, cw, mp->rhs, mp->P,
                       regdev(rv, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1], ba->dlo[2], 1.0 / ba->dx[0],
                       1.0 / ba->dx[1], 1.0 / ba->dx[2]);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

static int mac_correct(ksgf_macproj* mp, double* u_mac, double* v_mac, double* w_mac, hipStream_t st)
{
    const ksgf_boxarray* ba = mp->ba;
    double* mac[3] = {u_mac, v_mac, w_mac};
    MFDev mu[3];
    int e, err = KSGF_OK;
    for (int d = 0; d < 3; ++d)
        if ((e = mfdev(ba, mac[d], KSGF_XFACE + d, 1, 1, &mu[d]))) return e;
    const LevelGeom& L0 = mp->P;
    {
        const int glo[3] = {0, 0, 0}, ghi[3] = {1, 1, 1};
        const Region* rf = get_region(ba, glo, ghi, &err);
        if (!rf) return err;
        hipLaunchKernelGGL(k_mac_correct3, dim3(rf->nchunks), dim3(256), 0, st, mu[0], mu[1], mu[2],
                           (const double*)mp->phi, L0, regdev(rf, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1],
                           ba->dlo[2], mp->beta * (1.0 / ba->dx[0]), mp->beta * (1.0 / ba->dx[1]),
                           mp->beta * (1.0 / ba->dx[2]));
        KSGF_LAUNCH_CHECK();
    }
    return KSGF_OK;
}

int ksgf_macproj_project_begin(ksgf_macproj* mp, const double* u_mac, const double* v_mac, const double* w_mac,
                               double rtol, double atol, int ncycles, ksgf_stream_t stream)
{
    if (!mp || !u_mac || !v_mac || !w_mac || ncycles < 0 || !(rtol >= 0.0) || !(atol >= 0.0)) {
        set_error("ksgf_macproj_project_begin: bad argument");
        return KSGF_ERR_INVALID;
    }
    if (mp->use_vb) {
        set_error("ksgf_macproj_project_begin: face coefficients solve with host-side tests only");
        return KSGF_ERR_UNSUPPORTED;
    }
    hipStream_t st = (hipStream_t)stream;
    int e = mac_rhs(mp, u_mac, v_mac, w_mac, st);
    if (e) return e;
    return mg_solve_begin(mp, rtol, atol, std::min(ncycles, mp->opt.maxiter), st);
}

int ksgf_macproj_project_finish(ksgf_macproj* mp, ksgf_mlmg_info* info, ksgf_stream_t stream)
{
    if (!mp) {
        set_error("ksgf_macproj_project_finish: null argument");
        return KSGF_ERR_INVALID;
    }
    if (mp->use_vb) {
        set_error("ksgf_macproj_project_finish: face coefficients solve with host-side tests only");
        return KSGF_ERR_UNSUPPORTED;
    }
    return mg_solve_finish(mp, info, (hipStream_t)stream);
}

int ksgf_macproj_project_correct(ksgf_macproj* mp, double* u_mac, double* v_mac, double* w_mac, ksgf_stream_t stream)
{
    if (!mp || !u_mac || !v_mac || !w_mac) {
        set_error("ksgf_macproj_project_correct: null argument");
        return KSGF_ERR_INVALID;
    }
    if (mp->use_vb) {
        set_error("ksgf_macproj_project_correct: face coefficients solve with host-side tests only");
        return KSGF_ERR_UNSUPPORTED;
    }
    hipStream_t st = (hipStream_t)stream;
    launch_ghosts<double>(mp->phi, mp->P, st);
    return mac_correct(mp, u_mac, v_mac, w_mac, st);
}

int ksgf_macproj_solve(ksgf_macproj* mp, const double* rhs, double* phi, double rtol, double atol, ksgf_mlmg_info* info,
                       ksgf_stream_t stream)
{
    if (!mp || !rhs || !phi) {
        set_error("ksgf_macproj_solve: null argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    if (mp->use_vb) return vb_solve(mp->vb, rhs, phi, rtol, atol, info, st);
    const LevelGeom& L0 = mp->P;
    long long n = (long long)L0.n0 * L0.n1 * L0.n2;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, st, rhs, mp->rhs, L0, 1);
    KSGF_LAUNCH_CHECK();
    int ret = mg_solve(mp, rtol, atol, info, st);
    if (ret < 0) return ret;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, st, (const double*)mp->phi, phi, L0, 0);
    KSGF_LAUNCH_CHECK();
    return ret;
}


// ---------------------------------------------------------------------------
// MLABecLaplacian with constant coefficients over the same multigrid
// ---------------------------------------------------------------------------
struct ksgf_abeclap {
    ksgf_macproj* mp = nullptr;
    double alpha = 1.0, a = 1.0, beta = 1.0, b = 1.0;
    int base_singular = 1;  // singular without the alpha term (no Dirichlet side)
    bool dirty = false;     // scalars changed since the hierarchy's coefficients were set
};

// the multigrid's coefficients from the current scalars (a solve needs
// beta b > 0 and alpha a >= 0; an apply takes any scalars)
static int abec_refill(ksgf_abeclap* ab)
{
    ksgf_macproj* mp = ab->mp;
    if (!(ab->beta * ab->b > 0.0) || !(ab->alpha * ab->a >= 0.0)) {
        set_error("ksgf_abeclap_solve: the solve needs beta b > 0 and alpha a >= 0");
        return KSGF_ERR_INVALID;
    }
    mp->alpha = ab->alpha * ab->a;
    mp->singular = (mp->alpha > 0.0) ? 0 : ab->base_singular;
    ab->dirty = false;
    return ksgf_macproj_set_beta(mp, ab->beta * ab->b);
}

int ksgf_abeclap_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc, double alpha, double a,
                        double beta, double b, const ksgf_mlmg_options* opts, ksgf_abeclap** out)
{
    if (!ba || !lobc || !hibc || !out || !(b > 0.0) || !(a >= 0.0) || !std::isfinite(alpha) || !std::isfinite(beta)) {
        set_error("ksgf_abeclap_create: bad argument (b > 0 and a >= 0 required)");
        return KSGF_ERR_INVALID;
    }
    *out = nullptr;
    ksgf_abeclap* ab = new ksgf_abeclap();
    ab->alpha = alpha;
    ab->a = a;
    ab->beta = beta;
    ab->b = b;
    // the hierarchy is built with a positive beta; the scalars are applied
    // before the first solve
    int e = ksgf_macproj_create(ba, lobc, hibc, (beta * b > 0.0) ? beta * b : b, opts, &ab->mp);
    if (e) {
        delete ab;
        return e;
    }
    ab->base_singular = ab->mp->singular;
    ab->dirty = true;
    *out = ab;
    return KSGF_OK;
}

int ksgf_abeclap_set_scalars(ksgf_abeclap* ab, double alpha, double beta)
{
    if (!ab || !std::isfinite(alpha) || !std::isfinite(beta)) {
        set_error("ksgf_abeclap_set_scalars: bad argument");
        return KSGF_ERR_INVALID;
    }
    ab->alpha = alpha;
    ab->beta = beta;
    ab->dirty = true;
    return KSGF_OK;
}

int ksgf_abeclap_solve(ksgf_abeclap* ab, double* sol, int ng_sol, const double* rhs, int ng_rhs, int ncomp,
                       int comp, double rtol, double atol, ksgf_mlmg_info* info, ksgf_stream_t stream)
{
    if (!ab || !sol || !rhs || ncomp <= 0 || comp < 0 || comp >= ncomp || ng_sol < 0 || ng_rhs < 0) {
        set_error("ksgf_abeclap_solve: bad argument");
        return KSGF_ERR_INVALID;
    }
    ksgf_macproj* mp = ab->mp;
    const ksgf_boxarray* ba = mp->ba;
    hipStream_t st = (hipStream_t)stream;
    CMFDev cr;
    MFDev ms;
    int e;
    if (ab->dirty && (e = abec_refill(ab))) return e;
    if ((e = cmfdev(ba, rhs, KSGF_CELL, ncomp, ng_rhs, &cr)) || (e = mfdev(ba, sol, KSGF_CELL, ncomp, ng_sol, &ms)))
        return e;
    int err = KSGF_OK;
    const int z[3] = {0, 0, 0};
    const Region* rv = get_region(ba, z, z, &err);
    if (!rv) return err;
    // the multigrid solves div(beta b grad phi) - alpha a phi = -rhs
    hipLaunchKernelGGL(k_mf_gather, dim3(rv->nchunks), dim3(256), 0, st, cr, comp, -1.0, mp->rhs, mp->P,
                       regdev(rv, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1], ba->dlo[2]);
    KSGF_LAUNCH_CHECK();
    int ret = mg_solve(mp, rtol, atol, info, st);
    if (ret < 0) return ret;
    hipLaunchKernelGGL(k_mf_scatter, dim3(rv->nchunks), dim3(256), 0, st, ms, comp, (const double*)mp->phi, mp->P,
                       regdev(rv, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1], ba->dlo[2]);
    KSGF_LAUNCH_CHECK();
    return ret;
}

int ksgf_abeclap_apply(ksgf_abeclap* ab, double* out, int ng_out, const double* in, int ng_in, int ncomp, int comp,
                       ksgf_stream_t stream)
{
    if (!ab || !out || !in || ncomp <= 0 || comp < 0 || comp >= ncomp || ng_out < 0 || ng_in < 0) {
        set_error("ksgf_abeclap_apply: bad argument");
        return KSGF_ERR_INVALID;
    }
    ksgf_macproj* mp = ab->mp;
    const ksgf_boxarray* ba = mp->ba;
    hipStream_t st = (hipStream_t)stream;
    CMFDev ci;
    MFDev mo;
    int e;
    if ((e = cmfdev(ba, in, KSGF_CELL, ncomp, ng_in, &ci)) || (e = mfdev(ba, out, KSGF_CELL, ncomp, ng_out, &mo)))
        return e;
    int err = KSGF_OK;
    const int z[3] = {0, 0, 0};
    const Region* rv = get_region(ba, z, z, &err);
    if (!rv) return err;
    // x into the level-0 array (phi is scratch between solves), BC ghosts, and
    // the stencil with the current scalars (of either sign)
    LevelGeom A = mp->P;
    const double h0[3] = {ba->dx[0], ba->dx[1], ba->dx[2]};
    fill_coefs(A, h0, ab->beta * ab->b, ab->alpha * ab->a);
    hipLaunchKernelGGL(k_mf_gather, dim3(rv->nchunks), dim3(256), 0, st, ci, comp, 1.0, mp->phi, mp->P,
                       regdev(rv, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1], ba->dlo[2]);
    launch_ghosts<double>(mp->phi, mp->P, st);
    hipLaunchKernelGGL(k_abec_apply, dim3(rv->nchunks), dim3(256), 0, st, (const double*)mp->phi, mo, comp, A,
                       regdev(rv, ba->nbox), ba->dbox, ba->dlo[0], ba->dlo[1], ba->dlo[2]);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_abeclap_destroy(ksgf_abeclap* ab)
{
    if (!ab) return KSGF_OK;
    ksgf_macproj_destroy(ab->mp);
    delete ab;
    return KSGF_OK;
}

int ksgf_macproj_get_phi(ksgf_macproj* mp, double* phi, ksgf_stream_t stream)
{
    if (!mp || !phi) {
        set_error("ksgf_macproj_get_phi: null argument");
        return KSGF_ERR_INVALID;
    }
    if (mp->use_vb) return vb_get_phi(mp->vb, phi, (hipStream_t)stream);
    const LevelGeom& L0 = mp->P;
    long long n = (long long)L0.n0 * L0.n1 * L0.n2;
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, (const double*)mp->phi, phi,
                       L0, 0);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

}  // extern "C"

// ===========================================================================
// Many independent Poisson problems solved together.  Each problem has its
// own hierarchy (a projector over a one-box domain, autotuned as usual); the
// batch runs every multigrid stage of every problem as one launch over a
// ragged table of (problem, tile) blocks, the coarse part as one workgroup
// per problem, and the FP64 outer residual of all problems in one launch.
// Per-problem state (active flag, current phi buffer, iteration count,
// norm, target) lives on the device, so the launch sequence of one
// iteration never changes: it is captured once into a hipGraph and replayed
// until no problem is active.
// ===========================================================================
namespace ksgf {

template <typename T>
struct BLevel {  // one problem at one grid level
    PkArgs<T> pk;  // x = cor[l], b = res[l], xc = cor[l+1]
    RowTiles th;   // half-sweep tiles over packed positions
    RowTiles tr;   // restriction tiles over the coarse level
    T* cres;       // res[l+1]
    T w;
    int r0, r1, r2;
    int prob;
};

template <typename T>
struct BTop {  // one problem's FP64 outer residual (level 0)
    LevelGeom P, L0;
    double* phi[2];
    const double* rhs;
    const T* cor;
    T* r;
    RowTiles t;
    int prob;
};

template <typename T>
struct BCoarse {
    const LevelPtrs<T>* lp;
    BottomWork w;
    int lc, nlev, singular, prob;
};

struct BState {
    int* active;     // 1 while the problem iterates
    int* cur;        // phi buffer holding the current iterate
    int* iters;
    double* norm;    // last |r|_inf
    double* target;  // max(atol, rtol |rhs|_inf)
    double* bound;   // kDivergeFactor |rhs|_inf
    double* part;    // per-block maxima
    const int* part0;  // first partial slot of each problem
    int maxiter;
};

// logical block b of a batched launch: table entry and tile within it
struct BlkMap {
    int e, t;
};

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_pk_half_batch(const BLevel<T>* __restrict__ lv, const BlkMap* __restrict__ bm,
                                                       int color, BState S)
{
    const BlkMap m = bm[xcd_remap((int)blockIdx.x, (int)gridDim.x)];
    const BLevel<T> L = lv[m.e];
    if (!S.active[L.prob]) return;
    int h, j, k0, k1;
    row_tile(L.th, m.t, (int)threadIdx.x, h, j, k0, k1, L.pk.L.n2);
    if (h >= L.pk.L.prow || j >= L.pk.L.n1) return;
    if (MODE == kPkProlong && L.pk.C.packed) {
        for (int k = k0; k < k1; ++k) pk_relax<T, MODE, true>(L.pk, color, h, j, k);
    } else {
        for (int k = k0; k < k1; ++k) pk_relax<T, MODE, false>(L.pk, color, h, j, k);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_restrict_batch(const BLevel<T>* __restrict__ lv, const BlkMap* __restrict__ bm,
                                                        BState S)
{
    const BlkMap m = bm[xcd_remap((int)blockIdx.x, (int)gridDim.x)];
    const BLevel<T> L = lv[m.e];
    if (!S.active[L.prob]) return;
    const LevelGeom& C = L.pk.C;
    int I, J, K0, K1;
    row_tile(L.tr, m.t, (int)threadIdx.x, I, J, K0, K1, C.n2);
    if (I >= C.n0 || J >= C.n1) return;
    for (int K = K0; K < K1; ++K)
        L.cres[AnyIdx(C, I, J, K)] =
            pk_restrict_cell<T>(L.pk.x, L.pk.b, L.pk.L, I, J, K, L.r0, L.r1, L.r2, L.w);
}

template <typename T>
__global__ void __launch_bounds__(kCoarseThreads) k_coarse_batch(const BCoarse<T>* __restrict__ cb, BState S,
                                                                 double rtol, double atol, int maxiter, int nub,
                                                                 int nuf, int nu1, int nu2)
{
    __shared__ double sh[kCoarseThreads / 64];
    __shared__ __align__(16) char pool[kLdsPoolBytes];
    const BCoarse<T>& c = cb[blockIdx.x];
    if (!S.active[c.prob]) return;
    coarse_vcycle_body<T>(c.lp, c.lc, c.nlev, c.w, c.singular, rtol, atol, maxiter, nub, nuf, nu1, nu2, sh, pool);
}

// FP64 outer residual of every problem; with_cor: phi(next) = phi + cor
template <typename T>
__global__ void __launch_bounds__(256) k_update_residual_batch(const BTop<T>* __restrict__ tp,
                                                               const BlkMap* __restrict__ bm, int with_cor, BState S)
{
    const BlkMap mb = bm[xcd_remap((int)blockIdx.x, (int)gridDim.x)];
    const BTop<T>& t = tp[mb.e];
    if (!S.active[t.prob]) return;
    const bool c = S.cur[t.prob] != 0;
    double* const p0 = t.phi[0];
    double* const p1 = t.phi[1];
    const double m = update_residual_tile<T, T, true>(c ? p1 : p0, with_cor ? t.cor : (const T*)nullptr, c ? p0 : p1,
                                                      t.rhs, t.r, t.P, t.L0, t.t, mb.t);
    __shared__ double sh[4];
    double v = m;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) S.part[S.part0[t.prob] + mb.t] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// one workgroup per problem: the new norm, the phi buffer flip, the
// iteration count and whether the problem keeps iterating.  first: the
// norm of the right-hand side (sets the target, no flip, no count).
__global__ void __launch_bounds__(256) k_batch_finalize(const int* __restrict__ nparts, int first, double rtol,
                                                        double atol, BState S)
{
    const int p = blockIdx.x;
    if (!S.active[p]) return;
    double m = 0.0;
    for (int q = (int)threadIdx.x; q < nparts[p]; q += (int)blockDim.x) m = fmax(m, S.part[S.part0[p] + q]);
    __shared__ double sh[4];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
        S.norm[p] = m;
        if (first) {
            S.target[p] = fmax(atol, rtol * m);
            S.bound[p] = kDivergeFactor * m;
            S.iters[p] = 0;
            S.active[p] = (m > S.target[p] && S.maxiter > 0) ? 1 : 0;
        } else {
            S.cur[p] ^= 1;
            S.iters[p] += 1;
            S.active[p] = (m > S.target[p] && m <= S.bound[p] && S.iters[p] < S.maxiter) ? 1 : 0;
        }
    }
}

}  // namespace ksgf

struct ksgf_poisson_batch {
    int np = 0;
    bool mixed = false;
    std::vector<ksgf_boxarray*> bas;
    std::vector<ksgf_macproj*> mps;
    int lmax = 0;  // grid levels of the deepest problem
    struct Tab {
        void* lv = nullptr;  // BLevel<T>[ne]
        void* bm_h = nullptr;  // BlkMap per half-sweep block
        void* bm_r = nullptr;  // BlkMap per restriction block
        int ne = 0, nb_h = 0, nb_r = 0;
    };
    std::vector<Tab> lv;
    void* top = nullptr;
    void* top_bm = nullptr;
    int top_nb = 0;
    void* coarse = nullptr;
    int ncoarse = 0;
    int* nparts = nullptr;
    int* part0 = nullptr;
    double* part = nullptr;
    int *active = nullptr, *cur = nullptr, *iters = nullptr;
    double *norm = nullptr, *target = nullptr, *bound = nullptr;
    int* h_state = nullptr;     // pinned: active, cur, iters (3*np)
    double* h_norm = nullptr;   // pinned: norm, target, bound (3*np)
    hipStream_t cap = nullptr;  // capture stream
    hipGraphExec_t exec = nullptr;
    ksgf_mlmg_options opt{};
};

namespace ksgf {

static BState batch_state(const ksgf_poisson_batch* pb)
{
    BState S;
    S.active = pb->active;
    S.cur = pb->cur;
    S.iters = pb->iters;
    S.norm = pb->norm;
    S.target = pb->target;
    S.bound = pb->bound;
    S.part = pb->part;
    S.part0 = pb->part0;
    S.maxiter = pb->opt.maxiter;
    return S;
}

template <typename T>
static std::vector<T*>& cor_of(ksgf_macproj* mp);
template <>
std::vector<float*>& cor_of<float>(ksgf_macproj* mp)
{
    return mp->cor32;
}
template <>
std::vector<double*>& cor_of<double>(ksgf_macproj* mp)
{
    return mp->cor;
}
template <typename T>
static std::vector<T*>& res_of(ksgf_macproj* mp);
template <>
std::vector<float*>& res_of<float>(ksgf_macproj* mp)
{
    return mp->res32;
}
template <>
std::vector<double*>& res_of<double>(ksgf_macproj* mp)
{
    return mp->res;
}
template <typename T>
static const LevelPtrs<T>* dlp_of(ksgf_macproj* mp);
template <>
const LevelPtrs<float>* dlp_of<float>(ksgf_macproj* mp)
{
    return mp->dlp32;
}
template <>
const LevelPtrs<double>* dlp_of<double>(ksgf_macproj* mp)
{
    return mp->dlp64;
}

template <typename X>
static int upload(void** d, const std::vector<X>& v)
{
    if (v.empty()) return KSGF_OK;
    if (hipMalloc(d, sizeof(X) * v.size()) != hipSuccess) {
        set_error("poisson batch: hipMalloc");
        return KSGF_ERR_NOMEM;
    }
    KSGF_HIP_CHECK(hipMemcpy(*d, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice));
    return KSGF_OK;
}

// device tables of every stage of the batch
template <typename T>
static int batch_tables(ksgf_poisson_batch* pb)
{
    int e;
    pb->lmax = 0;
    for (auto* mp : pb->mps) pb->lmax = std::max(pb->lmax, mp->lc);
    pb->lv.assign(pb->lmax, {});
    for (int l = 0; l < pb->lmax; ++l) {
        std::vector<BLevel<T>> t;
        std::vector<BlkMap> mh, mr;
        for (int p = 0; p < pb->np; ++p) {
            ksgf_macproj* mp = pb->mps[p];
            if (l >= mp->lc) continue;
            BLevel<T> b{};
            b.pk = pk_args<T>(mp, l, cor_of<T>(mp)[l], res_of<T>(mp)[l], cor_of<T>(mp)[l + 1]);
            b.th = row_tiles(b.pk.L.prow, b.pk.L.n1, b.pk.L.n2, mp->kz_sweep[l]);
            const LevelGeom& C = mp->G[l + 1];
            b.tr = row_tiles(C.n0, C.n1, C.n2, kRestrictKz);
            if (b.tr.ntx * b.tr.nty * b.tr.nkc < 2048) b.tr = row_tiles(C.n0, C.n1, C.n2, 1);
            b.cres = res_of<T>(mp)[l + 1];
            const int* r = &mp->ratio[3 * l];
            b.r0 = r[0];
            b.r1 = r[1];
            b.r2 = r[2];
            b.w = (T)(1.0 / (double)(r[0] * r[1] * r[2]));
            b.prob = p;
            t.push_back(b);
            for (int q = 0; q < b.th.ntx * b.th.nty * b.th.nkc; ++q) mh.push_back({(int)t.size() - 1, q});
            for (int q = 0; q < b.tr.ntx * b.tr.nty * b.tr.nkc; ++q) mr.push_back({(int)t.size() - 1, q});
        }
        auto& tab = pb->lv[l];
        tab.ne = (int)t.size();
        tab.nb_h = (int)mh.size();
        tab.nb_r = (int)mr.size();
        if ((e = upload(&tab.lv, t)) || (e = upload(&tab.bm_h, mh)) || (e = upload(&tab.bm_r, mr)))
            return e;
    }
    std::vector<BTop<T>> top;
    std::vector<BlkMap> tbm;
    std::vector<int> nparts, part0;
    for (int p = 0; p < pb->np; ++p) {
        ksgf_macproj* mp = pb->mps[p];
        BTop<T> b{};
        b.P = mp->P;
        b.L0 = mp->G[0];
        b.phi[0] = mp->phi;
        b.phi[1] = mp->phi2;
        b.rhs = mp->rhs;
        b.cor = cor_of<T>(mp)[0];
        b.r = res_of<T>(mp)[0];
        b.t = row_tiles(mp->P.n0, mp->P.n1, mp->P.n2, mp->kz_res);
        b.prob = p;
        const int nb = b.t.ntx * b.t.nty * b.t.nkc;
        part0.push_back((int)tbm.size());
        nparts.push_back(nb);
        for (int q = 0; q < nb; ++q) tbm.push_back({p, q});
        top.push_back(b);
    }
    pb->top_nb = (int)tbm.size();
    if ((e = upload(&pb->top, top)) || (e = upload(&pb->top_bm, tbm)) ||
        (e = upload((void**)&pb->nparts, nparts)) || (e = upload((void**)&pb->part0, part0)))
        return e;
    if (hipMalloc(&pb->part, sizeof(double) * pb->top_nb) != hipSuccess) {
        set_error("poisson batch: hipMalloc");
        return KSGF_ERR_NOMEM;
    }
    std::vector<BCoarse<T>> cb;
    for (int p = 0; p < pb->np; ++p) {
        ksgf_macproj* mp = pb->mps[p];
        BCoarse<T> c{};
        c.lp = dlp_of<T>(mp);
        c.w = mp->bw;
        c.lc = mp->lc;
        c.nlev = mp->nlev;
        c.singular = mp->singular;
        c.prob = p;
        cb.push_back(c);
    }
    pb->ncoarse = (int)cb.size();
    return upload(&pb->coarse, cb);
}

// one iteration of every active problem: V-cycle, FP64 residual, finalize
template <typename T>
static void batch_iteration(ksgf_poisson_batch* pb, hipStream_t st)
{
    const BState S = batch_state(pb);
    const ksgf_mlmg_options& o = pb->opt;
    auto half = [&](int l, int mode) {
        const auto& t = pb->lv[l];
        for (int color = 0; color < 2; ++color) {
            const BLevel<T>* lv = (const BLevel<T>*)t.lv;
            const BlkMap* bm = (const BlkMap*)t.bm_h;
            if (mode == kPkZero)
                hipLaunchKernelGGL((k_pk_half_batch<T, kPkZero>), dim3(t.nb_h), dim3(256), 0, st, lv, bm, color, S);
            else if (mode == kPkProlong)
                hipLaunchKernelGGL((k_pk_half_batch<T, kPkProlong>), dim3(t.nb_h), dim3(256), 0, st, lv, bm, color, S);
            else
                hipLaunchKernelGGL((k_pk_half_batch<T, kPkPlain>), dim3(t.nb_h), dim3(256), 0, st, lv, bm, color, S);
        }
    };
    for (int l = 0; l < pb->lmax; ++l) {
        for (int s = 0; s < o.num_pre_smooth; ++s) half(l, s == 0 ? kPkZero : kPkPlain);
        const auto& t = pb->lv[l];
        hipLaunchKernelGGL(k_restrict_batch<T>, dim3(t.nb_r), dim3(256), 0, st, (const BLevel<T>*)t.lv,
                           (const BlkMap*)t.bm_r, S);
    }
    hipLaunchKernelGGL(k_coarse_batch<T>, dim3(pb->ncoarse), dim3(kCoarseThreads), 0, st,
                       (const BCoarse<T>*)pb->coarse, S, o.bottom_rtol, o.bottom_atol, o.bottom_maxiter,
                       o.num_bottom_smooth, o.num_final_smooth, o.num_pre_smooth, o.num_post_smooth);
    for (int l = pb->lmax - 1; l >= 0; --l)
        for (int s = 0; s < o.num_post_smooth; ++s) half(l, s == 0 ? kPkProlong : kPkPlain);
    hipLaunchKernelGGL(k_update_residual_batch<T>, dim3(pb->top_nb), dim3(256), 0, st, (const BTop<T>*)pb->top,
                       (const BlkMap*)pb->top_bm, 1, S);
    hipLaunchKernelGGL(k_batch_finalize, dim3(pb->np), dim3(256), 0, st, (const int*)pb->nparts, 0, 0.0, 0.0, S);
}

template <typename T>
static int batch_solve(ksgf_poisson_batch* pb, const double* const* rhs, double* const* phi, double rtol, double atol,
                       int use_graph, ksgf_mlmg_info* infos, hipStream_t st)
{
    const BState S = batch_state(pb);
    // set up: pack, zero phi, subtract means, rhs norms and first residuals
    std::vector<int> ones(pb->np, 1), zeros(pb->np, 0);
    KSGF_HIP_CHECK(hipMemcpyAsync(pb->active, ones.data(), sizeof(int) * pb->np, hipMemcpyHostToDevice, st));
    KSGF_HIP_CHECK(hipMemcpyAsync(pb->cur, zeros.data(), sizeof(int) * pb->np, hipMemcpyHostToDevice, st));
    for (int p = 0; p < pb->np; ++p) {
        ksgf_macproj* mp = pb->mps[p];
        const LevelGeom& P = mp->P;
        const long long n = (long long)P.n0 * P.n1 * P.n2;
        hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, st, rhs[p], mp->rhs, P, 1);
        launch_fill<double>(mp->phi, P.tot, 0.0, st);
        if (mp->singular) launch_sub_mean(mp->rhs, P, mp->part, st);
        KSGF_HIP_CHECK(hipMemsetAsync(mp->bw.stats, 0, 2 * sizeof(int), st));
    }
    hipLaunchKernelGGL(k_update_residual_batch<T>, dim3(pb->top_nb), dim3(256), 0, st, (const BTop<T>*)pb->top,
                       (const BlkMap*)pb->top_bm, 0, S);
    hipLaunchKernelGGL(k_batch_finalize, dim3(pb->np), dim3(256), 0, st, (const int*)pb->nparts, 1, rtol, atol, S);
    KSGF_LAUNCH_CHECK();
    auto read_state = [&]() -> int {
        KSGF_HIP_CHECK(hipMemcpyAsync(pb->h_state, pb->active, sizeof(int) * pb->np, hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(
            hipMemcpyAsync(pb->h_state + pb->np, pb->cur, sizeof(int) * pb->np, hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(
            hipMemcpyAsync(pb->h_state + 2 * pb->np, pb->iters, sizeof(int) * pb->np, hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(hipMemcpyAsync(pb->h_norm, pb->norm, sizeof(double) * pb->np, hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(
            hipMemcpyAsync(pb->h_norm + pb->np, pb->target, sizeof(double) * pb->np, hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(hipMemcpyAsync(pb->h_norm + 2 * pb->np, pb->bound, sizeof(double) * pb->np,
                                      hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(hipStreamSynchronize(st));
        return KSGF_OK;
    };
    int e = read_state();
    if (e) return e;
    std::vector<double> r0(pb->h_norm, pb->h_norm + pb->np);
    if (use_graph && !pb->exec) {
        // capture one iteration on the batch's own stream
        KSGF_HIP_CHECK(hipStreamBeginCapture(pb->cap, hipStreamCaptureModeThreadLocal));
        batch_iteration<T>(pb, pb->cap);
        hipGraph_t g;
        KSGF_HIP_CHECK(hipStreamEndCapture(pb->cap, &g));
        const hipError_t ie = hipGraphInstantiate(&pb->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            pb->exec = nullptr;
            set_error(std::string("poisson batch: hipGraphInstantiate: ") + hipGetErrorString(ie));
            return KSGF_ERR_HIP;
        }
    }
    for (int it = 0; it < pb->opt.maxiter; ++it) {
        bool any = false;
        for (int p = 0; p < pb->np; ++p) any = any || pb->h_state[p];
        if (!any) break;
        if (use_graph)
            KSGF_HIP_CHECK(hipGraphLaunch(pb->exec, st));
        else
            batch_iteration<T>(pb, st);
        KSGF_LAUNCH_CHECK();
        if ((e = read_state())) return e;
    }
    int ret = KSGF_OK;
    std::string diverged;
    for (int p = 0; p < pb->np; ++p) {
        ksgf_macproj* mp = pb->mps[p];
        const LevelGeom& P = mp->P;
        const long long n = (long long)P.n0 * P.n1 * P.n2;
        const double* cur = pb->h_state[pb->np + p] ? mp->phi2 : mp->phi;
        hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, st, cur, phi[p], P, 0);
        KSGF_HIP_CHECK(hipMemcpyAsync(mp->hstats, mp->bw.stats, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(hipStreamSynchronize(st));
        ksgf_mlmg_info inf{};
        inf.iters = pb->h_state[2 * pb->np + p];
        inf.r0 = r0[p];
        inf.rfinal = pb->h_norm[p];
        inf.converged = pb->h_norm[p] <= pb->h_norm[pb->np + p];
        inf.nlevels = mp->nlev;
        inf.outer_iters = pb->mixed ? inf.iters : 0;
        inf.bottom_iters = mp->hstats[0];
        inf.bottom_fail = mp->hstats[1];
        if (infos) infos[p] = inf;
        if (!inf.converged) {
            if (!(inf.rfinal <= pb->h_norm[2 * pb->np + p]))
                diverged += (diverged.empty() ? "" : ", ") + std::to_string(p);
            else if (ret == KSGF_OK)
                ret = KSGF_ERR_NOT_CONVERGED;
        }
    }
    if (!diverged.empty()) {
        set_error("poisson batch: the residual of problem(s) " + diverged + " exceeds 1e20 |rhs| (diverging)");
        return KSGF_ERR_DIVERGED;
    }
    return ret;
}

}  // namespace ksgf

extern "C" {

int ksgf_poisson_batch_destroy(ksgf_poisson_batch* pb)
{
    if (!pb) return KSGF_OK;
    if (pb->exec) (void)hipGraphExecDestroy(pb->exec);
    if (pb->cap) (void)hipStreamDestroy(pb->cap);
    for (auto& t : pb->lv) {
        if (t.lv) (void)hipFree(t.lv);
        if (t.bm_h) (void)hipFree(t.bm_h);
        if (t.bm_r) (void)hipFree(t.bm_r);
    }
    void* ps[] = {pb->top, pb->top_bm, pb->coarse, pb->nparts, pb->part0, pb->part,
                  pb->active, pb->cur, pb->iters, pb->norm, pb->target, pb->bound};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (pb->h_state) (void)hipHostFree(pb->h_state);
    if (pb->h_norm) (void)hipHostFree(pb->h_norm);
    for (auto* mp : pb->mps) ksgf_macproj_destroy(mp);
    for (auto* ba : pb->bas) ksgf_boxarray_destroy(ba);
    delete pb;
    return KSGF_OK;
}

int ksgf_poisson_batch_create(int nprob, const int* n, const double* dx, const int* lobc, const int* hibc,
                              const double* beta, const ksgf_mlmg_options* opts, ksgf_poisson_batch** out)
{
    if (nprob <= 0 || !n || !dx || !lobc || !hibc || !beta || !out) {
        set_error("ksgf_poisson_batch_create: bad argument");
        return KSGF_ERR_INVALID;
    }
    auto* pb = new ksgf_poisson_batch();
    pb->np = nprob;
    if (opts)
        pb->opt = *opts;
    else
        ksgf_mlmg_default_options(&pb->opt);
    pb->mixed = (pb->opt.precision == KSGF_MG_MIXED);
    if (pb->opt.num_pre_smooth < 1 || pb->opt.num_post_smooth < 1) {
        delete pb;
        set_error("ksgf_poisson_batch_create: the batch needs at least one pre- and one post-smoothing sweep");
        return KSGF_ERR_UNSUPPORTED;
    }
    int e;
    for (int p = 0; p < nprob; ++p) {
        ksgf_geometry g{};
        int box[6];
        for (int d = 0; d < 3; ++d) {
            g.domain_lo[d] = 0;
            g.domain_hi[d] = n[3 * p + d] - 1;
            g.prob_lo[d] = 0.0;
            g.prob_hi[d] = dx[3 * p + d] * n[3 * p + d];
            g.is_periodic[d] = (lobc[3 * p + d] == KSGF_LO_PERIODIC);
            box[d] = 0;
            box[3 + d] = n[3 * p + d] - 1;
        }
        ksgf_boxarray* ba = nullptr;
        if ((e = ksgf_boxarray_create(&g, box, 1, &ba))) {
            ksgf_poisson_batch_destroy(pb);
            return e;
        }
        pb->bas.push_back(ba);
        ksgf_macproj* mp = nullptr;
        if ((e = ksgf_macproj_create(ba, lobc + 3 * p, hibc + 3 * p, beta[p], &pb->opt, &mp))) {
            ksgf_poisson_batch_destroy(pb);
            return e;
        }
        pb->mps.push_back(mp);
        for (int l = 0; l < mp->lc; ++l)
            if (!mp->G[l].packed || mp->P.ghost_mode) {
                ksgf_poisson_batch_destroy(pb);
                set_error("ksgf_poisson_batch_create: problem " + std::to_string(p) +
                          " has a grid level with an odd x extent or odd periodic extent; solve it on its own");
                return KSGF_ERR_UNSUPPORTED;
            }
    }
    if ((e = pb->mixed ? batch_tables<float>(pb) : batch_tables<double>(pb))) {
        ksgf_poisson_batch_destroy(pb);
        return e;
    }
    const size_t np = (size_t)nprob;
    if (hipMalloc(&pb->active, sizeof(int) * np) != hipSuccess || hipMalloc(&pb->cur, sizeof(int) * np) != hipSuccess ||
        hipMalloc(&pb->iters, sizeof(int) * np) != hipSuccess || hipMalloc(&pb->norm, sizeof(double) * np) != hipSuccess ||
        hipMalloc(&pb->target, sizeof(double) * np) != hipSuccess ||
        hipMalloc(&pb->bound, sizeof(double) * np) != hipSuccess ||
        hipHostMalloc(&pb->h_state, sizeof(int) * 3 * np, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&pb->h_norm, sizeof(double) * 3 * np, hipHostMallocDefault) != hipSuccess ||
        hipStreamCreateWithFlags(&pb->cap, hipStreamNonBlocking) != hipSuccess) {
        ksgf_poisson_batch_destroy(pb);
        set_error("ksgf_poisson_batch_create: allocation");
        return KSGF_ERR_NOMEM;
    }
    *out = pb;
    return KSGF_OK;
}

int ksgf_poisson_batch_solve(ksgf_poisson_batch* pb, const double* const* rhs, double* const* phi, double rtol,
                             double atol, int use_graph, ksgf_mlmg_info* infos, ksgf_stream_t stream)
{
    if (!pb || !rhs || !phi) {
        set_error("ksgf_poisson_batch_solve: null argument");
        return KSGF_ERR_INVALID;
    }
    for (int p = 0; p < pb->np; ++p)
        if (!rhs[p] || !phi[p]) {
            set_error("ksgf_poisson_batch_solve: null rhs/phi pointer");
            return KSGF_ERR_INVALID;
        }
    hipStream_t st = (hipStream_t)stream;
    return pb->mixed ? batch_solve<float>(pb, rhs, phi, rtol, atol, use_graph, infos, st)
                     : batch_solve<double>(pb, rhs, phi, rtol, atol, use_graph, infos, st);
}

}  // extern "C"
