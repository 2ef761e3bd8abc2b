// Nodal projection with a constant sigma:
//   incflo::ApplyProjection -> Hydro::NodalProjector(vel, sigma, geom, lpinfo)
//   + setDomainBC + MLMGOptions + project(rtol, atol) + getPhi / getGradPhi
//   (src/projection/incflo_apply_nodal_projection.cpp:149-465,
//    src/core/MLMGOptions.cpp:118-130).
//
// Unknowns are the nodes of the cell grid, one domain-wide array per
// multigrid level (Fortran order, no ghosts): N_d = n_d nodes in a periodic
// direction (node n_d is node 0), n_d + 1 otherwise.  Periodic wrap, the
// Neumann mirror (node -1 reads node 1) and Dirichlet nodes (held at zero)
// are resolved in the stencil index arithmetic, so no ghost fills run
// between sweeps.
//
// Operator: the trilinear (Q1) finite-element stiffness matrix over the cell
// volume with the sign of div(sigma grad) -- a 27-point stencil whose eight
// weights depend only on sigma and the cell sizes.  A Neumann boundary row is
// the full stencil with mirrored neighbours (twice the finite-element row)
// and its right-hand side is doubled to match.
//
// V-cycle (MLMG defaults nu1 = nu2 = 2): each smoothing step is four weighted
// Jacobi sweeps with omega = 1/lambda_max(D^-1 A) (2/3 on a cubic grid); the
// residual is restricted by full weighting (the transpose of trilinear
// interpolation over the coarse/fine volume ratio) and the coarse correction
// comes back by trilinear interpolation.  For nested Q1 spaces these are the
// Galerkin transfers, so the rediscretised coarse stencil is the Galerkin
// coarse operator.  The coarsest level is solved by unpreconditioned
// BiCGStab in one workgroup.
//
// The stencil, transfer and correction expressions are evaluated in the
// operand order of the CPU oracle (oracle/ksgf_oracle.c, nodal section) and
// built without contraction; reductions (norm, weighted mean, BiCGStab dot
// products) are blocked, so iterates agree with the oracle to rounding.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ksgf_internal.h"

namespace ksgf {

constexpr int kNdSweeps = 4;       // Jacobi sweeps per smoothing step
constexpr int kNdBottomThreads = 1024;
constexpr int kNdMaxLev = 40;

struct NLev {
    int N[3];
    int per[3];
    int dlo[3], dhi[3];  // Dirichlet sides
    int nlo[3], nhi[3];  // Neumann sides
    int r[3];            // coarsening ratio to the next level
    double w[8];         // centre, faces x/y/z, edges xy/xz/yz, corners
    double wd;           // omega / centre
    int tot;
};

// neighbour node along one direction: periodic wrap, or the mirror at a
// non-periodic side (only a boundary node has a neighbour outside)
__device__ __forceinline__ int nd_nb(int i, int s, int N, int per)
{
    int m = i + s;
    if (per) return m < 0 ? m + N : (m >= N ? m - N : m);
    return m < 0 ? 1 : (m >= N ? N - 2 : m);
}

__device__ __forceinline__ bool nd_dirichlet(const NLev& L, int i, int j, int k)
{
    return (L.dlo[0] && i == 0) || (L.dhi[0] && i == L.N[0] - 1) || (L.dlo[1] && j == 0) ||
           (L.dhi[1] && j == L.N[1] - 1) || (L.dlo[2] && k == 0) || (L.dhi[2] && k == L.N[2] - 1);
}

__device__ __forceinline__ int nd_idx(const NLev& L, int i, int j, int k) { return i + L.N[0] * (j + L.N[1] * k); }

// (L x) at node (i,j,k), same grouping as the oracle
__device__ __forceinline__ double nd_apply_at(const NLev& L, const double* __restrict__ x, int i, int j, int k)
{
    const int im = nd_nb(i, -1, L.N[0], L.per[0]), ip = nd_nb(i, 1, L.N[0], L.per[0]);
    const int jm = nd_nb(j, -1, L.N[1], L.per[1]), jp = nd_nb(j, 1, L.N[1], L.per[1]);
    const int km = nd_nb(k, -1, L.N[2], L.per[2]), kp = nd_nb(k, 1, L.N[2], L.per[2]);
    const int sy = L.N[0], sz = L.N[0] * L.N[1];
    const int y0 = j * sy, ym = jm * sy, yp = jp * sy;
    const int z0 = k * sz, zm = km * sz, zp = kp * sz;
#define X(a, b, c) x[(a) + (b) + (c)]
    const double fx = X(im, y0, z0) + X(ip, y0, z0);
    const double fy = X(i, ym, z0) + X(i, yp, z0);
    const double fz = X(i, y0, zm) + X(i, y0, zp);
    const double exy = (X(im, ym, z0) + X(ip, ym, z0)) + (X(im, yp, z0) + X(ip, yp, z0));
    const double exz = (X(im, y0, zm) + X(ip, y0, zm)) + (X(im, y0, zp) + X(ip, y0, zp));
    const double eyz = (X(i, ym, zm) + X(i, yp, zm)) + (X(i, ym, zp) + X(i, yp, zp));
    const double cr = ((X(im, ym, zm) + X(ip, ym, zm)) + (X(im, yp, zm) + X(ip, yp, zm))) +
                      ((X(im, ym, zp) + X(ip, ym, zp)) + (X(im, yp, zp) + X(ip, yp, zp)));
    const double c0 = X(i, y0, z0);
#undef X
    return (((L.w[0] * c0 + L.w[1] * fx) + (L.w[2] * fy + L.w[3] * fz)) +
            ((L.w[4] * exy + L.w[5] * exz) + (L.w[6] * eyz + L.w[7] * cr)));
}

__device__ __forceinline__ double nd_abs_or_inf(double v)
{
    const double a = fabs(v);
    return a == a ? a : __longlong_as_double(0x7ff0000000000000LL);
}

// block max of non-negative values into one word (bit order = value order)
__device__ __forceinline__ void nd_block_max(double m, unsigned long long* out)
{
    __shared__ double sh[4];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    if ((t & 63) == 0) sh[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        double q = sh[0];
        const int nw = (int)(blockDim.x * blockDim.y) >> 6;
        for (int w = 1; w < nw; ++w) q = fmax(q, sh[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(q));
    }
}

// launch shape of the level kernels: 64 x 4 nodes per workgroup, one plane
__device__ __forceinline__ bool nd_node(const NLev& L, int& i, int& j, int& k)
{
    i = blockIdx.x * 64 + threadIdx.x;
    j = blockIdx.y * 4 + threadIdx.y;
    k = blockIdx.z;
    return i < L.N[0] && j < L.N[1];
}

// one Jacobi sweep: xo = xi + (b - L xi) * wd; zero guess: xo = b * wd
template <bool ZERO>
__global__ void __launch_bounds__(256) k_nd_jacobi(double* __restrict__ xo, const double* __restrict__ xi,
                                                   const double* __restrict__ b, NLev L)
{
    int i, j, k;
    if (!nd_node(L, i, j, k)) return;
    const int c = nd_idx(L, i, j, k);
    if (nd_dirichlet(L, i, j, k)) {
        xo[c] = 0.0;
        return;
    }
    if (ZERO)
        xo[c] = b[c] * L.wd;
    else
        xo[c] = xi[c] + (b[c] - nd_apply_at(L, xi, i, j, k)) * L.wd;
}

// ---------------------------------------------------------------------------
// z-marching stencil kernels.  The stencil's grouping is a sum over planes:
// per plane the centre c, the x pair, the y pair and the four xy diagonals;
// plane k contributes (c, x, y, xy) with weights (w0, w1, w2, w4) and planes
// k-1 and k+1 contribute the sums of theirs with weights (w3, w5, w6, w7).
// A thread owns one (i, j) column over kz planes and keeps the sums of three
// planes in registers, so each node costs the 3 x 3 loads of one new plane
// instead of 27.  The sums are formed in the oracle's operand order, so the
// results are those of nd_apply_at bit for bit.
// ---------------------------------------------------------------------------
struct NdPlane {
    double c, x, y, xy;
};

template <bool SUM2>
__device__ __forceinline__ double nd_val(const double* __restrict__ a, const double* __restrict__ a2, int q)
{
    return SUM2 ? a[q] + a2[q] : a[q];
}

template <bool SUM2>
__device__ __forceinline__ NdPlane nd_plane(const double* __restrict__ a, const double* __restrict__ a2, int z, int i,
                                            int im, int ip, int y0, int ym, int yp)
{
    NdPlane s;
    s.c = nd_val<SUM2>(a, a2, z + y0 + i);
    s.x = nd_val<SUM2>(a, a2, z + y0 + im) + nd_val<SUM2>(a, a2, z + y0 + ip);
    s.y = nd_val<SUM2>(a, a2, z + ym + i) + nd_val<SUM2>(a, a2, z + yp + i);
    s.xy = (nd_val<SUM2>(a, a2, z + ym + im) + nd_val<SUM2>(a, a2, z + ym + ip)) +
           (nd_val<SUM2>(a, a2, z + yp + im) + nd_val<SUM2>(a, a2, z + yp + ip));
    return s;
}

__device__ __forceinline__ double nd_apply_planes(const NLev& L, const NdPlane& m, const NdPlane& o, const NdPlane& p)
{
    const double fz = m.c + p.c, exz = m.x + p.x, eyz = m.y + p.y, cr = m.xy + p.xy;
    return (((L.w[0] * o.c + L.w[1] * o.x) + (L.w[2] * o.y + L.w[3] * fz)) +
            ((L.w[4] * o.xy + L.w[5] * exz) + (L.w[6] * eyz + L.w[7] * cr)));
}

// MODE 0  Jacobi sweep:  out = x + (b - L x) * wd          (x2, r unused)
// MODE 1  residual:      out = b - L x                      (x2, r unused)
// MODE 2  outer update:  out = x + x2,  r = b - L (x + x2),  |r|_inf -> norm
template <int MODE>
__global__ void __launch_bounds__(256) k_nd_march(double* __restrict__ out, const double* __restrict__ x,
                                                  const double* __restrict__ x2, const double* __restrict__ b,
                                                  double* __restrict__ r, NLev L, int kz,
                                                  unsigned long long* __restrict__ norm)
{
    constexpr bool SUM2 = (MODE == 2);
    // 1-D grid of (x tile, y tile, z chunk), x fastest, remapped so that each
    // XCD owns a contiguous run of rows and shares their halo rows in its L2
    const int ntx = (L.N[0] + 63) >> 6, nty = (L.N[1] + 3) >> 2;
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int tx = t % ntx, q = t / ntx;
    const int i = tx * 64 + threadIdx.x;
    const int j = (q % nty) * 4 + threadIdx.y;
    const int k0 = (q / nty) * kz;
    const int k1 = min(k0 + kz, L.N[2]);
    double m = 0.0;
    if (i < L.N[0] && j < L.N[1]) {
        const int im = nd_nb(i, -1, L.N[0], L.per[0]), ip = nd_nb(i, 1, L.N[0], L.per[0]);
        const int jm = nd_nb(j, -1, L.N[1], L.per[1]), jp = nd_nb(j, 1, L.N[1], L.per[1]);
        const int sy = L.N[0], sz = L.N[0] * L.N[1];
        const int y0 = j * sy, ym = jm * sy, yp = jp * sy;
        const bool dxy = (L.dlo[0] && i == 0) || (L.dhi[0] && i == L.N[0] - 1) || (L.dlo[1] && j == 0) ||
                         (L.dhi[1] && j == L.N[1] - 1);
        NdPlane pm = nd_plane<SUM2>(x, x2, nd_nb(k0, -1, L.N[2], L.per[2]) * sz, i, im, ip, y0, ym, yp);
        NdPlane po = nd_plane<SUM2>(x, x2, k0 * sz, i, im, ip, y0, ym, yp);
        // two planes per step: both planes' loads are issued before either
        // node is formed (a store between them would otherwise make the next
        // plane's loads wait for it)
        auto node = [&](int k, const NdPlane& a, const NdPlane& o, const NdPlane& p, double bc) {
            const int c = k * sz + y0 + i;
            const bool dir = dxy || (L.dlo[2] && k == 0) || (L.dhi[2] && k == L.N[2] - 1);
            if (MODE == 0) {
                out[c] = dir ? 0.0 : o.c + (bc - nd_apply_planes(L, a, o, p)) * L.wd;
            } else if (MODE == 1) {
                out[c] = dir ? 0.0 : bc - nd_apply_planes(L, a, o, p);
            } else {
                out[c] = o.c;
                const double v = dir ? 0.0 : bc - nd_apply_planes(L, a, o, p);
                r[c] = v;
                m = fmax(m, nd_abs_or_inf(v));
            }
        };
        int k = k0;
        for (; k + 1 < k1; k += 2) {
            const NdPlane pp = nd_plane<SUM2>(x, x2, nd_nb(k, 1, L.N[2], L.per[2]) * sz, i, im, ip, y0, ym, yp);
            const NdPlane pq = nd_plane<SUM2>(x, x2, nd_nb(k + 1, 1, L.N[2], L.per[2]) * sz, i, im, ip, y0, ym, yp);
            const double b0 = b[k * sz + y0 + i], b1 = b[(k + 1) * sz + y0 + i];
            node(k, pm, po, pp, b0);
            node(k + 1, po, pp, pq, b1);
            pm = pp;
            po = pq;
        }
        if (k < k1) {
            const NdPlane pp = nd_plane<SUM2>(x, x2, nd_nb(k, 1, L.N[2], L.per[2]) * sz, i, im, ip, y0, ym, yp);
            node(k, pm, po, pp, b[k * sz + y0 + i]);
        }
    }
    if (MODE == 2) nd_block_max(m, norm);
}

// The same sweeps for levels at least 64 nodes wide, with each plane of a
// workgroup's 64 x 4 column tile staged through LDS: the workgroup loads the
// tile and its one-node halo (66 x 6 nodes, periodic images and mirrors
// resolved as nd_nb does) once, one or two loads per thread, and every
// thread forms its plane sums from LDS.  The register form issues nine
// global loads per node and plane, which kept the texture-address units
// about three quarters busy and the sweep at a third of HBM bandwidth.
// Plane p of the march lives in buffer p & 1; plane p + 2 is loaded into
// registers one iteration ahead, so one barrier per plane separates its LDS
// store from the reads.  Plane sums, node update and operand order are
// k_nd_march's, so the results are its bits.
constexpr int kNdLx = 66;  // tile row of 64 nodes and its halo

// storage index of a halo position g (one beyond either end at most) of a
// direction with N nodes; positions further out are read by no node and
// map to 0
__device__ __forceinline__ int nd_halo(int g, int N, int per)
{
    if (g >= 0 && g < N) return g;
    if (g == -1) return nd_nb(0, -1, N, per);
    if (g == N) return nd_nb(N - 1, 1, N, per);
    return 0;
}

template <int MODE, int RY>
__global__ void __launch_bounds__(256) k_nd_march_lds(double* __restrict__ out, const double* __restrict__ x,
                                                      const double* __restrict__ x2, const double* __restrict__ b,
                                                      double* __restrict__ r, NLev L, int kz,
                                                      unsigned long long* __restrict__ norm)
{
    // RY rows per thread: a 64 x 4RY tile, threads (x, y) own rows y, y + 4, ...
    constexpr bool SUM2 = (MODE == 2);
    constexpr int TY = 4 * RY, LY = TY + 2, LN = kNdLx * LY, NS = (LN + 255) / 256;
    __shared__ double sp[2][LN];
    const int ntx = (L.N[0] + 63) >> 6, nty = (L.N[1] + TY - 1) / TY;
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int tx = t % ntx, q = t / ntx;
    const int x0 = tx * 64, y0 = (q % nty) * TY;
    const int k0 = (q / nty) * kz;
    const int k1 = min(k0 + kz, L.N[2]);
    const int tid = (int)threadIdx.x + 64 * (int)threadIdx.y;
    const int i = x0 + (int)threadIdx.x;
    const int sy = L.N[0], sz = L.N[0] * L.N[1];
    // the LDS slots this thread fills: in-plane storage offsets
    int so[NS];
    bool sv[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const int s = tid + 256 * m;
        sv[m] = s < LN;
        const int hx = s % kNdLx, hy = s / kNdLx;
        so[m] = sv[m] ? nd_halo(x0 - 1 + hx, L.N[0], L.per[0]) + sy * nd_halo(y0 - 1 + hy, L.N[1], L.per[1]) : 0;
    }
    // plane p of the march: p = 0 is k0's low neighbour, p >= 1 is
    // nd_nb(k0 + p - 2, +1) (p = 1: k0 itself)
    auto plane_k = [&](int p) {
        return p == 0 ? nd_nb(k0, -1, L.N[2], L.per[2]) : (p == 1 ? k0 : nd_nb(k0 + p - 2, 1, L.N[2], L.per[2]));
    };
    auto gload = [&](int p, double (&v)[NS]) {
        const int z = plane_k(p) * sz;
#pragma unroll
        for (int m = 0; m < NS; ++m) v[m] = sv[m] ? nd_val<SUM2>(x, x2, z + so[m]) : 0.0;
    };
    auto lstore = [&](int p, const double (&v)[NS]) {
#pragma unroll
        for (int m = 0; m < NS; ++m)
            if (sv[m]) sp[p & 1][tid + 256 * m] = v[m];
    };
    auto psum = [&](int p, int c0) {
        const double* a = sp[p & 1];
        NdPlane s;
        s.c = a[c0];
        s.x = a[c0 - 1] + a[c0 + 1];
        s.y = a[c0 - kNdLx] + a[c0 + kNdLx];
        s.xy = (a[c0 - kNdLx - 1] + a[c0 - kNdLx + 1]) + (a[c0 + kNdLx - 1] + a[c0 + kNdLx + 1]);
        return s;
    };
    int c0[RY], cxy[RY];
    bool act[RY], dxy[RY];
#pragma unroll
    for (int w = 0; w < RY; ++w) {
        const int ty = (int)threadIdx.y + 4 * w, j = y0 + ty;
        c0[w] = (ty + 1) * kNdLx + (int)threadIdx.x + 1;
        act[w] = i < L.N[0] && j < L.N[1];
        cxy[w] = act[w] ? j * sy + i : 0;
        dxy[w] = (L.dlo[0] && i == 0) || (L.dhi[0] && i == L.N[0] - 1) || (L.dlo[1] && j == 0) ||
                 (L.dhi[1] && j == L.N[1] - 1);
    }
    double m = 0.0;
    const int nk = k1 - k0, np = nk + 2;  // nodes per column, planes of the march
    // planes n + 3 and n + 4 are in flight in two register sets while plane
    // n + 2 goes to LDS (the loop takes two planes per trip, so each set has
    // a fixed name)
    double va[NS], vb[NS];
    gload(0, va);
    lstore(0, va);
    gload(1, va);
    lstore(1, va);
    __syncthreads();
    NdPlane pm[RY], po[RY];
#pragma unroll
    for (int w = 0; w < RY; ++w) {
        pm[w] = psum(0, c0[w]);
        po[w] = psum(1, c0[w]);
    }
    gload(2, va);
    if (3 < np) gload(3, vb);
    __syncthreads();
    auto step = [&](int n, double (&v)[NS]) {
        const int k = k0 + n;
        lstore(n + 2, v);
        double bc[RY];
#pragma unroll
        for (int w = 0; w < RY; ++w) bc[w] = act[w] ? b[k * sz + cxy[w]] : 0.0;
        if (n + 4 < np) gload(n + 4, v);
        __syncthreads();
        const bool dz = (L.dlo[2] && k == 0) || (L.dhi[2] && k == L.N[2] - 1);
#pragma unroll
        for (int w = 0; w < RY; ++w) {
            const NdPlane pp = psum(n + 2, c0[w]);
            if (act[w]) {
                const int c = k * sz + cxy[w];
                const bool dir = dxy[w] || dz;
                if (MODE == 0) {
                    out[c] = dir ? 0.0 : po[w].c + (bc[w] - nd_apply_planes(L, pm[w], po[w], pp)) * L.wd;
                } else if (MODE == 1) {
                    out[c] = dir ? 0.0 : bc[w] - nd_apply_planes(L, pm[w], po[w], pp);
                } else {
                    out[c] = po[w].c;
                    const double rv = dir ? 0.0 : bc[w] - nd_apply_planes(L, pm[w], po[w], pp);
                    r[c] = rv;
                    m = fmax(m, nd_abs_or_inf(rv));
                }
            }
            pm[w] = po[w];
            po[w] = pp;
        }
    };
    for (int n = 0; n < nk; n += 2) {
        step(n, va);
        if (n + 1 < nk) step(n + 1, vb);
    }
    if (MODE == 2) nd_block_max(m, norm);
}

// |a|_inf over a level (Dirichlet nodes hold zero): a grid-stride pass of a
// bounded grid, so one atomic per workgroup merges kNdNormBlocks maxima (one
// workgroup per 256 nodes serialised 66 K atomics on one word at 256^3)
constexpr int kNdNormBlocks = 1024;
__global__ void __launch_bounds__(256) k_nd_norm(const double* __restrict__ a, NLev L,
                                                 unsigned long long* __restrict__ norm)
{
    double m = 0.0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < L.tot; c += (long long)gridDim.x * blockDim.x)
        m = fmax(m, nd_abs_or_inf(a[c]));
    nd_block_max(m, norm);
}

// full weighting: weights (1/4, 1/2, 1/4) along each coarsened direction
__global__ void __launch_bounds__(256) k_nd_restrict(double* __restrict__ c, const double* __restrict__ f, NLev F,
                                                     NLev C)
{
    int I, J, K;
    if (!nd_node(C, I, J, K)) return;
    const int o = nd_idx(C, I, J, K);
    if (nd_dirichlet(C, I, J, K)) {
        c[o] = 0.0;
        return;
    }
    const int ci[3] = {I, J, K};
    int idx[3][3];
    double wt[3][3];
    int cnt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (F.r[d] == 2) {
            const int fi = 2 * ci[d];
            idx[d][0] = nd_nb(fi, -1, F.N[d], F.per[d]);
            idx[d][1] = fi;
            idx[d][2] = nd_nb(fi, 1, F.N[d], F.per[d]);
            wt[d][0] = 0.25;
            wt[d][1] = 0.5;
            wt[d][2] = 0.25;
            cnt[d] = 3;
        } else {
            idx[d][0] = idx[d][1] = idx[d][2] = ci[d];
            wt[d][0] = 1.0;
            wt[d][1] = wt[d][2] = 0.0;
            cnt[d] = 1;
        }
    }
    double s = 0.0;
    for (int a = 0; a < cnt[2]; ++a) {
        double sy = 0.0;
        for (int b = 0; b < cnt[1]; ++b) {
            double sx = 0.0;
            for (int e = 0; e < cnt[0]; ++e) sx += wt[0][e] * f[nd_idx(F, idx[0][e], idx[1][b], idx[2][a])];
            sy += wt[1][b] * sx;
        }
        s += wt[2][a] * sy;
    }
    c[o] = s;
}

// trilinear interpolation of the coarse correction, added to the fine one
__global__ void __launch_bounds__(256) k_nd_prolong_add(double* __restrict__ f, const double* __restrict__ c, NLev F,
                                                        NLev C)
{
    int i, j, k;
    if (!nd_node(F, i, j, k)) return;
    if (nd_dirichlet(F, i, j, k)) return;
    const int fi[3] = {i, j, k};
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (F.r[d] == 2) {
            lo[d] = fi[d] >> 1;
            hi[d] = (fi[d] & 1) ? ((lo[d] + 1 == C.N[d]) ? 0 : lo[d] + 1) : lo[d];
        } else {
            lo[d] = hi[d] = fi[d];
        }
    }
    const bool ox = lo[0] != hi[0], oy = lo[1] != hi[1], oz = lo[2] != hi[2];
    double z0, z1;
    {
        double a = c[nd_idx(C, lo[0], lo[1], lo[2])], b = c[nd_idx(C, hi[0], lo[1], lo[2])];
        double e = c[nd_idx(C, lo[0], hi[1], lo[2])], g = c[nd_idx(C, hi[0], hi[1], lo[2])];
        double y0 = ox ? 0.5 * (a + b) : a, y1 = ox ? 0.5 * (e + g) : e;
        z0 = oy ? 0.5 * (y0 + y1) : y0;
    }
    {
        double a = c[nd_idx(C, lo[0], lo[1], hi[2])], b = c[nd_idx(C, hi[0], lo[1], hi[2])];
        double e = c[nd_idx(C, lo[0], hi[1], hi[2])], g = c[nd_idx(C, hi[0], hi[1], hi[2])];
        double y0 = ox ? 0.5 * (a + b) : a, y1 = ox ? 0.5 * (e + g) : e;
        z1 = oy ? 0.5 * (y0 + y1) : y0;
    }
    const double v = oz ? 0.5 * (z0 + z1) : z0;
    const int o = nd_idx(F, i, j, k);
    f[o] += v;
}

// weight of a node in the solvability condition: 2^-(Neumann planes)
__device__ __forceinline__ double nd_weight(const NLev& L, int i, int j, int k)
{
    double w = 1.0;
    if (L.nlo[0] && i == 0) w *= 0.5;
    if (L.nhi[0] && i == L.N[0] - 1) w *= 0.5;
    if (L.nlo[1] && j == 0) w *= 0.5;
    if (L.nhi[1] && j == L.N[1] - 1) w *= 0.5;
    if (L.nlo[2] && k == 0) w *= 0.5;
    if (L.nhi[2] && k == L.N[2] - 1) w *= 0.5;
    return w;
}

// per-workgroup partial sums of w * a (fixed order within the workgroup)
__global__ void __launch_bounds__(256) k_nd_wsum(const double* __restrict__ a, NLev L, double* __restrict__ part)
{
    __shared__ double sh[256];
    int i, j, k;
    double s = 0.0;
    if (nd_node(L, i, j, k)) s = nd_weight(L, i, j, k) * a[nd_idx(L, i, j, k)];
    const int t = threadIdx.x + 64 * threadIdx.y;
    sh[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) sh[t] += sh[t + h];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = sh[0];
}

// sum of the partials in a fixed order, then mean = sum / wsum
__global__ void __launch_bounds__(1024) k_nd_mean(const double* __restrict__ part, int n, double wsum,
                                                  double* __restrict__ mean)
{
    __shared__ double sh[1024];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 1024) s += part[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *mean = sh[0] / wsum;
}

__global__ void __launch_bounds__(256) k_nd_sub(double* __restrict__ a, NLev L, const double* __restrict__ mean,
                                                int zero_dirichlet)
{
    int i, j, k;
    if (!nd_node(L, i, j, k)) return;
    const int c = nd_idx(L, i, j, k);
    if (zero_dirichlet && nd_dirichlet(L, i, j, k))
        a[c] = 0.0;
    else if (mean)
        a[c] -= *mean;
}

__global__ void __launch_bounds__(256) k_nd_copy(double* __restrict__ a, const double* __restrict__ b, int n)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) a[t] = b[t];
}

__global__ void __launch_bounds__(256) k_nd_zero(double* __restrict__ a, int n)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) a[t] = 0.0;
}

__global__ void k_nd_reset(unsigned long long* w) { *w = 0ull; }

// ---------------------------------------------------------------------------
// bottom level in one workgroup: (singular) weighted mean removal of the
// right-hand side, BiCGStab from a zero guess, then nsmooth Jacobi sweeps --
// the steps of the oracle's bottom solve
// ---------------------------------------------------------------------------
struct NdBottom {
    double *r, *rh, *p, *v, *s, *t, *tmp;
};

__device__ __forceinline__ void nd_decode(const NLev& L, int c, int& i, int& j, int& k)
{
    i = c % L.N[0];
    const int q = c / L.N[0];
    j = q % L.N[1];
    k = q / L.N[1];
}

__device__ double nd_bsum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = kNdBottomThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
__device__ double nd_bmax(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = kNdBottomThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ void nd_bapply(const NLev& L, const double* x, double* y)
{
    for (int c = threadIdx.x; c < L.tot; c += kNdBottomThreads) {
        int i, j, k;
        nd_decode(L, c, i, j, k);
        y[c] = nd_dirichlet(L, i, j, k) ? 0.0 : nd_apply_at(L, x, i, j, k);
    }
    __syncthreads();
}

__device__ void nd_bottom_body(double* __restrict__ x, double* __restrict__ b, const NLev& L, const NdBottom& W,
                               int singular, double wsum, double rtol, double atol, int maxiter, int nsmooth_ok,
                               int nsmooth_fail, int* __restrict__ stats, double* sh)
{
    const int T = L.tot;
    if (singular) {
        double s = 0.0;
        for (int c = threadIdx.x; c < T; c += kNdBottomThreads) {
            int i, j, k;
            nd_decode(L, c, i, j, k);
            s += nd_weight(L, i, j, k) * b[c];
        }
        const double m = nd_bsum(s, sh) / wsum;
        for (int c = threadIdx.x; c < T; c += kNdBottomThreads) b[c] -= m;
        __syncthreads();
    }
    double mx = 0.0;
    for (int c = threadIdx.x; c < T; c += kNdBottomThreads) {
        x[c] = 0.0;
        W.r[c] = b[c];
        W.rh[c] = b[c];
        mx = fmax(mx, nd_abs_or_inf(b[c]));
    }
    const double rnorm0 = nd_bmax(mx, sh);
    const double tol = fmax(rtol * rnorm0, atol);
    int ret = 1, it = 0;
    if (rnorm0 == 0.0 || rnorm0 <= tol) {
        ret = 0;
    } else {
        double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
        for (it = 1; it <= maxiter; ++it) {
            double a = 0.0;
            for (int c = threadIdx.x; c < T; c += kNdBottomThreads) a += W.rh[c] * W.r[c];
            const double rho = nd_bsum(a, sh);
            if (rho == 0.0) { ret = 2; break; }
            if (it == 1) {
                for (int c = threadIdx.x; c < T; c += kNdBottomThreads) W.p[c] = W.r[c];
            } else {
                const double beta = (rho / rho_prev) * (alpha / omega);
                for (int c = threadIdx.x; c < T; c += kNdBottomThreads)
                    W.p[c] = W.r[c] + beta * (W.p[c] - omega * W.v[c]);
            }
            __syncthreads();
            nd_bapply(L, W.p, W.v);
            a = 0.0;
            for (int c = threadIdx.x; c < T; c += kNdBottomThreads) a += W.rh[c] * W.v[c];
            const double rv = nd_bsum(a, sh);
            if (rv == 0.0) { ret = 2; break; }
            alpha = rho / rv;
            mx = 0.0;
            for (int c = threadIdx.x; c < T; c += kNdBottomThreads) {
                W.s[c] = W.r[c] - alpha * W.v[c];
                mx = fmax(mx, nd_abs_or_inf(W.s[c]));
            }
            if (nd_bmax(mx, sh) <= tol) {
                for (int c = threadIdx.x; c < T; c += kNdBottomThreads) x[c] += alpha * W.p[c];
                __syncthreads();
                ret = 0;
                break;
            }
            nd_bapply(L, W.s, W.t);
            double tt = 0.0, ts = 0.0;
            for (int c = threadIdx.x; c < T; c += kNdBottomThreads) {
                tt += W.t[c] * W.t[c];
                ts += W.t[c] * W.s[c];
            }
            const double ttt = nd_bsum(tt, sh);
            if (ttt == 0.0) { ret = 2; break; }
            omega = nd_bsum(ts, sh) / ttt;
            mx = 0.0;
            for (int c = threadIdx.x; c < T; c += kNdBottomThreads) {
                x[c] += alpha * W.p[c] + omega * W.s[c];
                W.r[c] = W.s[c] - omega * W.t[c];
                mx = fmax(mx, nd_abs_or_inf(W.r[c]));
            }
            if (nd_bmax(mx, sh) <= tol) { ret = 0; break; }
            if (omega == 0.0) { ret = 2; break; }
            rho_prev = rho;
        }
        if (it > maxiter) it = maxiter;
    }
    __syncthreads();
    const int ns = kNdSweeps * (ret == 0 ? nsmooth_ok : nsmooth_fail);
    for (int sw = 0; sw < ns; ++sw) {
        for (int c = threadIdx.x; c < T; c += kNdBottomThreads) {
            int i, j, k;
            nd_decode(L, c, i, j, k);
            W.tmp[c] = nd_dirichlet(L, i, j, k) ? 0.0 : x[c] + (b[c] - nd_apply_at(L, x, i, j, k)) * L.wd;
        }
        __syncthreads();
        for (int c = threadIdx.x; c < T; c += kNdBottomThreads) x[c] = W.tmp[c];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[0] += it;
        stats[1] += (ret != 0) ? 1 : 0;
    }
}

__global__ void __launch_bounds__(kNdBottomThreads) k_nd_bottom(double* __restrict__ x, double* __restrict__ b, NLev L,
                                                                NdBottom W, int singular, double wsum, double rtol,
                                                                double atol, int maxiter, int nsmooth_ok,
                                                                int nsmooth_fail, int* __restrict__ stats)
{
    __shared__ double sh[kNdBottomThreads];
    nd_bottom_body(x, b, L, W, singular, wsum, rtol, atol, maxiter, nsmooth_ok, nsmooth_fail, stats, sh);
}

// ---------------------------------------------------------------------------
// The coarse levels (at most kNdCoarseNodes nodes each) of a V-cycle in one
// workgroup: the grid kernels' sweeps, residual, transfers and bottom solve,
// node loops over the workgroup with a barrier after every pass.  A grid
// kernel per pass spent about 4 us on launch and drain for a few hundred
// nodes.  Same arithmetic as the grid kernels (nd_apply_at is the plane-sum
// stencil bit for bit), so the iterates are theirs.
// ---------------------------------------------------------------------------
constexpr int kNdCoarseNodes = 4096;

struct NdCoarseLev {
    double *cor, *res, *tmp;
};

// one Jacobi sweep xo = x + (b - L x) wd (zero: xo = b wd), Dirichlet nodes 0
__device__ __forceinline__ void ndc_sweep(const NLev& L, const double* x, const double* __restrict__ b, double* xo, bool zero)
{
    for (int c = threadIdx.x; c < L.tot; c += kNdBottomThreads) {
        int i, j, k;
        nd_decode(L, c, i, j, k);
        if (nd_dirichlet(L, i, j, k))
            xo[c] = 0.0;
        else
            xo[c] = zero ? b[c] * L.wd : x[c] + (b[c] - nd_apply_at(L, x, i, j, k)) * L.wd;
    }
    __syncthreads();
}

// nsweep sweeps ending in x (t: scratch), as nd_smooth
__device__ __forceinline__ void ndc_smooth(const NLev& L, double* x, const double* __restrict__ b, double* t, int nsweep, bool zero)
{
    if (nsweep == 0) {
        if (zero) {
            for (int c = threadIdx.x; c < L.tot; c += kNdBottomThreads) x[c] = 0.0;
            __syncthreads();
        }
        return;
    }
    double* src = x;
    double* dst = t;
    for (int s = 0; s < nsweep; ++s) {
        ndc_sweep(L, src, b, dst, zero && s == 0);
        double* q = src;
        src = dst;
        dst = q;
    }
    if (src != x) {
        for (int c = threadIdx.x; c < L.tot; c += kNdBottomThreads) x[c] = src[c];
        __syncthreads();
    }
}

// out = b - L x, Dirichlet nodes 0
__device__ __forceinline__ void ndc_residual(const NLev& L, const double* x, const double* __restrict__ b, double* out)
{
    for (int c = threadIdx.x; c < L.tot; c += kNdBottomThreads) {
        int i, j, k;
        nd_decode(L, c, i, j, k);
        out[c] = nd_dirichlet(L, i, j, k) ? 0.0 : b[c] - nd_apply_at(L, x, i, j, k);
    }
    __syncthreads();
}

// k_nd_restrict's full weighting at coarse node o
__device__ __forceinline__ double nd_restrict_at(const double* f, const NLev& F, int I, int J, int K)
{
    const int ci[3] = {I, J, K};
    int idx[3][3];
    double wt[3][3];
    int cnt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (F.r[d] == 2) {
            const int fi = 2 * ci[d];
            idx[d][0] = nd_nb(fi, -1, F.N[d], F.per[d]);
            idx[d][1] = fi;
            idx[d][2] = nd_nb(fi, 1, F.N[d], F.per[d]);
            wt[d][0] = 0.25;
            wt[d][1] = 0.5;
            wt[d][2] = 0.25;
            cnt[d] = 3;
        } else {
            idx[d][0] = idx[d][1] = idx[d][2] = ci[d];
            wt[d][0] = 1.0;
            wt[d][1] = wt[d][2] = 0.0;
            cnt[d] = 1;
        }
    }
    double s = 0.0;
    for (int a = 0; a < cnt[2]; ++a) {
        double sy = 0.0;
        for (int bb = 0; bb < cnt[1]; ++bb) {
            double sx = 0.0;
            for (int e = 0; e < cnt[0]; ++e) sx += wt[0][e] * f[nd_idx(F, idx[0][e], idx[1][bb], idx[2][a])];
            sy += wt[1][bb] * sx;
        }
        s += wt[2][a] * sy;
    }
    return s;
}

// k_nd_prolong_add's trilinear value at fine node (i, j, k)
__device__ __forceinline__ double nd_prolong_at(const double* c, const NLev& F, const NLev& C, int i, int j, int k)
{
    const int fi[3] = {i, j, k};
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (F.r[d] == 2) {
            lo[d] = fi[d] >> 1;
            hi[d] = (fi[d] & 1) ? ((lo[d] + 1 == C.N[d]) ? 0 : lo[d] + 1) : lo[d];
        } else {
            lo[d] = hi[d] = fi[d];
        }
    }
    const bool ox = lo[0] != hi[0], oy = lo[1] != hi[1], oz = lo[2] != hi[2];
    double z0, z1;
    {
        double a = c[nd_idx(C, lo[0], lo[1], lo[2])], b = c[nd_idx(C, hi[0], lo[1], lo[2])];
        double e = c[nd_idx(C, lo[0], hi[1], lo[2])], g = c[nd_idx(C, hi[0], hi[1], lo[2])];
        double y0 = ox ? 0.5 * (a + b) : a, y1 = ox ? 0.5 * (e + g) : e;
        z0 = oy ? 0.5 * (y0 + y1) : y0;
    }
    {
        double a = c[nd_idx(C, lo[0], lo[1], hi[2])], b = c[nd_idx(C, hi[0], lo[1], hi[2])];
        double e = c[nd_idx(C, lo[0], hi[1], hi[2])], g = c[nd_idx(C, hi[0], hi[1], hi[2])];
        double y0 = ox ? 0.5 * (a + b) : a, y1 = ox ? 0.5 * (e + g) : e;
        z1 = oy ? 0.5 * (y0 + y1) : y0;
    }
    return oz ? 0.5 * (z0 + z1) : z0;
}

// LDS pool of k_nd_coarse: when every coarse level (correction, residual,
// scratch) and the bottom vectors fit, they live here for the whole launch;
// the first coarse level's residual is copied in and its correction out
constexpr int kNdPool = 5632;  // doubles (44 KB)

template <bool LDS>
__global__ void __launch_bounds__(kNdBottomThreads) k_nd_coarse(const NLev* __restrict__ Ls,
                                                                const NdCoarseLev* __restrict__ lv, int lc, int nlev,
                                                                NdBottom W, int singular, double wsum, double rtol,
                                                                double atol, int maxiter, int nsmooth_ok,
                                                                int nsmooth_fail, int nu1, int nu2,
                                                                int* __restrict__ stats)
{
    __shared__ double sh[kNdBottomThreads];
    __shared__ double pool[LDS ? kNdPool : 1];
    // level l's arrays: in the pool (3 tot each, levels lc.. in order, then
    // the bottom vectors) or the global ones
    auto lev = [&](int l) {
        NdCoarseLev v = lv[l];
        if (LDS) {
            int off = 0;
            for (int m = lc; m < l; ++m) off += 3 * Ls[m].tot;
            const int t = Ls[l].tot;
            v.cor = pool + off;
            v.res = pool + off + t;
            v.tmp = pool + off + 2 * t;
        }
        return v;
    };
    if (LDS) {
        const NdCoarseLev g = lv[lc], v = lev(lc);
        for (int c = threadIdx.x; c < Ls[lc].tot; c += kNdBottomThreads) v.res[c] = g.res[c];
        int off = 0;
        for (int m = lc; m < nlev; ++m) off += 3 * Ls[m].tot;
        const int tb = Ls[nlev - 1].tot;
        double* q = pool + off;
        W.r = q;
        W.rh = q + tb;
        W.p = q + 2 * tb;
        W.v = q + 3 * tb;
        W.s = q + 4 * tb;
        W.t = q + 5 * tb;
        W.tmp = q + 6 * tb;
        __syncthreads();
    }
    for (int l = lc; l < nlev - 1; ++l) {
        const NLev L = Ls[l], C = Ls[l + 1];
        const NdCoarseLev v = lev(l);
        ndc_smooth(L, v.cor, v.res, v.tmp, kNdSweeps * nu1, true);
        ndc_residual(L, v.cor, v.res, v.tmp);
        double* cres = lev(l + 1).res;
        for (int o = threadIdx.x; o < C.tot; o += kNdBottomThreads) {
            int I, J, K;
            nd_decode(C, o, I, J, K);
            cres[o] = nd_dirichlet(C, I, J, K) ? 0.0 : nd_restrict_at(v.tmp, L, I, J, K);
        }
        __syncthreads();
    }
    {
        const NLev Lb = Ls[nlev - 1];
        const NdCoarseLev v = lev(nlev - 1);
        nd_bottom_body(v.cor, v.res, Lb, W, singular, wsum, rtol, atol, maxiter, nsmooth_ok, nsmooth_fail, stats, sh);
        __syncthreads();
    }
    for (int l = nlev - 2; l >= lc; --l) {
        const NLev L = Ls[l], C = Ls[l + 1];
        const NdCoarseLev v = lev(l);
        const double* xc = lev(l + 1).cor;
        for (int c = threadIdx.x; c < L.tot; c += kNdBottomThreads) {
            int i, j, k;
            nd_decode(L, c, i, j, k);
            if (!nd_dirichlet(L, i, j, k)) v.cor[c] += nd_prolong_at(xc, L, C, i, j, k);
        }
        __syncthreads();
        ndc_smooth(L, v.cor, v.res, v.tmp, kNdSweeps * nu2, false);
    }
    if (LDS) {
        const NdCoarseLev g = lv[lc], v = lev(lc);
        for (int c = threadIdx.x; c < Ls[lc].tot; c += kNdBottomThreads) g.cor[c] = v.cor[c];
    }
}

// ---------------------------------------------------------------------------
// right-hand side and correction on the BoxArray
// ---------------------------------------------------------------------------
struct NdGeom {
    int N[3];
    int per[3];
    int nlo[3], nhi[3];
    int dlo[3], dhi[3];
    double f[3];  // 0.25 / dx
};

// nodal divergence of the cell velocity.  Launched over each box's cells grown
// by one at the high side; a thread owns node (i,j,k) if it lies in the box,
// or on the domain's high side of a non-periodic direction.  The eight cells
// (i-1|i, j-1|j, k-1|k) come from the box's FAB (ghosts as ApplyProjection
// leaves them: zero on physical boundaries, filled elsewhere).
__global__ void __launch_bounds__(256) k_nd_rhs(CMFDev vel, double* __restrict__ rhs, NdGeom g, RegDev r,
                                                const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const int* bx = box + 6 * b;
    const int gi[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (gi[d] > bx[3 + d] && (g.per[d] || gi[d] != g.dhi[d] + 1)) return;
    }
    const int id[3] = {i - g.dlo[0], j - g.dlo[1], k - g.dlo[2]};
    CFab f = fab_at(vel, box, b);
    const int c000 = f.idx(i - 1, j - 1, k - 1);
    const int sy = f.n0, sz = f.n0 * f.n1;
    const double* u = f.p + c000;
    const double* v = f.p + c000 + f.cs;
    const double* w = f.p + c000 + 2 * f.cs;
#define A(p, a, bb, c) p[(a) + sy * (bb) + sz * (c)]
    const double dux = ((A(u, 1, 0, 0) - A(u, 0, 0, 0)) + (A(u, 1, 1, 0) - A(u, 0, 1, 0))) +
                       ((A(u, 1, 0, 1) - A(u, 0, 0, 1)) + (A(u, 1, 1, 1) - A(u, 0, 1, 1)));
    const double dvy = ((A(v, 0, 1, 0) - A(v, 0, 0, 0)) + (A(v, 1, 1, 0) - A(v, 1, 0, 0))) +
                       ((A(v, 0, 1, 1) - A(v, 0, 0, 1)) + (A(v, 1, 1, 1) - A(v, 1, 0, 1)));
    const double dwz = ((A(w, 0, 0, 1) - A(w, 0, 0, 0)) + (A(w, 1, 0, 1) - A(w, 1, 0, 0))) +
                       ((A(w, 0, 1, 1) - A(w, 0, 1, 0)) + (A(w, 1, 1, 1) - A(w, 1, 1, 0)));
#undef A
    double dv = (dux * g.f[0] + dvy * g.f[1]) + dwz * g.f[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (g.nlo[d] && id[d] == 0) dv *= 2.0;
        if (g.nhi[d] && id[d] == g.N[d] - 1) dv *= 2.0;
    }
    rhs[id[0] + g.N[0] * (id[1] + g.N[1] * id[2])] = dv;
}

// vel -= sigma grad phi on valid cells; gp = grad phi (average of the four
// node differences per direction) if has_gp
__global__ void __launch_bounds__(256) k_nd_correct(MFDev vel, MFDev gp, int has_gp, const double* __restrict__ phi,
                                                    NdGeom g, double sigma, RegDev r, const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const int a0 = i - g.dlo[0], b0 = j - g.dlo[1], c0 = k - g.dlo[2];
    const int a1 = (a0 + 1 == g.N[0]) ? 0 : a0 + 1;
    const int b1 = (b0 + 1 == g.N[1]) ? 0 : b0 + 1;
    const int c1 = (c0 + 1 == g.N[2]) ? 0 : c0 + 1;
#define P(a, bb, c) phi[(a) + g.N[0] * ((bb) + g.N[1] * (c))]
    double gx = ((P(a1, b0, c0) - P(a0, b0, c0)) + (P(a1, b1, c0) - P(a0, b1, c0))) +
                ((P(a1, b0, c1) - P(a0, b0, c1)) + (P(a1, b1, c1) - P(a0, b1, c1)));
    double gy = ((P(a0, b1, c0) - P(a0, b0, c0)) + (P(a1, b1, c0) - P(a1, b0, c0))) +
                ((P(a0, b1, c1) - P(a0, b0, c1)) + (P(a1, b1, c1) - P(a1, b0, c1)));
    double gz = ((P(a0, b0, c1) - P(a0, b0, c0)) + (P(a1, b0, c1) - P(a1, b0, c0))) +
                ((P(a0, b1, c1) - P(a0, b1, c0)) + (P(a1, b1, c1) - P(a1, b1, c0)));
#undef P
    gx *= g.f[0];
    gy *= g.f[1];
    gz *= g.f[2];
    Fab fv = fab_at(vel, box, b);
    fv(i, j, k, 0) = fv(i, j, k, 0) - sigma * gx;
    fv(i, j, k, 1) = fv(i, j, k, 1) - sigma * gy;
    fv(i, j, k, 2) = fv(i, j, k, 2) - sigma * gz;
    if (has_gp) {
        Fab fg = fab_at(gp, box, b);
        fg(i, j, k, 0) = gx;
        fg(i, j, k, 1) = gy;
        fg(i, j, k, 2) = gz;
    }
}

// MultiFab::setBndry(val): every point of the FAB storage outside the box
__global__ void __launch_bounds__(256) k_set_bndry(MFDev m, int ncomp, double val, RegDev r,
                                                   const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const int* bx = box + 6 * b;
    const int hi0 = bx[3] + (m.itype == 1), hi1 = bx[4] + (m.itype == 2), hi2 = bx[5] + (m.itype == 3);
    if (i >= bx[0] && i <= hi0 && j >= bx[1] && j <= hi1 && k >= bx[2] && k <= hi2) return;
    Fab f = fab_at(m, box, b);
    for (int n = 0; n < ncomp; ++n) f(i, j, k, n) = val;
}

// vel += gp * (scale / rho) on valid cells (the non-incremental branch of
// ApplyProjection, incflo_apply_nodal_projection.cpp:209-233, without mesh
// mapping: fac = 1)
__global__ void __launch_bounds__(256) k_add_gradp(MFDev vel, CMFDev gp, CMFDev rho, double scale, RegDev r,
                                                   const int* __restrict__ box)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab fv = fab_at(vel, box, b);
    CFab fg = fab_at(gp, box, b), fr = fab_at(rho, box, b);
    const double soverrho = scale / fr(i, j, k);
    for (int n = 0; n < 3; ++n) fv(i, j, k, n) += fg(i, j, k, n) * soverrho;
}

}  // namespace ksgf

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct ksgf_nodalproj {
    const ksgf_boxarray* ba = nullptr;
    ksgf_mlmg_options opts;
    int lobc[3], hibc[3];
    double sigma = 1.0;
    int nlev = 0;
    ksgf::NLev L[ksgf::kNdMaxLev];
    double wsum[ksgf::kNdMaxLev];
    int singular = 0;
    // level arrays: correction, residual (the level's right-hand side), scratch
    std::vector<double*> cor, res, tmp;
    double *phi[2] = {nullptr, nullptr}, *rhs = nullptr;
    int cur = 0;  // phi[cur] holds the solution
    double* part = nullptr;
    int npart = 0;
    double* mean = nullptr;
    unsigned long long* norm = nullptr;
    int* stats = nullptr;
    ksgf::NdBottom bw{};
    int lc = 0;                            // first level of the one-workgroup coarse cycle (k_nd_coarse)
    ksgf::NLev* dL = nullptr;              // device copy of L
    ksgf::NdCoarseLev* dlv = nullptr;      // device table of the level arrays
    std::vector<double*> allocs;
};

namespace ksgf {

static void nd_weights(double sigma, const double* h, double* w)
{
    const double ax = sigma / (h[0] * h[0]), ay = sigma / (h[1] * h[1]), az = sigma / (h[2] * h[2]);
    const double s = (ax + ay) + az;
    w[0] = -(8.0 / 9.0) * s;
    w[1] = (2.0 / 9.0) * ((2.0 * ax - ay) - az);
    w[2] = (2.0 / 9.0) * ((2.0 * ay - ax) - az);
    w[3] = (2.0 / 9.0) * ((2.0 * az - ax) - ay);
    w[4] = (ax + ay) / 9.0 - az / 18.0;
    w[5] = (ax + az) / 9.0 - ay / 18.0;
    w[6] = (ay + az) / 9.0 - ax / 18.0;
    w[7] = s / 36.0;
}

// Jacobi weight 1 / lambda_max of D^-1 A; the symbol is multilinear in the
// cosines, so its extremes are at the corners of {0, pi}^3
static double nd_omega(const double* w)
{
    double lmax = 0.0;
    for (int m = 1; m < 8; ++m) {
        const double cx = (m & 1) ? -1.0 : 1.0, cy = (m & 2) ? -1.0 : 1.0, cz = (m & 4) ? -1.0 : 1.0;
        const double sym = ((w[0] + 2.0 * (w[1] * cx + w[2] * cy + w[3] * cz)) +
                            4.0 * (w[4] * cx * cy + w[5] * cx * cz + w[6] * cy * cz)) +
                           8.0 * w[7] * cx * cy * cz;
        const double lam = sym / w[0];
        if (lam > lmax) lmax = lam;
    }
    return 1.0 / lmax;
}

static void nd_set_coefs(ksgf_nodalproj* np)
{
    for (int l = 0; l < np->nlev; ++l) {
        NLev& L = np->L[l];
        double h[3];
        for (int d = 0; d < 3; ++d) {
            int f = 1;
            for (int q = 0; q < l; ++q) f *= np->L[q].r[d];
            h[d] = np->ba->dx[d] * f;
        }
        nd_weights(np->sigma, h, L.w);
        L.wd = nd_omega(L.w) / L.w[0];
    }
}

static dim3 nd_grid(const NLev& L) { return dim3((L.N[0] + 63) / 64, (L.N[1] + 3) / 4, L.N[2]); }
static const dim3 kNdBlock(64, 4, 1);

// planes per thread of the z-marching kernels: up to kz_max, keeping about
// 2048 workgroups or more on the large levels (KSGF_ND_KZ overrides kz_max)
static int nd_kz(const NLev& L)
{
    static const int kz_max = [] {
        const char* e = std::getenv("KSGF_ND_KZ");
        const int v = e ? std::atoi(e) : 0;
        return (v >= 1 && v <= 64) ? v : 32;
    }();
    const long long bxy = (long long)((L.N[0] + 63) / 64) * ((L.N[1] + 3) / 4);
    long long kz = (long long)L.N[2] * bxy / 2048;
    return (int)std::max(1LL, std::min((long long)kz_max, kz));
}

template <int MODE>
static void nd_launch_march(double* out, const double* x, const double* x2, const double* b, double* r, const NLev& L,
                            unsigned long long* norm, hipStream_t st)
{
    const int kz = nd_kz(L);
    const long long nb = (long long)((L.N[0] + 63) / 64) * ((L.N[1] + 3) / 4) * ((L.N[2] + kz - 1) / kz);
    if (L.N[0] >= 64 && L.N[1] >= 8) {
        const long long nb2 = (long long)((L.N[0] + 63) / 64) * ((L.N[1] + 7) / 8) * ((L.N[2] + kz - 1) / kz);
        hipLaunchKernelGGL((k_nd_march_lds<MODE, 2>), dim3((unsigned)nb2), kNdBlock, 0, st, out, x, x2, b, r, L, kz,
                           norm);
    } else if (L.N[0] >= 64 && L.N[1] >= 4) {
        hipLaunchKernelGGL((k_nd_march_lds<MODE, 1>), dim3((unsigned)nb), kNdBlock, 0, st, out, x, x2, b, r, L, kz,
                           norm);
    } else {
        hipLaunchKernelGGL(k_nd_march<MODE>, dim3((unsigned)nb), kNdBlock, 0, st, out, x, x2, b, r, L, kz, norm);
    }
}

static int nd_alloc(ksgf_nodalproj* np, double** p, size_t n)
{
    if (hipMalloc((void**)p, sizeof(double) * std::max<size_t>(n, 1)) != hipSuccess) {
        set_error("hipMalloc (nodal projector)");
        return KSGF_ERR_NOMEM;
    }
    np->allocs.push_back(*p);
    return KSGF_OK;
}

// the level table of the one-workgroup coarse cycle (levels and their
// coefficients: again after every coefficient change)
static int nd_upload_levels(ksgf_nodalproj* np)
{
    if (!np->dL) return KSGF_OK;
    KSGF_HIP_CHECK(hipMemcpy(np->dL, np->L, sizeof(NLev) * np->nlev, hipMemcpyHostToDevice));
    return KSGF_OK;
}

static void nd_weighted_mean_sub(ksgf_nodalproj* np, int l, double* a, hipStream_t st)
{
    const NLev& L = np->L[l];
    dim3 gr = nd_grid(L);
    hipLaunchKernelGGL(k_nd_wsum, gr, kNdBlock, 0, st, a, L, np->part);
    const int n = (int)(gr.x * gr.y * gr.z);
    hipLaunchKernelGGL(k_nd_mean, dim3(1), dim3(1024), 0, st, np->part, n, np->wsum[l], np->mean);
    hipLaunchKernelGGL(k_nd_sub, gr, kNdBlock, 0, st, a, L, np->mean, 0);
}

// nsweep Jacobi sweeps on x (zero guess if zero); the result ends in x
static void nd_smooth(ksgf_nodalproj* np, int l, double* x, const double* b, int nsweep, bool zero, hipStream_t st)
{
    const NLev& L = np->L[l];
    dim3 gr = nd_grid(L);
    if (nsweep == 0) {
        if (zero) hipLaunchKernelGGL(k_nd_zero, dim3((L.tot + 255) / 256), dim3(256), 0, st, x, L.tot);
        return;
    }
    double* t = np->tmp[l];
    // out-of-place sweeps alternating between x and t, starting where the
    // last one lands in x; a zero-guess first sweep reads nothing, so it may
    // write x itself (an odd count); otherwise an odd count ends with a copy
    double* src = x;
    double* dst = (nsweep % 2 == 1 && zero) ? x : t;
    for (int s = 0; s < nsweep; ++s) {
        if (s == 0 && zero)
            hipLaunchKernelGGL(k_nd_jacobi<true>, gr, kNdBlock, 0, st, dst, src, b, L);
        else
            nd_launch_march<0>(dst, src, nullptr, b, nullptr, L, nullptr, st);
        src = dst;
        dst = (src == x) ? t : x;
    }
    if (src != x) hipLaunchKernelGGL(k_nd_copy, dim3((L.tot + 255) / 256), dim3(256), 0, st, x, src, L.tot);
}

static void nd_vcycle(ksgf_nodalproj* np, int l, hipStream_t st)
{
    const ksgf_mlmg_options& o = np->opts;
    if (l >= np->lc && np->dlv) {
        long long need = 7LL * np->L[np->nlev - 1].tot;
        for (int m = l; m < np->nlev; ++m) need += 3LL * np->L[m].tot;
        if (need <= kNdPool)
            hipLaunchKernelGGL(k_nd_coarse<true>, dim3(1), dim3(kNdBottomThreads), 0, st, (const NLev*)np->dL,
                               (const NdCoarseLev*)np->dlv, l, np->nlev, np->bw, np->singular,
                               np->wsum[np->nlev - 1], o.bottom_rtol, o.bottom_atol, o.bottom_maxiter,
                               o.num_bottom_smooth, o.num_final_smooth, o.num_pre_smooth, o.num_post_smooth,
                               np->stats);
        else
            hipLaunchKernelGGL(k_nd_coarse<false>, dim3(1), dim3(kNdBottomThreads), 0, st, (const NLev*)np->dL,
                               (const NdCoarseLev*)np->dlv, l, np->nlev, np->bw, np->singular,
                               np->wsum[np->nlev - 1], o.bottom_rtol, o.bottom_atol, o.bottom_maxiter,
                               o.num_bottom_smooth, o.num_final_smooth, o.num_pre_smooth, o.num_post_smooth,
                               np->stats);
        return;
    }
    if (l == np->nlev - 1) {
        hipLaunchKernelGGL(k_nd_bottom, dim3(1), dim3(kNdBottomThreads), 0, st, np->cor[l], np->res[l], np->L[l],
                           np->bw, np->singular, np->wsum[l], o.bottom_rtol, o.bottom_atol, o.bottom_maxiter,
                           o.num_bottom_smooth, o.num_final_smooth, np->stats);
        return;
    }
    const NLev& L = np->L[l];
    dim3 gr = nd_grid(L);
    nd_smooth(np, l, np->cor[l], np->res[l], kNdSweeps * o.num_pre_smooth, true, st);
    nd_launch_march<1>(np->tmp[l], np->cor[l], nullptr, np->res[l], nullptr, L, nullptr, st);
    hipLaunchKernelGGL(k_nd_restrict, nd_grid(np->L[l + 1]), kNdBlock, 0, st, np->res[l + 1], np->tmp[l], L,
                       np->L[l + 1]);
    nd_vcycle(np, l + 1, st);
    hipLaunchKernelGGL(k_nd_prolong_add, gr, kNdBlock, 0, st, np->cor[l], np->cor[l + 1], L, np->L[l + 1]);
    nd_smooth(np, l, np->cor[l], np->res[l], kNdSweeps * o.num_post_smooth, false, st);
}

static int nd_read_norm(ksgf_nodalproj* np, hipStream_t st, double* out)
{
    unsigned long long bits = 0;
    KSGF_HIP_CHECK(hipMemcpyAsync(&bits, np->norm, sizeof(bits), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    std::memcpy(out, &bits, sizeof(double));
    return KSGF_OK;
}

// solve with np->rhs (level-0 right-hand side, device) already set; result in
// np->phi[np->cur]
static int nd_solve(ksgf_nodalproj* np, double rtol, double atol, ksgf_mlmg_info* info, hipStream_t st)
{
    const NLev& L0 = np->L[0];
    dim3 gr = nd_grid(L0);
    const int nb = (L0.tot + 255) / 256;
    hipLaunchKernelGGL(k_nd_sub, gr, kNdBlock, 0, st, np->rhs, L0, (const double*)nullptr, 1);
    if (np->singular) nd_weighted_mean_sub(np, 0, np->rhs, st);
    hipLaunchKernelGGL(k_nd_reset, dim3(1), dim3(1), 0, st, np->norm);
    hipLaunchKernelGGL(k_nd_norm, dim3(std::min(kNdNormBlocks, (L0.tot + 255) / 256)), dim3(256), 0, st, np->rhs, L0,
                       np->norm);
    KSGF_LAUNCH_CHECK();
    KSGF_HIP_CHECK(hipMemsetAsync(np->stats, 0, 2 * sizeof(int), st));
    np->cur = 0;
    hipLaunchKernelGGL(k_nd_zero, dim3(nb), dim3(256), 0, st, np->phi[0], L0.tot);
    hipLaunchKernelGGL(k_nd_copy, dim3(nb), dim3(256), 0, st, np->res[0], np->rhs, L0.tot);
    double rhsnorm = 0.0;
    int e = nd_read_norm(np, st, &rhsnorm);
    if (e) return e;
    double r = rhsnorm;
    const double target = std::max(atol, rtol * rhsnorm);
    ksgf_mlmg_info inf{};
    inf.r0 = r;
    inf.converged = (r <= target);
    inf.nlevels = np->nlev;
    int diverged = 0;
    if (!inf.converged) {
        for (int it = 1; it <= np->opts.maxiter; ++it) {
            nd_vcycle(np, 0, st);
            hipLaunchKernelGGL(k_nd_reset, dim3(1), dim3(1), 0, st, np->norm);
            nd_launch_march<2>(np->phi[1 - np->cur], np->phi[np->cur], np->cor[0], np->rhs, np->res[0], L0, np->norm,
                               st);
            KSGF_LAUNCH_CHECK();
            np->cur = 1 - np->cur;
            e = nd_read_norm(np, st, &r);
            if (e) return e;
            inf.iters = it;
            if (r <= target) {
                inf.converged = 1;
                break;
            }
            // 1e20 |rhs| or not finite: MLMG::solve's "failing so lets stop here"
            if (!(r <= 1.0e20 * rhsnorm)) {
                diverged = 1;
                break;
            }
        }
    }
    int hs[2] = {0, 0};
    KSGF_HIP_CHECK(hipMemcpyAsync(hs, np->stats, sizeof(hs), hipMemcpyDeviceToHost, st));
    KSGF_HIP_CHECK(hipStreamSynchronize(st));
    inf.rfinal = r;
    inf.bottom_iters = hs[0];
    inf.bottom_fail = hs[1];
    if (info) *info = inf;
    if (diverged) {
        set_error("nodal MLMG diverged (residual above 1e20 |rhs| or not finite)");
        return KSGF_ERR_DIVERGED;
    }
    return inf.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
}

static NdGeom nd_geom(const ksgf_nodalproj* np)
{
    NdGeom g;
    for (int d = 0; d < 3; ++d) {
        g.N[d] = np->L[0].N[d];
        g.per[d] = np->L[0].per[d];
        g.nlo[d] = np->L[0].nlo[d];
        g.nhi[d] = np->L[0].nhi[d];
        g.dlo[d] = np->ba->dlo[d];
        g.dhi[d] = np->ba->dhi[d];
        g.f[d] = 0.25 / np->ba->dx[d];
    }
    return g;
}

static int nd_compute_rhs(ksgf_nodalproj* np, const double* vel, int ng_vel, double* rhs, hipStream_t st)
{
    CMFDev v;
    int e = cmfdev(np->ba, vel, KSGF_CELL, 3, ng_vel, &v);
    if (e) return e;
    int err = KSGF_OK;
    const int glo[3] = {0, 0, 0}, ghi[3] = {1, 1, 1};
    const Region* r = get_region(np->ba, glo, ghi, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_nd_rhs, dim3(r->nchunks), dim3(kChunk), 0, st, v, rhs, nd_geom(np), regdev(r, np->ba->nbox),
                       np->ba->dbox);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

}  // namespace ksgf

using namespace ksgf;

extern "C" {

int ksgf_nodalproj_destroy(ksgf_nodalproj* np)
{
    if (!np) return KSGF_OK;
    for (double* p : np->allocs) (void)hipFree(p);
    if (np->norm) (void)hipFree(np->norm);
    if (np->stats) (void)hipFree(np->stats);
    delete np;
    return KSGF_OK;
}

int ksgf_nodalproj_create(const ksgf_boxarray* ba, const int* lobc, const int* hibc, double sigma,
                          const ksgf_mlmg_options* opts, ksgf_nodalproj** out)
{
    if (!ba || !lobc || !hibc || !out) {
        set_error("ksgf_nodalproj_create: null argument");
        return KSGF_ERR_INVALID;
    }
    *out = nullptr;
    for (int d = 0; d < 3; ++d) {
        for (int bc : {lobc[d], hibc[d]})
            if (bc != KSGF_LO_PERIODIC && bc != KSGF_LO_NEUMANN && bc != KSGF_LO_DIRICHLET) {
                set_error("ksgf_nodalproj_create: domain BC must be periodic, Neumann or Dirichlet (inflow is not provided)");
                return bc >= 0 ? KSGF_ERR_UNSUPPORTED : KSGF_ERR_INVALID;
            }
        if ((lobc[d] == KSGF_LO_PERIODIC) != (hibc[d] == KSGF_LO_PERIODIC) ||
            (lobc[d] == KSGF_LO_PERIODIC) != (ba->periodic[d] != 0)) {
            set_error("ksgf_nodalproj_create: periodic BCs must match the geometry on both sides");
            return KSGF_ERR_INVALID;
        }
    }
    if (!(sigma > 0.0)) {
        set_error("ksgf_nodalproj_create: sigma must be positive");
        return KSGF_ERR_INVALID;
    }
    ksgf_nodalproj* np = new ksgf_nodalproj();
    np->ba = ba;
    if (opts)
        np->opts = *opts;
    else
        ksgf_mlmg_default_options(&np->opts);
    if (np->opts.precision != KSGF_MG_FP64) {
        delete np;
        set_error("ksgf_nodalproj_create: the nodal solver runs in FP64 only");
        return KSGF_ERR_UNSUPPORTED;
    }
    np->sigma = sigma;
    np->singular = 1;
    for (int d = 0; d < 3; ++d) {
        np->lobc[d] = lobc[d];
        np->hibc[d] = hibc[d];
        if (lobc[d] == KSGF_LO_DIRICHLET || hibc[d] == KSGF_LO_DIRICHLET) np->singular = 0;
    }
    // hierarchy: the MAC projector's coarsening rule on the cell counts
    int cn[3];
    double ch[3];
    for (int d = 0; d < 3; ++d) {
        cn[d] = ba->dhi[d] - ba->dlo[d] + 1;
        ch[d] = ba->dx[d];
    }
    int nl = 0;
    for (;;) {
        NLev& L = np->L[nl];
        long long tot = 1;
        for (int d = 0; d < 3; ++d) {
            L.per[d] = lobc[d] == KSGF_LO_PERIODIC;
            L.N[d] = L.per[d] ? cn[d] : cn[d] + 1;
            L.dlo[d] = lobc[d] == KSGF_LO_DIRICHLET;
            L.dhi[d] = hibc[d] == KSGF_LO_DIRICHLET;
            L.nlo[d] = lobc[d] == KSGF_LO_NEUMANN;
            L.nhi[d] = hibc[d] == KSGF_LO_NEUMANN;
            L.r[d] = 1;
            tot *= L.N[d];
        }
        if (tot >= (1LL << 31) || L.N[2] > 65535 || (L.N[1] + 3) / 4 > 65535) {
            delete np;
            set_error("ksgf_nodalproj_create: level too large");
            return KSGF_ERR_INVALID;
        }
        L.tot = (int)tot;
        // sum of the solvability weights 2^-(Neumann planes), exact
        double ws[3];
        for (int d = 0; d < 3; ++d) ws[d] = L.N[d] - 0.5 * (L.nlo[d] + L.nhi[d]);
        np->wsum[nl] = ws[0] * ws[1] * ws[2];
        ++nl;
        if (nl >= kNdMaxLev || nl > np->opts.max_coarsening_level) break;
        int can[3], any = 0;
        for (int d = 0; d < 3; ++d) can[d] = (cn[d] % 2 == 0 && cn[d] / 2 >= 2);
        if (!np->opts.do_semicoarsening) {
            if (!(can[0] && can[1] && can[2])) break;
            for (int d = 0; d < 3; ++d) L.r[d] = 2;
            any = 1;
        } else {
            double hmin = 1e300;
            for (int d = 0; d < 3; ++d) hmin = std::min(hmin, ch[d]);
            for (int d = 0; d < 3; ++d)
                if (can[d] && ch[d] <= 1.5 * hmin) {
                    L.r[d] = 2;
                    any = 1;
                }
            if (!any) break;
        }
        for (int d = 0; d < 3; ++d) {
            cn[d] /= L.r[d];
            ch[d] *= L.r[d];
        }
    }
    np->nlev = nl;
    nd_set_coefs(np);
    int e = KSGF_OK;
    np->cor.resize(nl);
    np->res.resize(nl);
    np->tmp.resize(nl);
    for (int l = 0; l < nl && !e; ++l) {
        const size_t n = (size_t)np->L[l].tot;
        if ((e = nd_alloc(np, &np->cor[l], n)) || (e = nd_alloc(np, &np->res[l], n)) ||
            (e = nd_alloc(np, &np->tmp[l], n)))
            break;
    }
    const size_t n0 = (size_t)np->L[0].tot;
    const NLev& Lb = np->L[nl - 1];
    dim3 gr = nd_grid(np->L[0]);
    np->npart = (int)(gr.x * gr.y * gr.z);
    if (!e && ((e = nd_alloc(np, &np->phi[0], n0)) || (e = nd_alloc(np, &np->phi[1], n0)) ||
               (e = nd_alloc(np, &np->rhs, n0)) || (e = nd_alloc(np, &np->part, (size_t)np->npart)) ||
               (e = nd_alloc(np, &np->mean, 1)) || (e = nd_alloc(np, &np->bw.r, Lb.tot)) ||
               (e = nd_alloc(np, &np->bw.rh, Lb.tot)) || (e = nd_alloc(np, &np->bw.p, Lb.tot)) ||
               (e = nd_alloc(np, &np->bw.v, Lb.tot)) || (e = nd_alloc(np, &np->bw.s, Lb.tot)) ||
               (e = nd_alloc(np, &np->bw.t, Lb.tot)) || (e = nd_alloc(np, &np->bw.tmp, Lb.tot)))) {
    }
    if (!e && (hipMalloc((void**)&np->norm, sizeof(unsigned long long)) != hipSuccess ||
               hipMalloc((void**)&np->stats, 2 * sizeof(int)) != hipSuccess)) {
        set_error("hipMalloc (nodal projector)");
        e = KSGF_ERR_NOMEM;
    }
    if (!e && hipMemset(np->phi[0], 0, sizeof(double) * n0) != hipSuccess) {
        set_error("hipMemset (nodal projector)");
        e = KSGF_ERR_HIP;
    }
    if (!e) {
        // levels of at most kNdCoarseNodes nodes, from the bottom up, run in
        // the one-workgroup coarse cycle
        np->lc = np->nlev - 1;
        while (np->lc > 0 && np->L[np->lc - 1].tot <= kNdCoarseNodes) np->lc--;
        double *a = nullptr, *t = nullptr;
        if (!(e = nd_alloc(np, &a, (sizeof(NLev) * np->nlev + 7) / 8)) &&
            !(e = nd_alloc(np, &t, (sizeof(NdCoarseLev) * np->nlev + 7) / 8))) {
            std::vector<NdCoarseLev> lv(np->nlev);
            for (int l = 0; l < np->nlev; ++l) lv[l] = {np->cor[l], np->res[l], np->tmp[l]};
            np->dL = (NLev*)a;
            np->dlv = (NdCoarseLev*)t;
            if (hipMemcpy(np->dlv, lv.data(), sizeof(NdCoarseLev) * np->nlev, hipMemcpyHostToDevice) != hipSuccess) {
                set_error("hipMemcpy (nodal projector)");
                e = KSGF_ERR_HIP;
            }
            if (!e) e = nd_upload_levels(np);
        }
    }
    if (e) {
        ksgf_nodalproj_destroy(np);
        return e;
    }
    *out = np;
    return KSGF_OK;
}

int ksgf_nodalproj_set_sigma(ksgf_nodalproj* np, double sigma)
{
    if (!np || !(sigma > 0.0)) {
        set_error("ksgf_nodalproj_set_sigma: bad argument");
        return KSGF_ERR_INVALID;
    }
    np->sigma = sigma;
    nd_set_coefs(np);
    return nd_upload_levels(np);
}

int ksgf_nodalproj_node_dims(const ksgf_nodalproj* np, int* N)
{
    if (!np || !N) {
        set_error("ksgf_nodalproj_node_dims: null argument");
        return KSGF_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d) N[d] = np->L[0].N[d];
    return KSGF_OK;
}

int ksgf_nodalproj_compute_rhs(ksgf_nodalproj* np, const double* vel, int ng_vel, double* rhs, ksgf_stream_t stream)
{
    if (!np || !vel || !rhs || ng_vel < 1) {
        set_error("ksgf_nodalproj_compute_rhs: bad argument (vel needs one ghost layer)");
        return KSGF_ERR_INVALID;
    }
    return nd_compute_rhs(np, vel, ng_vel, rhs, (hipStream_t)stream);
}

int ksgf_nodalproj_solve(ksgf_nodalproj* np, const double* rhs, double* phi, double rtol, double atol,
                         ksgf_mlmg_info* info, ksgf_stream_t stream)
{
    if (!np || !rhs) {
        set_error("ksgf_nodalproj_solve: null argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    const int n = np->L[0].tot;
    KSGF_HIP_CHECK(hipMemcpyAsync(np->rhs, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    int ret = nd_solve(np, rtol, atol, info, st);
    if (ret < 0) return ret;
    if (phi) KSGF_HIP_CHECK(hipMemcpyAsync(phi, np->phi[np->cur], sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    return ret;
}

int ksgf_nodalproj_project(ksgf_nodalproj* np, double* vel, int ng_vel, double* gradphi, int ng_gp, double rtol,
                           double atol, ksgf_mlmg_info* info, ksgf_stream_t stream)
{
    if (!np || !vel || ng_vel < 1 || ng_gp < 0) {
        set_error("ksgf_nodalproj_project: bad argument (vel needs one ghost layer)");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    int e = nd_compute_rhs(np, vel, ng_vel, np->rhs, st);
    if (e) return e;
    int ret = nd_solve(np, rtol, atol, info, st);
    if (ret < 0) return ret;
    MFDev v, g{};
    if ((e = mfdev(np->ba, vel, KSGF_CELL, 3, ng_vel, &v))) return e;
    if (gradphi && (e = mfdev(np->ba, gradphi, KSGF_CELL, 3, ng_gp, &g))) return e;
    int err = KSGF_OK;
    const int z[3] = {0, 0, 0};
    const Region* r = get_region(np->ba, z, z, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_nd_correct, dim3(r->nchunks), dim3(kChunk), 0, st, v, g, gradphi ? 1 : 0,
                       np->phi[np->cur], nd_geom(np), np->sigma, regdev(r, np->ba->nbox), np->ba->dbox);
    KSGF_LAUNCH_CHECK();
    return ret;
}

int ksgf_nodalproj_get_phi(ksgf_nodalproj* np, double* phi, ksgf_stream_t stream)
{
    if (!np || !phi) {
        set_error("ksgf_nodalproj_get_phi: null argument");
        return KSGF_ERR_INVALID;
    }
    KSGF_HIP_CHECK(hipMemcpyAsync(phi, np->phi[np->cur], sizeof(double) * np->L[0].tot, hipMemcpyDeviceToDevice,
                                  (hipStream_t)stream));
    return KSGF_OK;
}

int ksgf_incflo_add_gradp(const ksgf_boxarray* ba, double* vel, int ng_vel, const double* gp, int ng_gp,
                          const double* rho, int ng_rho, double scaling_factor, ksgf_stream_t stream)
{
    if (!ba || !vel || !gp || !rho || ng_vel < 0 || ng_gp < 0 || ng_rho < 0) {
        set_error("ksgf_incflo_add_gradp: bad argument");
        return KSGF_ERR_INVALID;
    }
    MFDev v;
    CMFDev g, rr;
    int e;
    if ((e = mfdev(ba, vel, KSGF_CELL, 3, ng_vel, &v)) || (e = cmfdev(ba, gp, KSGF_CELL, 3, ng_gp, &g)) ||
        (e = cmfdev(ba, rho, KSGF_CELL, 1, ng_rho, &rr)))
        return e;
    int err = KSGF_OK;
    const int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_add_gradp, dim3(r->nchunks), dim3(kChunk), 0, (hipStream_t)stream, v, g, rr, scaling_factor,
                       regdev(r, ba->nbox), ba->dbox);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_set_bndry(const ksgf_boxarray* ba, double* mf, int itype, int ncomp, int ngrow, double val,
                   ksgf_stream_t stream)
{
    if (!ba || !mf || itype < 0 || itype > 3 || ncomp <= 0 || ngrow < 0) {
        set_error("ksgf_set_bndry: bad argument");
        return KSGF_ERR_INVALID;
    }
    if (ngrow == 0) return KSGF_OK;
    MFDev m;
    int e = mfdev(ba, mf, itype, ncomp, ngrow, &m);
    if (e) return e;
    int err = KSGF_OK;
    const Region* r = fab_region(ba, itype, ngrow, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_set_bndry, dim3(r->nchunks), dim3(kChunk), 0, (hipStream_t)stream, m, ncomp, val,
                       regdev(r, ba->nbox), ba->dbox);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

}  // extern "C"
