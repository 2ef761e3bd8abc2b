// Composite two-level MAC projection: Hydro::MacProjector over levels 0 and 1
// (icns_advection.cpp:88-95 builds it on Geom(0..finest), :331 setUMAC with
// every level's face velocities, :353 project), i.e. AMReX MLMG's composite
// solve of div(beta grad phi) = div(u*) on the coarse level outside the fine
// level and the fine level inside it, at refinement ratio 2.
//
// Discretisation (restated from AMReX's published method; its source is not
// part of the reference):
//  * a fine cell's neighbour outside the fine level and inside the domain is
//    a coarse-fine ghost: the quadratic through the coarse value b (located at
//    the coarse cell centre, one fine cell beyond the face) and the two fine
//    cells inside, 8/15 b + 2/3 u0 - 1/5 u1; b interpolates the coarse data
//    tangentially (centred slopes and curvature, one-sided slope where a
//    coarse neighbour is covered or outside, the cross term where the four
//    diagonal neighbours are usable);
//  * an uncovered coarse cell next to the fine level takes the average of the
//    four fine fluxes on each coarse-fine face (reflux), so the composite
//    operator conserves;
//  * covered coarse cells carry no unknown; they end at the average of their
//    children.
// Iteration: damped Jacobi (omega 6/7) on the fine cells (each MLMG
// smoothing step a pair of sweeps: 2 nu1 before, 2 nu2 after), a coarse
// correction of the composite residual (the fine residual averaged under the
// fine level) by one V-cycle of the single-level projector on the coarse
// grid, piecewise-constant prolongation to the fine cells, damped Jacobi
// again, until max |composite residual| <= max(atol, rtol r0).
//
// Data: coarse unknowns, right-hand side and residual are domain-wide arrays
// (the coarse level covers the domain; neighbours resolve by index
// arithmetic with periodic wrap); fine ones are MultiFabs of the fine
// BoxArray (phi with one ghost layer; the fine stencil passes evaluate a
// face ghost where they read it, from a per-ghost source code: another fine
// box's cell, a wall, or the coarse-fine quadratic with the coarse
// interpolation tabulated once per coarse correction; the Jacobi sweeps
// write the other of two phi buffers; the ghost layer itself is written once,
// for the final gradient).  Every expression is evaluated in the operand order of
// the CPU oracle (oracle/ksgf_oracle.c, composite two-level section).
#include <algorithm>
#include <cstdint>
#include <cstring>

#include "ksgf_internal.h"

using namespace ksgf;

namespace {

constexpr double kOmega = 6.0 / 7.0;

struct TLGeom {
    int n[3], nf[3], per[3], lobc[3], hibc[3];
    double hc[3], hf[3], ihc[3], ihf[3], beta;  // spacings and their inverses
};

__device__ __forceinline__ long long tl_ci(const TLGeom& T, int i, int j, int k)
{
    return i + (long long)T.n[0] * (j + (long long)T.n[1] * k);
}
__device__ __forceinline__ long long tl_fi(const TLGeom& T, int i, int j, int k)
{
    return i + (long long)T.nf[0] * (j + (long long)T.nf[1] * k);
}
__device__ __forceinline__ int wrapi(int x, int n) { return ((x % n) + n) % n; }

// coarse value at Q (wrapped); ok = 0 where covered or outside the domain
__device__ __forceinline__ double tl_cval(const TLGeom& T, const double* pc, const unsigned char* cv, const int* Q,
                                          int& ok)
{
    int q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q[d] = Q[d];
        if (q[d] < 0 || q[d] >= T.n[d]) {
            if (!T.per[d]) {
                ok = 0;
                return 0.0;
            }
            q[d] = wrapi(q[d], T.n[d]);
        }
    }
    const long long c = tl_ci(T, q[0], q[1], q[2]);
    if (cv[c]) {
        ok = 0;
        return 0.0;
    }
    ok = 1;
    return pc[c];
}

__device__ double tl_cinterp(const TLGeom& T, const double* pc, const unsigned char* cv, const int* Q, int dn,
                             const double* x)
{
    int ok;
    const double uc = tl_cval(T, pc, cv, Q, ok);
    const int t1 = (dn == 0) ? 1 : 0, t2 = (dn == 2) ? 1 : 2;
    const int tt[2] = {t1, t2};
    double term[2];
    for (int a = 0; a < 2; ++a) {
        const int t = tt[a];
        int P[3] = {Q[0], Q[1], Q[2]}, M[3] = {Q[0], Q[1], Q[2]}, okp, okm;
        P[t] += 1;
        M[t] -= 1;
        const double up = tl_cval(T, pc, cv, P, okp), um = tl_cval(T, pc, cv, M, okm);
        double d1 = 0.0, d2 = 0.0;
        if (okp && okm) {
            d1 = 0.5 * (up - um);
            d2 = 0.5 * ((up - 2.0 * uc) + um);
        } else if (okp) {
            d1 = up - uc;
        } else if (okm) {
            d1 = uc - um;
        }
        term[a] = x[t] * (d1 + x[t] * d2);
    }
    double dxy = 0.0;
    {
        double v[4];
        int okc[4];
        for (int m = 0; m < 4; ++m) {
            int C[3] = {Q[0], Q[1], Q[2]};
            C[t1] += (m & 1) ? 1 : -1;
            C[t2] += (m & 2) ? 1 : -1;
            v[m] = tl_cval(T, pc, cv, C, okc[m]);
        }
        if (okc[0] && okc[1] && okc[2] && okc[3]) dxy = 0.25 * ((v[3] - v[2]) - (v[1] - v[0]));
    }
    return ((uc + term[0]) + term[1]) + x[t1] * x[t2] * dxy;
}

// coarse covered mask from the fine owner map (fine boxes are aligned to 2)
__global__ void __launch_bounds__(256) k_tl_cover(const int* __restrict__ fown, unsigned char* __restrict__ cv,
                                                  TLGeom T)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = (long long)T.n[0] * T.n[1] * T.n[2];
    if (t >= nc) return;
    const int i = (int)(t % T.n[0]), j = (int)((t / T.n[0]) % T.n[1]), k = (int)(t / ((long long)T.n[0] * T.n[1]));
    cv[t] = fown[tl_fi(T, 2 * i, 2 * j, 2 * k)] >= 0 ? 1 : 0;
}

// A face ghost of a fine FAB (edges and corners are not read by the 7-point
// stencil) is another fine box's cell, a physical-boundary ghost, or a
// coarse-fine ghost.  The ghost list holds, for every fine box, the six
// one-cell slabs outside its faces, x fastest within a slab, in the order
// (x lo, x hi, y lo, y hi, z lo, z hi); an entry is (fine box, linear index
// of the ghost in its FAB storage of ngrow 1).
struct TLGhost {
    int b, d, s;  // box, normal direction, side (-1 low, +1 high)
    int q[3];     // ghost position
};
__device__ __forceinline__ TLGhost tl_ghost_at(const int* __restrict__ fbox, int2 ent)
{
    TLGhost g;
    g.b = ent.x;
    const int* bx = fbox + 6 * g.b;
    const int n0 = bx[3] - bx[0] + 3, n1 = bx[4] - bx[1] + 3;
    g.q[0] = bx[0] - 1 + ent.y % n0;
    g.q[1] = bx[1] - 1 + (ent.y / n0) % n1;
    g.q[2] = bx[2] - 1 + ent.y / (n0 * n1);
    g.d = 0;
    g.s = 0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        if (g.q[e] < bx[e]) {
            g.d = e;
            g.s = -1;
        } else if (g.q[e] > bx[3 + e]) {
            g.d = e;
            g.s = 1;
        }
    }
    return g;
}

// the coarse interpolation b of a coarse-fine ghost at fine position w
// (inside the domain, wrapped), normal direction d
__device__ __forceinline__ double tl_cf_bq(const TLGeom& T, const double* pc, const unsigned char* cv, const int* w,
                                           int d)
{
    int Q[3];
    double x[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        Q[e] = w[e] / 2;
        x[e] = ((double)(w[e] - 2 * Q[e]) + 0.5) * 0.5 - 0.5;
    }
    x[d] = 0.0;
    return tl_cinterp(T, pc, cv, Q, d, x);
}

// the value of a coarse-fine ghost from b and the two fine cells inside:
// u0 next to the face, u1 one further in
__device__ __forceinline__ double tl_cf_ghost(double bq, double u0, double u1)
{
    return ((8.0 / 15.0) * bq + (2.0 / 3.0) * u0) - 0.2 * u1;
}

// One thread per entry of the ghost list: fills the ghost in the FABs of pf.
__global__ void __launch_bounds__(256) k_tl_ghosts(MFDev pf, const int* __restrict__ fbox,
                                                   const int2* __restrict__ gl, long long ngl,
                                                   const int* __restrict__ fown, const double* __restrict__ pc,
                                                   const unsigned char* __restrict__ cv, TLGeom T)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ngl) return;
    const TLGhost g = tl_ghost_at(fbox, gl[t]);
    const int d = g.d, s = g.s;
    Fab f = fab_at(pf, fbox, g.b);
    int i0[3] = {g.q[0], g.q[1], g.q[2]};
    i0[d] -= s;
    const double u0 = f(i0[0], i0[1], i0[2]);
    double v;
    int w[3] = {g.q[0], g.q[1], g.q[2]};
    if (w[d] < 0 || w[d] >= T.nf[d]) {
        if (!T.per[d]) {
            const int bc = (s < 0) ? T.lobc[d] : T.hibc[d];
            f(g.q[0], g.q[1], g.q[2]) = (bc == KSGF_LO_DIRICHLET) ? -u0 : u0;
            return;
        }
        w[d] = wrapi(w[d], T.nf[d]);
    }
    const int o = fown[tl_fi(T, w[0], w[1], w[2])];
    if (o >= 0) {
        Fab h = fab_at(pf, fbox, o);
        v = h(w[0], w[1], w[2]);
    } else {
        int i1[3] = {i0[0], i0[1], i0[2]};
        i1[d] -= s;
        v = tl_cf_ghost(tl_cf_bq(T, pc, cv, w, d), u0, f(i1[0], i1[1], i1[2]));
    }
    f(g.q[0], g.q[1], g.q[2]) = v;
}

// Ghost source codes, one per entry of the ghost list, so the stencil passes
// evaluate the ghosts where they read them: code >= 0 is the storage offset
// of the source cell (another fine box's) in the fine phi MultiFab (the two
// phi buffers share the layout); kGhNeumann and kGhDirichlet are physical
// walls (ghost u0 or -u0); kGhCF is a coarse-fine ghost, whose coarse
// interpolation the table bq holds at the same entry.
constexpr long long kGhNeumann = -1, kGhDirichlet = -2, kGhCF = -3;

__global__ void __launch_bounds__(256) k_tl_gsrc(CMFDev pf, const int* __restrict__ fbox,
                                                 const int2* __restrict__ gl, long long ngl,
                                                 const int* __restrict__ fown, TLGeom T, long long* __restrict__ code)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ngl) return;
    const TLGhost g = tl_ghost_at(fbox, gl[t]);
    const int d = g.d, s = g.s;
    int w[3] = {g.q[0], g.q[1], g.q[2]};
    if (w[d] < 0 || w[d] >= T.nf[d]) {
        if (!T.per[d]) {
            const int bc = (s < 0) ? T.lobc[d] : T.hibc[d];
            code[t] = (bc == KSGF_LO_DIRICHLET) ? kGhDirichlet : kGhNeumann;
            return;
        }
        w[d] = wrapi(w[d], T.nf[d]);
    }
    const int o = fown[tl_fi(T, w[0], w[1], w[2])];
    if (o >= 0) {
        const CFab h = fab_at(pf, fbox, o);
        code[t] = pf.pre[o] + h.idx(w[0], w[1], w[2]);
    } else {
        code[t] = kGhCF;
    }
}

// the coarse interpolations of the coarse-fine ghosts (entries cfl of the
// ghost list) from the current coarse solution, stored at their entries
__global__ void __launch_bounds__(256) k_tl_bq(const int* __restrict__ fbox, const int2* __restrict__ gl,
                                               const long long* __restrict__ cfl, long long ncf,
                                               const double* __restrict__ pc, const unsigned char* __restrict__ cv,
                                               TLGeom T, double* __restrict__ bq)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncf) return;
    const long long e = cfl[t];
    const TLGhost g = tl_ghost_at(fbox, gl[e]);
    int w[3] = {g.q[0], g.q[1], g.q[2]};
    w[g.d] = wrapi(w[g.d], T.nf[g.d]);
    bq[e] = tl_cf_bq(T, pc, cv, w, g.d);
}

// One ghost slab (a face of a fine box), classified on the host from the
// codes of its entries.  kFaceAffine: every ghost is a cell of one other
// fine box, at offset off + a sa + b sb for slab coordinates (a, b) (the two
// tangential directions in x, y, z order); the other kinds hold for the
// whole slab (walls, coarse-fine) or fall back to the codes (mixed).  Non-
// affine slabs keep off = sa = sb = 0, so an unconditional affine load reads
// the buffer's first value.
enum { kFaceAffine = 0, kFaceNeumann = 1, kFaceDirichlet = 2, kFaceCF = 3, kFaceMixed = 4 };
struct TLFace {
    long long off;
    int sa, sb, kind, pad;
};

struct GhDev {
    const long long* code;  // ghost source code of every ghost-list entry
    const long long* off;   // first ghost-list entry of each fine box
    const TLFace* face;     // six slabs per fine box
    const double* bq;       // coarse interpolations at the coarse-fine entries
};

// a ghost from its code (entry t): u0 is the fine cell next to the face, u1
// the one further in
__device__ __forceinline__ double tl_gval(long long code, long long t, const double* __restrict__ base, double u0,
                                          double u1, const double* __restrict__ bq)
{
    if (code >= 0) return base[code];
    if (code == kGhNeumann) return u0;
    if (code == kGhDirichlet) return -u0;
    return tl_cf_ghost(bq[t], u0, u1);
}

// a ghost of a non-affine slab F (entry t)
__device__ __forceinline__ double tl_fghost(const TLFace& F, long long t, const GhDev& gh,
                                            const double* __restrict__ base, double u0, double u1)
{
    if (F.kind == kFaceNeumann) return u0;
    if (F.kind == kFaceDirichlet) return -u0;
    if (F.kind == kFaceCF) return tl_cf_ghost(gh.bq[t], u0, u1);
    return tl_gval(gh.code[t], t, base, u0, u1, gh.bq);
}

// a ghost of slab F at slab coordinates (a, b), entry t
__device__ __forceinline__ double tl_face_ghost(const TLFace& F, int a, int b, long long t, const GhDev& gh,
                                                const double* __restrict__ base, double u0, double u1)
{
    if (F.kind == kFaceAffine) return base[F.off + a * (long long)F.sa + b * (long long)F.sb];
    return tl_fghost(F, t, gh, base, u0, u1);
}

// The fine stencil passes: one thread per segment of kTlKz cells of one
// (i, j) column of a fine box, walking up in z with the cell and its z
// neighbours in registers, its FABs located once.  A neighbour outside the
// box is evaluated where it is read, from its slab's descriptor or its ghost
// code (another box's cell of the input buffer, a wall, or the coarse-fine
// quadratic through the input's cells), so no ghost fill precedes a pass.  The residual is R = b - div(beta grad
// phi) (the oracle's operand order).  SMOOTH: one damped-Jacobi sweep,
// phi + w R, into the other phi buffer; otherwise R into rf.
constexpr int kTlKz = 16;

struct ColDev {
    const int* chunks;  // (box, first column-segment) per 256 threads
};

template <bool SMOOTH>
__global__ void __launch_bounds__(256) k_tl_march(CMFDev pin, CMFDev bf, MFDev out, const int* __restrict__ fbox,
                                                  TLGeom T, double w, ColDev cd, GhDev gh)
{
    const int b = cd.chunks[2 * blockIdx.x];
    const unsigned loc = (unsigned)cd.chunks[2 * blockIdx.x + 1] + threadIdx.x;
    const int* bx = fbox + 6 * b;
    const unsigned nx = (unsigned)(bx[3] - bx[0] + 1), ny = (unsigned)(bx[4] - bx[1] + 1);
    const int nz = bx[5] - bx[2] + 1, nseg = (nz + kTlKz - 1) / kTlKz;
    if (loc >= nx * ny * (unsigned)nseg) return;
    const unsigned q = loc / nx;
    const int ii = (int)(loc - q * nx), jj = (int)(q % ny), sg = (int)(q / ny);
    const int i = bx[0] + ii, j = bx[1] + jj;
    const int k0 = bx[2] + sg * kTlKz, k1 = min(k0 + kTlKz, bx[5] + 1);
    const CFab p = fab_at(pin, fbox, b);
    const CFab rb = fab_at(bf, fbox, b);
    const Fab o = fab_at(out, fbox, b);
    // input, right-hand side and output never overlap: restrict-qualified
    // bases let the compiler issue the loads of later planes early
    const double* __restrict__ pp = p.p;
    const double* __restrict__ bp = rb.p;
    double* __restrict__ op = o.p;
    const double* __restrict__ base = pin.p;
    const int sy = p.n0, sz = p.n0 * p.n1, sbz = rb.n0 * rb.n1, soz = o.n0 * o.n1;
    // this column's slabs and entries: x lo/hi (slab coordinates (j, k)),
    // y lo/hi ((i, k)), z lo/hi ((i, j)); boxes have at least two cells per
    // direction, so a lane is on at most one x and one y face
    const int nyz = (int)ny * nz, nxz = (int)nx * nz;
    const long long g0 = gh.off[b];
    const bool bw = ii == 0, be = ii == (int)nx - 1, bs = jj == 0, bn = jj == (int)ny - 1;
    const TLFace FX = gh.face[6 * b + (be ? 1 : 0)];
    const double* __restrict__ xg = base + FX.off + jj * (long long)FX.sa + (k0 - bx[2]) * (long long)FX.sb;
    const long long tx = g0 + (be ? nyz : 0) + jj;                // + ny (k - z0)
    const long long ty = g0 + 2 * nyz + (bn ? nxz : 0) + ii;      // + nx (k - z0)
    const long long tz = g0 + 2 * (nyz + nxz) + ii + (int)nx * jj;  // + nx ny on the top
    // the stencil loads the FAB's ghost slots (in bounds) everywhere and
    // replaces them by the ghost's value on the box faces
    int c = p.idx(i, j, k0), cb = rb.idx(i, j, k0), co = o.idx(i, j, k0);
    double u0 = pp[c];
    double um = pp[c - sz];
    if (k0 == bx[2]) um = tl_face_ghost(gh.face[6 * b + 4], ii, jj, tz, gh, base, u0, pp[c + sz]);
    const int len = k1 - k0;
    int kr = k0 - bx[2];
    // not unrolled: 60 VGPRs (8 waves per SIMD); unrolled by 2 or 4 the
    // face loads take 88 / 111 VGPRs and it is no faster or slower
#pragma unroll 1
    for (int t = 0; t < len; ++t, ++kr) {
        double up = pp[c + sz];
        if (kr == nz - 1) up = tl_face_ghost(gh.face[6 * b + 5], ii, jj, tz + (int)nx * ny, gh, base, u0, um);
        double e = pp[c + 1], wv = pp[c - 1];
        // x faces are in every wave (a wave is a run of x): the affine
        // ghost is loaded without a branch (lanes off the face read the
        // same address as the face lane of their slab)
        {
            double gv = *xg;
            xg += FX.sb;
            if ((bw | be) && FX.kind != kFaceAffine)
                gv = tl_fghost(FX, tx + (long long)ny * kr, gh, base, u0, bw ? e : wv);
            if (bw) wv = gv;
            if (be) e = gv;
        }
        double nv = pp[c + sy], sv = pp[c - sy];
        if (bs | bn) {
            const double gv = tl_face_ghost(gh.face[6 * b + (bn ? 3 : 2)], ii, kr, ty + (long long)nx * kr, gh, base,
                                            u0, bs ? nv : sv);
            if (bs) sv = gv;
            if (bn) nv = gv;
        }
        double acc = 0.0;
        {
            const double Fp = T.beta * (e - u0) * T.ihf[0];
            const double Fm = T.beta * (u0 - wv) * T.ihf[0];
            acc += (Fp - Fm) * T.ihf[0];
        }
        {
            const double Fp = T.beta * (nv - u0) * T.ihf[1];
            const double Fm = T.beta * (u0 - sv) * T.ihf[1];
            acc += (Fp - Fm) * T.ihf[1];
        }
        {
            const double Fp = T.beta * (up - u0) * T.ihf[2];
            const double Fm = T.beta * (u0 - um) * T.ihf[2];
            acc += (Fp - Fm) * T.ihf[2];
        }
        const double R = bp[cb] - acc;
        op[co] = SMOOTH ? u0 + w * R : R;
        um = u0;
        u0 = up;
        c += sz;
        cb += sbz;
        co += soz;
    }
}

// position in the ghost list of box bx's ghost at (i, j, k), outside the box
// in direction d on the high (hi != 0) or low side
__device__ __forceinline__ long long tl_gidx(const int* bx, int d, int hi, int i, int j, int k)
{
    const int nx = bx[3] - bx[0] + 1, ny = bx[4] - bx[1] + 1, nz = bx[5] - bx[2] + 1;
    const long long nyz = (long long)ny * nz, nxz = (long long)nx * nz;
    if (d == 0) return (hi ? nyz : 0) + (j - bx[1]) + (long long)ny * (k - bx[2]);
    if (d == 1) return 2 * nyz + (hi ? nxz : 0) + (i - bx[0]) + (long long)nx * (k - bx[2]);
    return 2 * (nyz + nxz) + (hi ? (long long)nx * ny : 0) + (i - bx[0]) + (long long)nx * (j - bx[1]);
}

// composite residual on the coarse cells: uncovered R = b - Lc (averaged fine
// fluxes on coarse-fine faces), covered 0
__global__ void __launch_bounds__(256) k_tl_cres(const double* __restrict__ pc, const double* __restrict__ bc,
                                                 double* __restrict__ rc, const unsigned char* __restrict__ cv,
                                                 CMFDev pf, const int* __restrict__ fbox, const int* __restrict__ fown,
                                                 TLGeom T, GhDev gh)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = (long long)T.n[0] * T.n[1] * T.n[2];
    if (t >= nc) return;
    if (cv[t]) {
        rc[t] = 0.0;
        return;
    }
    const int C[3] = {(int)(t % T.n[0]), (int)((t / T.n[0]) % T.n[1]), (int)(t / ((long long)T.n[0] * T.n[1]))};
    const double uc = pc[t];
    double acc = 0.0;
    for (int d = 0; d < 3; ++d) {
        double F[2];
        for (int a = 0; a < 2; ++a) {
            const int s = a ? 1 : -1;
            int N[3] = {C[0], C[1], C[2]};
            N[d] += s;
            if (N[d] < 0 || N[d] >= T.n[d]) {
                if (!T.per[d]) {
                    const int bcd = (s < 0) ? T.lobc[d] : T.hibc[d];
                    F[a] = (bcd == KSGF_LO_DIRICHLET)
                               ? ((s > 0) ? T.beta * (-uc - uc) * T.ihc[d] : T.beta * (uc + uc) * T.ihc[d])
                               : 0.0;
                    continue;
                }
                N[d] = wrapi(N[d], T.n[d]);
            }
            const long long nci = tl_ci(T, N[0], N[1], N[2]);
            if (cv[nci]) {
                const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                double ff[4];
                for (int m = 0; m < 4; ++m) {
                    int fi[3];
                    fi[d] = (s > 0) ? 2 * N[d] : 2 * N[d] + 1;
                    fi[t1] = 2 * N[t1] + (m & 1);
                    fi[t2] = 2 * N[t2] + ((m >> 1) & 1);
                    // the flux through fi's face towards C: its ghost there is
                    // a coarse-fine ghost, evaluated from its code
                    const int o = fown[tl_fi(T, fi[0], fi[1], fi[2])];
                    const CFab pb = fab_at(pf, fbox, o);
                    int fn[3] = {fi[0], fi[1], fi[2]};
                    fn[d] += s;
                    const double u0 = pb(fi[0], fi[1], fi[2]);
                    const long long ge = gh.off[o] + tl_gidx(fbox + 6 * o, d, s < 0, fi[0], fi[1], fi[2]);
                    const double gv = tl_gval(gh.code[ge], ge, pf.p, u0, pb(fn[0], fn[1], fn[2]), gh.bq);
                    ff[m] = (s < 0) ? T.beta * (gv - u0) * T.ihf[d] : T.beta * (u0 - gv) * T.ihf[d];
                }
                F[a] = 0.25 * ((ff[0] + ff[1]) + (ff[2] + ff[3]));
            } else {
                F[a] = (s > 0) ? T.beta * (pc[nci] - uc) * T.ihc[d] : T.beta * (uc - pc[nci]) * T.ihc[d];
            }
        }
        acc += (F[1] - F[0]) * T.ihc[d];
    }
    rc[t] = bc[t] - acc;
}

// the average of the 8 children of every covered coarse cell: the coarse
// correction's right-hand side (from the fine residual) or the final coarse
// solution (from the fine solution)
__global__ void __launch_bounds__(256) k_tl_avg8(double* __restrict__ out, const unsigned char* __restrict__ cv,
                                                 CMFDev fm, const int* __restrict__ fbox, const int* __restrict__ fown,
                                                 TLGeom T)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = (long long)T.n[0] * T.n[1] * T.n[2];
    if (t >= nc || !cv[t]) return;
    const int i = (int)(t % T.n[0]), j = (int)((t / T.n[0]) % T.n[1]), k = (int)(t / ((long long)T.n[0] * T.n[1]));
    const int o = fown[tl_fi(T, 2 * i, 2 * j, 2 * k)];
    CFab f = fab_at(fm, fbox, o);
    double s8 = 0.0;
    for (int kk = 0; kk < 2; ++kk)
        for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) s8 += f(2 * i + ii, 2 * j + jj, 2 * k + kk);
    out[t] = 0.125 * s8;
}

// coarse correction: pc += ec everywhere, fine cells += ec of the parent
__global__ void __launch_bounds__(256) k_tl_addc(double* __restrict__ pc, const double* __restrict__ ec, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) pc[t] += ec[t];
}
__global__ void __launch_bounds__(256) k_tl_prolong(MFDev pf, const double* __restrict__ ec,
                                                    const int* __restrict__ fbox, TLGeom T, RegDev r)
{
    int b, i, j, k;
    if (!region_locate(r, fbox, b, i, j, k)) return;
    fab_at(pf, fbox, b)(i, j, k) += ec[tl_ci(T, i / 2, j / 2, k / 2)];
}

// div u over the valid cells: fine into a MultiFab, coarse into a
// domain-wide array (covered cells 0)
__global__ void __launch_bounds__(256) k_tl_div_f(MFDev out, CMFDev ux, CMFDev uy, CMFDev uz,
                                                  const int* __restrict__ box, double hx, double hy, double hz,
                                                  RegDev r)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab a = fab_at(ux, box, b), e = fab_at(uy, box, b), g = fab_at(uz, box, b);
    const double dx = (a(i + 1, j, k) - a(i, j, k)) / hx;
    const double dy = (e(i, j + 1, k) - e(i, j, k)) / hy;
    const double dz = (g(i, j, k + 1) - g(i, j, k)) / hz;
    fab_at(out, box, b)(i, j, k) = (dx + dy) + dz;
}
__global__ void __launch_bounds__(256) k_tl_div_c(double* __restrict__ out, const unsigned char* __restrict__ cv,
                                                  CMFDev ux, CMFDev uy, CMFDev uz, const int* __restrict__ box,
                                                  TLGeom T, RegDev r)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    const long long c = tl_ci(T, i, j, k);
    if (cv[c]) {
        out[c] = 0.0;
        return;
    }
    CFab a = fab_at(ux, box, b), e = fab_at(uy, box, b), g = fab_at(uz, box, b);
    const double dx = (a(i + 1, j, k) - a(i, j, k)) / T.hc[0];
    const double dy = (e(i, j + 1, k) - e(i, j, k)) / T.hc[1];
    const double dz = (g(i, j, k + 1) - g(i, j, k)) / T.hc[2];
    out[c] = (dx + dy) + dz;
}

// u -= beta grad phi on the valid faces of direction D: fine from the FAB
// (ghosts filled), coarse from the domain-wide solution with the domain BCs
template <int D>
__global__ void __launch_bounds__(256) k_tl_fix_f(MFDev u, CMFDev pf, const int* __restrict__ box, TLGeom T, RegDev r)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab p = fab_at(pf, box, b);
    const double hi = p(i, j, k);
    const double lo = p(i - (D == 0), j - (D == 1), k - (D == 2));
    fab_at(u, box, b)(i, j, k) -= T.beta * (hi - lo) * T.ihf[D];
}
template <int D>
__global__ void __launch_bounds__(256) k_tl_fix_c(MFDev u, const double* __restrict__ pc, const int* __restrict__ box,
                                                  TLGeom T, RegDev r)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
    lo[D] -= 1;
    double F;
    if (lo[D] < 0 || hi[D] >= T.n[D]) {
        if (!T.per[D]) {
            const bool low = lo[D] < 0;
            const double u0 = low ? pc[tl_ci(T, hi[0], hi[1], hi[2])] : pc[tl_ci(T, lo[0], lo[1], lo[2])];
            const int bcd = low ? T.lobc[D] : T.hibc[D];
            F = (bcd == KSGF_LO_DIRICHLET) ? (low ? T.beta * (u0 + u0) * T.ihc[D] : T.beta * (-u0 - u0) * T.ihc[D])
                                           : 0.0;
            fab_at(u, box, b)(i, j, k) -= F;
            return;
        }
        if (lo[D] < 0) lo[D] += T.n[D];
        if (hi[D] >= T.n[D]) hi[D] -= T.n[D];
    }
    F = T.beta * (pc[tl_ci(T, hi[0], hi[1], hi[2])] - pc[tl_ci(T, lo[0], lo[1], lo[2])]) * T.ihc[D];
    fab_at(u, box, b)(i, j, k) -= F;
}

// max |a| (non-negative doubles order as their bit patterns)
__global__ void __launch_bounds__(256) k_tl_absmax(const double* __restrict__ a, long long n,
                                                   unsigned long long* __restrict__ out)
{
    __shared__ double sm[256];
    double m = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const double v = fabs(a[t]);
        m = (v > m || v != v) ? v : m;
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = sm[threadIdx.x + w];
            if (o > sm[threadIdx.x] || o != o) sm[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(sm[0]));
}

// block partial sums for the composite mean (coarse over uncovered cells,
// fine over the valid cells of a MultiFab), one slot per block, summed in
// block order by k_tl_mean: the mean is the same bits on every run
__global__ void __launch_bounds__(256) k_tl_sum_c(const double* __restrict__ a, const unsigned char* __restrict__ cv,
                                                  long long n, double* __restrict__ out)
{
    __shared__ double sm[256];
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        if (!cv[t]) s += a[t];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}
__global__ void __launch_bounds__(256) k_tl_sum_f(CMFDev f, const int* __restrict__ box, RegDev r, int nchunks,
                                                  double* __restrict__ out)
{
    __shared__ double sm[256];
    double s = 0.0;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int b = r.chunks[2 * c];
        int i, j, k;
        if (decode_in_box(r, box, b, (unsigned)r.chunks[2 * c + 1] + threadIdx.x, i, j, k))
            s += fab_at(f, box, b)(i, j, k);
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}
__global__ void __launch_bounds__(256) k_tl_shift_c(double* __restrict__ a, const unsigned char* __restrict__ cv,
                                                    long long n, const double* __restrict__ mu, int all)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && (all || !cv[t])) a[t] -= *mu;
}
__global__ void __launch_bounds__(256) k_tl_shift_f(MFDev f, const int* __restrict__ box, RegDev r,
                                                    const double* __restrict__ mu)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    fab_at(f, box, b)(i, j, k) -= *mu;
}
constexpr int kTlSumC = 1024, kTlSumF = 2048;  // blocks of the composite-mean partial sums

__global__ void k_tl_mean(double* __restrict__ s, const double* __restrict__ pc, int nbc,
                          const double* __restrict__ pf, int nbf, double wc, double wf, double den)
{
    // s[0]: coarse sum, s[1]: fine sum (block partials in block order) ->
    // s[2] = (8 sc + sf) / den
    double sc = 0.0, sf = 0.0;
    for (int b = 0; b < nbc; ++b) sc += pc[b];
    for (int b = 0; b < nbf; ++b) sf += pf[b];
    s[0] = sc;
    s[1] = sf;
    s[2] = (wc * sc + wf * sf) / den;
}

}  // namespace

struct ksgf_macproj_levels {
    const ksgf_boxarray* cba = nullptr;
    const ksgf_boxarray* fba = nullptr;
    TLGeom T{};
    ksgf_mlmg_options opt{};
    int singular = 1;
    ksgf_macproj* crse = nullptr;  // single-level projector on the coarse grid: one V-cycle per correction
    const int* fown = nullptr;     // fine owner map (owned by the fine BoxArray)
    unsigned char* cv = nullptr;
    double *pc = nullptr, *bc = nullptr, *rc = nullptr, *ec = nullptr;
    double* pf[2] = {nullptr, nullptr};  // fine phi (ngrow 1), ping-pong of the Jacobi sweeps
    int cur = 0;                         // the buffer holding the current iterate
    double *bf = nullptr, *rf = nullptr;
    int2* gl = nullptr;  // face-ghost list of the fine FABs
    long long ngl = 0;
    long long* gcode = nullptr;  // ghost source code of every ghost-list entry
    long long* goff = nullptr;   // first ghost-list entry of each fine box
    TLFace* face = nullptr;      // six slab descriptors per fine box
    long long* cfl = nullptr;    // ghost-list entries of the coarse-fine ghosts
    double* bq = nullptr;        // their coarse interpolations (current coarse solution), by entry
    long long ncf = 0;
    int* cols = nullptr;  // column-segment launch table of the fine boxes (k_tl_march)
    int ncols_chunks = 0;
    long long nc = 0, nfg = 0, nf0 = 0;
    long long ncov = 0, nfine = 0;
    unsigned long long* dmax = nullptr;  // 2 slots
    double* dsum = nullptr;              // 3 slots + block partials (coarse, fine)
    unsigned long long* hmax = nullptr;  // pinned
};

namespace {

int tl_mfs(ksgf_macproj_levels* m, MFDev* pf, CMFDev* cpf, CMFDev* bf, MFDev* rf, CMFDev* crf)
{
    int e;
    for (int a = 0; a < 2; ++a)
        if ((e = mfdev(m->fba, m->pf[a], KSGF_CELL, 1, 1, &pf[a])) ||
            (e = cmfdev(m->fba, m->pf[a], KSGF_CELL, 1, 1, &cpf[a])))
            return e;
    if ((e = cmfdev(m->fba, m->bf, KSGF_CELL, 1, 0, bf)) || (e = mfdev(m->fba, m->rf, KSGF_CELL, 1, 0, rf)) ||
        (e = cmfdev(m->fba, m->rf, KSGF_CELL, 1, 0, crf)))
        return e;
    return KSGF_OK;
}

struct TLRun {
    ksgf_macproj_levels* m;
    hipStream_t st;
    MFDev pfm[2];
    CMFDev cpfm[2], bf, crf;
    MFDev rf;
    RegDev rv;  // fine valid cells
    int nv_chunks;

    const MFDev& pf() const { return pfm[m->cur]; }
    const CMFDev& cpf() const { return cpfm[m->cur]; }
    GhDev gh() const { return GhDev{m->gcode, m->goff, m->face, m->bq}; }
    // the face ghosts of the current phi buffer, written into its FABs (the
    // stencil passes evaluate them where they read them; this is for the
    // final gradient and the returned solution)
    void ghosts()
    {
        if (m->ngl == 0) return;
        hipLaunchKernelGGL(k_tl_ghosts, dim3(grid_for(m->ngl, 256)), dim3(256), 0, st, pf(), m->fba->dbox,
                           (const int2*)m->gl, m->ngl, m->fown, (const double*)m->pc, (const unsigned char*)m->cv,
                           m->T);
    }
    // the coarse interpolations of the coarse-fine ghosts, after every change
    // of the coarse solution
    void coarse_changed()
    {
        if (m->ncf == 0) return;
        hipLaunchKernelGGL(k_tl_bq, dim3(grid_for(m->ncf, 256)), dim3(256), 0, st, m->fba->dbox, (const int2*)m->gl,
                           (const long long*)m->cfl, m->ncf, (const double*)m->pc, (const unsigned char*)m->cv, m->T,
                           m->bq);
    }
    void fres()
    {
        hipLaunchKernelGGL(k_tl_march<false>, dim3(m->ncols_chunks), dim3(256), 0, st, cpf(), bf, rf, m->fba->dbox,
                           m->T, 0.0, ColDev{m->cols}, gh());
    }
    void cres()
    {
        hipLaunchKernelGGL(k_tl_cres, dim3(grid_for(m->nc, 256)), dim3(256), 0, st, (const double*)m->pc,
                           (const double*)m->bc, m->rc, (const unsigned char*)m->cv, cpf(), m->fba->dbox, m->fown,
                           m->T, gh());
    }
    // composite residual on both levels
    void residual()
    {
        fres();
        cres();
    }
    void smooth(int nsweep, double w)
    {
        for (int s = 0; s < nsweep; ++s) {
            hipLaunchKernelGGL(k_tl_march<true>, dim3(m->ncols_chunks), dim3(256), 0, st, cpf(), bf, pfm[1 - m->cur],
                               m->fba->dbox, m->T, w, ColDev{m->cols}, gh());
            m->cur ^= 1;
        }
    }
    int norm(double* out)
    {
        KSGF_HIP_CHECK(hipMemsetAsync(m->dmax, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_tl_absmax, dim3(std::min(1024, grid_for(m->nc, 256))), dim3(256), 0, st,
                           (const double*)m->rc, m->nc, m->dmax);
        hipLaunchKernelGGL(k_tl_absmax, dim3(std::min(1024, grid_for(m->nf0, 256))), dim3(256), 0, st,
                           (const double*)m->rf, m->nf0, m->dmax);
        KSGF_LAUNCH_CHECK();
        KSGF_HIP_CHECK(hipMemcpyAsync(m->hmax, m->dmax, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        KSGF_HIP_CHECK(hipStreamSynchronize(st));
        unsigned long long bits = *m->hmax;
        double v;
        memcpy(&v, &bits, sizeof(v));
        *out = v;
        return KSGF_OK;
    }
    // subtract the composite mean of (coarse a, fine f); all_coarse: also
    // from the covered coarse cells
    int sub_mean(double* a, MFDev f, CMFDev cf, int all_coarse)
    {
        const int nbc = std::min(kTlSumC, grid_for(m->nc, 256)), nbf = std::min(nv_chunks, kTlSumF);
        double* pc = m->dsum + 3;
        double* pf = pc + kTlSumC;
        hipLaunchKernelGGL(k_tl_sum_c, dim3(nbc), dim3(256), 0, st, (const double*)a, (const unsigned char*)m->cv,
                           m->nc, pc);
        hipLaunchKernelGGL(k_tl_sum_f, dim3(nbf), dim3(256), 0, st, cf, m->fba->dbox, rv, nv_chunks, pf);
        const double den = 8.0 * (double)(m->nc - m->ncov) + (double)m->nfine;
        hipLaunchKernelGGL(k_tl_mean, dim3(1), dim3(1), 0, st, m->dsum, (const double*)pc, nbc, (const double*)pf,
                           nbf, 8.0, 1.0, den);
        hipLaunchKernelGGL(k_tl_shift_c, dim3(grid_for(m->nc, 256)), dim3(256), 0, st, a, (const unsigned char*)m->cv,
                           m->nc, (const double*)(m->dsum + 2), all_coarse);
        hipLaunchKernelGGL(k_tl_shift_f, dim3(nv_chunks), dim3(256), 0, st, f, m->fba->dbox, rv,
                           (const double*)(m->dsum + 2));
        KSGF_LAUNCH_CHECK();
        return KSGF_OK;
    }
};

}  // namespace

extern "C" {

int ksgf_macproj_levels_create(const ksgf_boxarray* crse_ba, const ksgf_boxarray* fine_ba, const int* lobc,
                               const int* hibc, double beta, const ksgf_mlmg_options* opts, ksgf_macproj_levels** out)
{
    if (!crse_ba || !fine_ba || !lobc || !hibc || !out || !(beta > 0.0)) {
        set_error("ksgf_macproj_levels_create: bad argument");
        return KSGF_ERR_INVALID;
    }
    *out = nullptr;
    if (fine_ba->nbox <= 0) {
        set_error("ksgf_macproj_levels_create: the fine level has no boxes (use ksgf_macproj_create)");
        return KSGF_ERR_INVALID;
    }
    if (!crse_ba->covers_domain) {
        set_error("ksgf_macproj_levels_create: the coarse level must cover its domain");
        return KSGF_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d) {
        const int per = (lobc[d] == KSGF_LO_PERIODIC);
        if (fine_ba->dlo[d] != 2 * crse_ba->dlo[d] || fine_ba->dhi[d] + 1 != 2 * (crse_ba->dhi[d] + 1) ||
            crse_ba->dlo[d] != 0 || per != (hibc[d] == KSGF_LO_PERIODIC) || per != crse_ba->periodic[d] ||
            per != fine_ba->periodic[d]) {
            set_error("ksgf_macproj_levels_create: the fine domain must be the coarse domain (origin 0) refined by 2, "
                      "with matching periodicity");
            return KSGF_ERR_INVALID;
        }
    }
    for (int b = 0; b < fine_ba->nbox; ++b) {
        const int* x = &fine_ba->hbox[6 * b];
        for (int d = 0; d < 3; ++d)
            if ((x[d] & 1) || !(x[3 + d] & 1)) {
                set_error("ksgf_macproj_levels_create: fine boxes must be aligned to the refinement ratio 2");
                return KSGF_ERR_INVALID;
            }
    }
    ksgf_macproj_levels* m = new ksgf_macproj_levels();
    m->cba = crse_ba;
    m->fba = fine_ba;
    if (opts)
        m->opt = *opts;
    else
        ksgf_mlmg_default_options(&m->opt);
    TLGeom& T = m->T;
    m->singular = 1;
    for (int d = 0; d < 3; ++d) {
        T.n[d] = crse_ba->dhi[d] + 1;
        T.nf[d] = 2 * T.n[d];
        T.per[d] = (lobc[d] == KSGF_LO_PERIODIC);
        T.lobc[d] = lobc[d];
        T.hibc[d] = hibc[d];
        T.hc[d] = crse_ba->dx[d];
        T.hf[d] = 0.5 * crse_ba->dx[d];
        T.ihc[d] = 1.0 / T.hc[d];
        T.ihf[d] = 1.0 / T.hf[d];
        if (lobc[d] == KSGF_LO_DIRICHLET || hibc[d] == KSGF_LO_DIRICHLET) m->singular = 0;
    }
    T.beta = beta;
    m->nc = (long long)T.n[0] * T.n[1] * T.n[2];
    for (int b = 0; b < fine_ba->nbox; ++b) {
        const int* x = &fine_ba->hbox[6 * b];
        m->nfine += (long long)(x[3] - x[0] + 1) * (x[4] - x[1] + 1) * (x[5] - x[2] + 1);
    }
    m->ncov = m->nfine / 8;
    int e = KSGF_OK, err = KSGF_OK;
    ksgf_mlmg_options co = m->opt;
    co.maxiter = 1;
    co.precision = KSGF_MG_FP64;
    const Region* rg = fab_region(fine_ba, KSGF_CELL, 1, &err);
    const Region* r0 = fab_region(fine_ba, KSGF_CELL, 0, &err);
    if (!rg || !r0) {
        delete m;
        return err;
    }
    m->nfg = rg->total;
    m->nf0 = r0->total;
    if ((e = ksgf_macproj_create(crse_ba, lobc, hibc, beta, &co, &m->crse))) {
        delete m;
        return e;
    }
    auto fail = [&](int code) {
        ksgf_macproj_levels_destroy(m);
        return code;
    };
    if (hipMalloc(&m->cv, m->nc) != hipSuccess || hipMalloc(&m->pc, m->nc * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->bc, m->nc * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->rc, m->nc * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->ec, m->nc * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->pf[0], m->nfg * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->pf[1], m->nfg * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->bf, m->nf0 * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->rf, m->nf0 * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->dmax, 2 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&m->dsum, (3 + kTlSumC + kTlSumF) * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&m->hmax, sizeof(unsigned long long)) != hipSuccess) {
        set_error("ksgf_macproj_levels_create: out of device memory");
        return fail(KSGF_ERR_HIP);
    }
    if ((e = owner_map(fine_ba, nullptr, &m->fown))) return fail(e);
    // face-ghost list: for every fine box, the six one-cell slabs outside its
    // faces (edges and corners excluded), x fastest within a slab
    {
        std::vector<int2> h;
        std::vector<long long> goff(fine_ba->nbox);
        for (int b = 0; b < fine_ba->nbox; ++b) {
            goff[b] = (long long)h.size();
            const int* x = &fine_ba->hbox[6 * b];
            const int n0 = x[3] - x[0] + 3, n1 = x[4] - x[1] + 3;
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    int lo[3] = {x[0], x[1], x[2]}, hi[3] = {x[3], x[4], x[5]};
                    lo[d] = hi[d] = side ? x[3 + d] + 1 : x[d] - 1;
                    for (int k = lo[2]; k <= hi[2]; ++k)
                        for (int j = lo[1]; j <= hi[1]; ++j)
                            for (int i = lo[0]; i <= hi[0]; ++i)
                                h.push_back(make_int2(b, (i - x[0] + 1) + n0 * ((j - x[1] + 1) + n1 * (k - x[2] + 1))));
                }
        }
        m->ngl = (long long)h.size();
        // column segments: per box, nx * ny columns times ceil(nz / kTlKz)
        // segments, x fastest; chunks of 256 consecutive segments of one box
        std::vector<int> ch;
        for (int b = 0; b < fine_ba->nbox; ++b) {
            const int* x = &fine_ba->hbox[6 * b];
            const long long n = (long long)(x[3] - x[0] + 1) * (x[4] - x[1] + 1) * ((x[5] - x[2] + kTlKz) / kTlKz);
            for (long long s0 = 0; s0 < n; s0 += 256) {
                ch.push_back(b);
                ch.push_back((int)s0);
            }
        }
        m->ncols_chunks = (int)(ch.size() / 2);
        if (hipMalloc(&m->cols, ch.size() * sizeof(int)) != hipSuccess ||
            hipMemcpy(m->cols, ch.data(), ch.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("ksgf_macproj_levels_create: column table");
            return fail(KSGF_ERR_HIP);
        }
        if (m->ngl && (hipMalloc(&m->gl, h.size() * sizeof(int2)) != hipSuccess ||
                       hipMemcpy(m->gl, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)) {
            set_error("ksgf_macproj_levels_create: ghost list");
            return fail(KSGF_ERR_HIP);
        }
        // ghost source codes (on the device, from the owner map), then the
        // list of coarse-fine entries and the slab descriptors on the host
        CMFDev cpf;
        if ((e = cmfdev(fine_ba, m->pf[0], KSGF_CELL, 1, 1, &cpf))) return fail(e);
        std::vector<long long> code(m->ngl), cfl;
        if (hipMalloc(&m->goff, goff.size() * sizeof(long long)) != hipSuccess ||
            hipMemcpy(m->goff, goff.data(), goff.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
            hipMalloc(&m->gcode, std::max<long long>(1, m->ngl) * sizeof(long long)) != hipSuccess) {
            set_error("ksgf_macproj_levels_create: ghost codes");
            return fail(KSGF_ERR_HIP);
        }
        if (m->ngl) {
            hipLaunchKernelGGL(k_tl_gsrc, dim3(grid_for(m->ngl, 256)), dim3(256), 0, nullptr, cpf, fine_ba->dbox,
                               (const int2*)m->gl, m->ngl, m->fown, T, m->gcode);
            if (hipMemcpy(code.data(), m->gcode, code.size() * sizeof(long long), hipMemcpyDeviceToHost) !=
                hipSuccess) {
                set_error("ksgf_macproj_levels_create: ghost codes");
                return fail(KSGF_ERR_HIP);
            }
        }
        for (long long t = 0; t < m->ngl; ++t)
            if (code[t] == kGhCF) cfl.push_back(t);
        // slab descriptors: affine where every entry is a cell of one box at
        // an offset affine in the slab coordinates, else one kind for the
        // whole slab or mixed
        std::vector<TLFace> face(6 * (size_t)fine_ba->nbox);
        for (int b = 0; b < fine_ba->nbox; ++b) {
            const int* x = &fine_ba->hbox[6 * b];
            const int n[3] = {x[3] - x[0] + 1, x[4] - x[1] + 1, x[5] - x[2] + 1};
            long long t0 = goff[b];
            for (int d = 0; d < 3; ++d) {
                const int na = (d == 0) ? n[1] : n[0], nb = (d == 2) ? n[1] : n[2];
                for (int side = 0; side < 2; ++side, t0 += (long long)na * nb) {
                    TLFace& F = face[6 * b + 2 * d + side];
                    F = TLFace{0, 0, 0, kFaceMixed, 0};
                    const long long* cd = code.data() + t0;
                    const long long c0 = cd[0];
                    bool same = true;
                    for (long long e = 1; e < (long long)na * nb && same; ++e) same = (cd[e] == c0);
                    if (same && c0 == kGhNeumann) F.kind = kFaceNeumann;
                    if (same && c0 == kGhDirichlet) F.kind = kFaceDirichlet;
                    if (same && c0 == kGhCF) F.kind = kFaceCF;
                    if (c0 < 0 || cd[1] < 0 || cd[na] < 0) continue;
                    const long long sa = cd[1] - c0, sb = cd[na] - c0;
                    if (sa <= 0 || sb <= 0 || sa > INT32_MAX || sb > INT32_MAX) continue;
                    bool aff = true;
                    for (int bb = 0; bb < nb && aff; ++bb)
                        for (int aa = 0; aa < na && aff; ++aa)
                            aff = (cd[aa + (long long)na * bb] == c0 + aa * sa + bb * sb);
                    if (aff) F = TLFace{c0, (int)sa, (int)sb, kFaceAffine, 0};
                }
            }
        }
        if (hipMalloc(&m->face, face.size() * sizeof(TLFace)) != hipSuccess ||
            hipMemcpy(m->face, face.data(), face.size() * sizeof(TLFace), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("ksgf_macproj_levels_create: ghost slab table");
            return fail(KSGF_ERR_HIP);
        }
        m->ncf = (long long)cfl.size();
        if (hipMalloc(&m->cfl, std::max<long long>(1, m->ncf) * sizeof(long long)) != hipSuccess ||
            hipMalloc(&m->bq, std::max<long long>(1, m->ngl) * sizeof(double)) != hipSuccess ||
            (m->ncf && hipMemcpy(m->cfl, cfl.data(), cfl.size() * sizeof(long long), hipMemcpyHostToDevice) !=
                           hipSuccess)) {
            set_error("ksgf_macproj_levels_create: coarse-fine ghost table");
            return fail(KSGF_ERR_HIP);
        }
    }
    hipLaunchKernelGGL(k_tl_cover, dim3(grid_for(m->nc, 256)), dim3(256), 0, nullptr, m->fown, m->cv, T);
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("ksgf_macproj_levels_create: covered mask");
        return fail(KSGF_ERR_HIP);
    }
    *out = m;
    return KSGF_OK;
}

int ksgf_macproj_levels_project(ksgf_macproj_levels* m, double* const* umac_c, double* const* umac_f, int ng,
                                double rtol, double atol, ksgf_mlmg_info* info, ksgf_stream_t stream)
{
    if (!m || !umac_c || !umac_f || ng < 0) {
        set_error("ksgf_macproj_levels_project: bad argument");
        return KSGF_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d)
        if (!umac_c[d] || !umac_f[d]) {
            set_error("ksgf_macproj_levels_project: null face MultiFab");
            return KSGF_ERR_INVALID;
        }
    hipStream_t st = (hipStream_t)stream;
    const ksgf_boxarray* cba = m->cba;
    const ksgf_boxarray* fba = m->fba;
    int e, err = KSGF_OK;
    TLRun R;
    R.m = m;
    R.st = st;
    if ((e = tl_mfs(m, R.pfm, R.cpfm, &R.bf, &R.rf, &R.crf))) return e;
    const int z[3] = {0, 0, 0};
    const Region* rv = get_region(fba, z, z, &err);
    const Region* rcv = get_region(cba, z, z, &err);
    if (!rv || !rcv) return err;
    R.rv = regdev(rv, fba->nbox);
    R.nv_chunks = rv->nchunks;
    MFDev uf[3], uc[3];
    CMFDev cuf[3], cuc[3];
    for (int d = 0; d < 3; ++d)
        if ((e = mfdev(fba, umac_f[d], KSGF_XFACE + d, 1, ng, &uf[d])) ||
            (e = cmfdev(fba, umac_f[d], KSGF_XFACE + d, 1, ng, &cuf[d])) ||
            (e = mfdev(cba, umac_c[d], KSGF_XFACE + d, 1, ng, &uc[d])) ||
            (e = cmfdev(cba, umac_c[d], KSGF_XFACE + d, 1, ng, &cuc[d])))
            return e;
    // one velocity per coarse-fine face: the fine average on the coarse faces
    // under and around the fine level
    if ((e = ksgf_average_down_faces(fba, umac_f[0], umac_f[1], umac_f[2], ng, cba, umac_c[0], umac_c[1], umac_c[2],
                                     ng, 1, 2, stream)))
        return e;
    // right-hand sides
    const TLGeom& T = m->T;
    hipLaunchKernelGGL(k_tl_div_f, dim3(R.nv_chunks), dim3(256), 0, st, MFDev{m->bf, R.bf.pre, KSGF_CELL, 0, 1},
                       cuf[0], cuf[1], cuf[2], fba->dbox, T.hf[0], T.hf[1], T.hf[2], R.rv);
    hipLaunchKernelGGL(k_tl_div_c, dim3(rcv->nchunks), dim3(256), 0, st, m->bc, (const unsigned char*)m->cv, cuc[0],
                       cuc[1], cuc[2], cba->dbox, T, regdev(rcv, cba->nbox));
    KSGF_LAUNCH_CHECK();
    if (m->singular && (e = R.sub_mean(m->bc, MFDev{m->bf, R.bf.pre, KSGF_CELL, 0, 1}, R.bf, 0))) return e;
    KSGF_HIP_CHECK(hipMemsetAsync(m->pc, 0, m->nc * sizeof(double), st));
    m->cur = 0;
    KSGF_HIP_CHECK(hipMemsetAsync(m->pf[0], 0, m->nfg * sizeof(double), st));
    R.coarse_changed();
    const double diag = -2.0 * T.beta * ((1.0 / (T.hf[0] * T.hf[0]) + 1.0 / (T.hf[1] * T.hf[1])) +
                                         1.0 / (T.hf[2] * T.hf[2]));
    const double w = kOmega / diag;
    R.residual();
    KSGF_LAUNCH_CHECK();
    double r0 = 0.0, r = 0.0;
    if ((e = R.norm(&r0))) return e;
    const double target = std::max(atol, rtol * r0);
    r = r0;
    ksgf_mlmg_info inf{};
    inf.r0 = r0;
    inf.converged = (r <= target);
    inf.nlevels = 2;
    int ret = KSGF_OK;
    for (int it = 1; !inf.converged && it <= m->opt.maxiter; ++it) {
        R.smooth(2 * m->opt.num_pre_smooth, w);
        R.residual();
        hipLaunchKernelGGL(k_tl_avg8, dim3(grid_for(m->nc, 256)), dim3(256), 0, st, m->rc,
                           (const unsigned char*)m->cv, R.crf, fba->dbox, m->fown, T);
        KSGF_LAUNCH_CHECK();
        ksgf_mlmg_info ci{};
        if ((e = macproj_vcycles(m->crse, m->rc, m->ec, &ci, st))) return e;
        inf.bottom_iters += ci.bottom_iters;
        hipLaunchKernelGGL(k_tl_addc, dim3(grid_for(m->nc, 256)), dim3(256), 0, st, m->pc, (const double*)m->ec, m->nc);
        hipLaunchKernelGGL(k_tl_prolong, dim3(R.nv_chunks), dim3(256), 0, st, R.pf(), (const double*)m->ec, fba->dbox,
                           T, R.rv);
        R.coarse_changed();
        R.smooth(2 * m->opt.num_post_smooth, w);
        R.residual();
        KSGF_LAUNCH_CHECK();
        if ((e = R.norm(&r))) return e;
        inf.iters = it;
        if (r <= target)
            inf.converged = 1;
        else if (!(r <= 1.0e20 * r0)) {
            ret = KSGF_ERR_DIVERGED;
            break;
        }
    }
    inf.rfinal = r;
    if (info) *info = inf;
    if (ret) {
        set_error("ksgf_macproj_levels_project: the composite iteration diverged");
        return ret;
    }
    // covered coarse cells: the average of their children; singular: the
    // composite mean removed
    hipLaunchKernelGGL(k_tl_avg8, dim3(grid_for(m->nc, 256)), dim3(256), 0, st, m->pc, (const unsigned char*)m->cv,
                       R.cpf(), fba->dbox, m->fown, T);
    KSGF_LAUNCH_CHECK();
    if (m->singular && (e = R.sub_mean(m->pc, R.pf(), R.cpf(), 1))) return e;
    // u -= beta grad phi on both levels, then the fine average on the coarse
    // faces under and around the fine level
    R.ghosts();
    for (int d = 0; d < 3; ++d) {
        int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        ghi[d] = 1;
        const Region* ff = get_region(fba, glo, ghi, &err);
        const Region* fc = get_region(cba, glo, ghi, &err);
        if (!ff || !fc) return err;
        RegDev rf = regdev(ff, fba->nbox), rc = regdev(fc, cba->nbox);
        switch (d) {
        case 0:
            hipLaunchKernelGGL(k_tl_fix_f<0>, dim3(ff->nchunks), dim3(256), 0, st, uf[0], R.cpf(), fba->dbox, T, rf);
            hipLaunchKernelGGL(k_tl_fix_c<0>, dim3(fc->nchunks), dim3(256), 0, st, uc[0], (const double*)m->pc,
                               cba->dbox, T, rc);
            break;
        case 1:
            hipLaunchKernelGGL(k_tl_fix_f<1>, dim3(ff->nchunks), dim3(256), 0, st, uf[1], R.cpf(), fba->dbox, T, rf);
            hipLaunchKernelGGL(k_tl_fix_c<1>, dim3(fc->nchunks), dim3(256), 0, st, uc[1], (const double*)m->pc,
                               cba->dbox, T, rc);
            break;
        default:
            hipLaunchKernelGGL(k_tl_fix_f<2>, dim3(ff->nchunks), dim3(256), 0, st, uf[2], R.cpf(), fba->dbox, T, rf);
            hipLaunchKernelGGL(k_tl_fix_c<2>, dim3(fc->nchunks), dim3(256), 0, st, uc[2], (const double*)m->pc,
                               cba->dbox, T, rc);
            break;
        }
        KSGF_LAUNCH_CHECK();
    }
    if ((e = ksgf_average_down_faces(fba, umac_f[0], umac_f[1], umac_f[2], ng, cba, umac_c[0], umac_c[1], umac_c[2],
                                     ng, 1, 2, stream)))
        return e;
    return inf.converged ? KSGF_OK : KSGF_ERR_NOT_CONVERGED;
}

int ksgf_macproj_levels_get_phi(ksgf_macproj_levels* m, double* phi_c, double* phi_f, int ng_f, ksgf_stream_t stream)
{
    if (!m || !phi_c || !phi_f || ng_f != 1) {
        set_error("ksgf_macproj_levels_get_phi: bad argument (the fine solution has one ghost layer)");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    KSGF_HIP_CHECK(hipMemcpyAsync(phi_c, m->pc, m->nc * sizeof(double), hipMemcpyDeviceToDevice, st));
    KSGF_HIP_CHECK(hipMemcpyAsync(phi_f, m->pf[m->cur], m->nfg * sizeof(double), hipMemcpyDeviceToDevice, st));
    return KSGF_OK;
}

int ksgf_macproj_levels_destroy(ksgf_macproj_levels* m)
{
    if (!m) return KSGF_OK;
    if (m->crse) ksgf_macproj_destroy(m->crse);
    hipFree(m->cv);
    hipFree(m->pc);
    hipFree(m->bc);
    hipFree(m->rc);
    hipFree(m->ec);
    hipFree(m->pf[0]);
    hipFree(m->pf[1]);
    hipFree(m->gl);
    hipFree(m->gcode);
    hipFree(m->goff);
    hipFree(m->face);
    hipFree(m->cfl);
    hipFree(m->bq);
    hipFree(m->cols);
    hipFree(m->bf);
    hipFree(m->rf);
    hipFree(m->dmax);
    hipFree(m->dsum);
    if (m->hmax) hipHostFree(m->hmax);
    delete m;
    return KSGF_OK;
}

}  // extern "C"
