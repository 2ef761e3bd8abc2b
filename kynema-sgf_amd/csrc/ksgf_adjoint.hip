// Reverse mode (adjoint) of the advect + project step of a passive scalar
// and the velocity on a triply periodic one-box domain, with checkpointed
// recomputation.
//
// The forward step is the step of the Godunov predictor without diffusion
// and nodal projection (incflo_advance.cpp:186-438 as driven by the C++ host
// layer) for density 1 and zero pressure gradient, at a fixed dt:
//   u_mac  = ExtrapVelToFaces(u)                         icns_advection.H:211-216
//   u_mac <- MacProjector::project(u_mac)                icns_advection.cpp:331
//   s'     = s + dt * (-div F(s, u_mac))                 AdvOp_Godunov.H:231-297
//   u'     = u + dt * (-div F(u, u_mac))                 icns_advection.H:359-422
// with F the conservative PPM fluxes (ComputeFluxesOnBoxFromState).  The
// reference has no adjoint; this is the differentiation workload's own.
//
// Adjoint of one step, in reverse order of the forward calls:
//   predictor update and divergence: linear, transposed exactly;
//   fluxes: the three Godunov stages are reversed one at a time (finals,
//     corner couplings, normal traces) from the forward stage arrays, which
//     are recomputed before each reversal.  Upwind and Burgers selections
//     and every limiter branch are frozen at the forward state; the PPM
//     trace of a cell is differentiated with dual numbers over its six
//     stencil values and advecting velocities (ksgf_dual.h), the other
//     stages with hand-derived partials;
//   projection: P = I - G L^{-1} D with G = -beta D^T is symmetric, so its
//     adjoint is the same projection applied to the face adjoint.
// Adjoint contributions to a shared point are added with device atomics,
// so gradients agree between runs to rounding, not bit for bit.
//
// Arrays of the adjoint are domain-wide and periodic (n0*n1*n2 per
// component, x fastest, no ghosts); the face arrays hold the n_d distinct
// faces of a periodic direction (face i = low face of cell i).  Forward
// values are read from the one-box MultiFabs (with their ghost layers)
// through periodic wrapping.
#include "ksgf_internal.h"
#include "ksgf_ppm.h"

namespace ksgf {

struct PGeom {
    int n0, n1, n2;
    int N;
};

__device__ __forceinline__ int wrp(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// read-only view of a one-box FAB (ngrow g, nodal in direction nd or -1)
struct FV {
    const double* p;
    int ng, m0, m1, cs;
};

static FV make_fv(const double* p, const PGeom& g, int ngrow, int nd)
{
    FV v;
    v.p = p;
    v.ng = ngrow;
    v.m0 = g.n0 + 2 * ngrow + (nd == 0);
    v.m1 = g.n1 + 2 * ngrow + (nd == 1);
    const int m2 = g.n2 + 2 * ngrow + (nd == 2);
    v.cs = v.m0 * v.m1 * m2;
    return v;
}
static FV make_fv(const MFDev& m, const PGeom& g)
{
    return make_fv(m.p, g, m.ngrow, m.itype - 1);
}

// value at periodic (i, j, k) (any integers within one period of the box)
__device__ __forceinline__ double rd(const FV& v, const PGeom& g, int i, int j, int k, int n = 0)
{
    i = wrp(i, g.n0);
    j = wrp(j, g.n1);
    k = wrp(k, g.n2);
    return v.p[(i + v.ng) + v.m0 * ((j + v.ng) + v.m1 * (k + v.ng)) + n * v.cs];
}

// index into a periodic domain-wide array
__device__ __forceinline__ int pidx(const PGeom& g, int i, int j, int k, int n = 0)
{
    return wrp(i, g.n0) + g.n0 * (wrp(j, g.n1) + g.n1 * wrp(k, g.n2)) + n * g.N;
}

__device__ __forceinline__ bool point_of(const PGeom& g, int& i, int& j, int& k)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= g.N) return false;
    const int q = t / g.n0;
    i = t - q * g.n0;
    k = q / g.n1;
    j = q - k * g.n1;
    return true;
}

template <int D>
__device__ __forceinline__ void shift(int& i, int& j, int& k, int s)
{
    if (D == 0) i += s;
    if (D == 1) j += s;
    if (D == 2) k += s;
}

__device__ __forceinline__ void aadd(double* p, double v)
{
    if (v != 0.0) atomicAdd(p, v);
}

// ---------------------------------------------------------------------------
// divergence and predictor update:  x' = x + dt * mult * sum_d dxi_d (F_d(+) - F_d(-))
// gives  Fbar_d(f) = dt * mult * dxi_d * (xbar(cell below f) - xbar(cell above f))
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256) k_adj_flux_from_update(const double* __restrict__ xb, double* __restrict__ fb,
                                                              int ncomp, PGeom g, double c)
{
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    int qi = i, qj = j, qk = k;
    shift<D>(qi, qj, qk, -1);
    for (int n = 0; n < ncomp; ++n) fb[pidx(g, i, j, k, n)] = c * (xb[pidx(g, qi, qj, qk, n)] - xb[pidx(g, i, j, k, n)]);
}

// ---------------------------------------------------------------------------
// flux mode (ComputeFluxesOnBoxFromState with known MAC velocities,
// conservative, no divu, forces not in the transverse terms)
// ---------------------------------------------------------------------------

// finals of direction D: flux = u * upwind(stl, sth, u)
template <int D>
__global__ void __launch_bounds__(256) k_adj_final(PGeom g, int ncomp, const double* __restrict__ fbar, FV mac0, FV mac1,
                                                   FV mac2, FV q, FV face, FV crnA, FV crnB, double c2a, double c2b,
                                                   double* __restrict__ lob, double* __restrict__ hib,
                                                   double* __restrict__ crnAb, double* __restrict__ crnBb,
                                                   double* __restrict__ macb0, double* __restrict__ macb1,
                                                   double* __restrict__ macb2, double* __restrict__ qb)
{
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const FV macs[3] = {mac0, mac1, mac2};
    double* macbs[3] = {macb0, macb1, macb2};
    const FV& vD = macs[D];
    const FV& v1 = macs[T1];
    const FV& v2 = macs[T2];
    int qi = i, qj = j, qk = k;
    shift<D>(qi, qj, qk, -1);
    int q1i = qi, q1j = qj, q1k = qk, p1i = i, p1j = j, p1k = k;
    shift<T1>(q1i, q1j, q1k, 1);
    shift<T1>(p1i, p1j, p1k, 1);
    int q2i = qi, q2j = qj, q2k = qk, p2i = i, p2j = j, p2k = k;
    shift<T2>(q2i, q2j, q2k, 1);
    shift<T2>(p2i, p2j, p2k, 1);
    const double u = rd(vD, g, i, j, k);
    const double v1ql = rd(v1, g, qi, qj, qk), v1qh = rd(v1, g, q1i, q1j, q1k);
    const double v1pl = rd(v1, g, i, j, k), v1ph = rd(v1, g, p1i, p1j, p1k);
    const double v2ql = rd(v2, g, qi, qj, qk), v2qh = rd(v2, g, q2i, q2j, q2k);
    const double v2pl = rd(v2, g, i, j, k), v2ph = rd(v2, g, p2i, p2j, p2k);
    double dlo, dhi;
    upwind_partials(u, dlo, dhi);
    double ub = 0.0;
    double v1qlb = 0, v1qhb = 0, v1plb = 0, v1phb = 0, v2qlb = 0, v2qhb = 0, v2plb = 0, v2phb = 0;
    for (int n = 0; n < ncomp; ++n) {
        const int f = pidx(g, i, j, k, n);
        const double fb = fbar[f];
        if (fb == 0.0) continue;
        const double e = rd(face, g, i, j, k, n);
        ub += e * fb;
        const double eb = u * fb;
        const double stlb = dlo * eb, sthb = dhi * eb;
        lob[f] += stlb;
        hib[f] += sthb;
        const double a_ql = rd(crnA, g, qi, qj, qk, n), a_qh = rd(crnA, g, q1i, q1j, q1k, n);
        const double a_pl = rd(crnA, g, i, j, k, n), a_ph = rd(crnA, g, p1i, p1j, p1k, n);
        const double b_ql = rd(crnB, g, qi, qj, qk, n), b_qh = rd(crnB, g, q2i, q2j, q2k, n);
        const double b_pl = rd(crnB, g, i, j, k, n), b_ph = rd(crnB, g, p2i, p2j, p2k, n);
        const double ql = rd(q, g, qi, qj, qk, n), qh = rd(q, g, i, j, k, n);
        // stl = lo - c2a (v1qh a_qh - v1ql a_ql) - c2b (v2qh b_qh - v2ql b_ql)
        //          + c2a ql (v1qh - v1ql) + c2b ql (v2qh - v2ql)
        v1qhb += stlb * (-c2a * a_qh + c2a * ql);
        v1qlb += stlb * (c2a * a_ql - c2a * ql);
        v2qhb += stlb * (-c2b * b_qh + c2b * ql);
        v2qlb += stlb * (c2b * b_ql - c2b * ql);
        aadd(&crnAb[pidx(g, q1i, q1j, q1k, n)], stlb * (-c2a * v1qh));
        aadd(&crnAb[pidx(g, qi, qj, qk, n)], stlb * (c2a * v1ql));
        aadd(&crnBb[pidx(g, q2i, q2j, q2k, n)], stlb * (-c2b * v2qh));
        aadd(&crnBb[pidx(g, qi, qj, qk, n)], stlb * (c2b * v2ql));
        aadd(&qb[pidx(g, qi, qj, qk, n)], stlb * (c2a * (v1qh - v1ql) + c2b * (v2qh - v2ql)));
        v1phb += sthb * (-c2a * a_ph + c2a * qh);
        v1plb += sthb * (c2a * a_pl - c2a * qh);
        v2phb += sthb * (-c2b * b_ph + c2b * qh);
        v2plb += sthb * (c2b * b_pl - c2b * qh);
        aadd(&crnAb[pidx(g, p1i, p1j, p1k, n)], sthb * (-c2a * v1ph));
        aadd(&crnAb[pidx(g, i, j, k, n)], sthb * (c2a * v1pl));
        aadd(&crnBb[pidx(g, p2i, p2j, p2k, n)], sthb * (-c2b * v2ph));
        aadd(&crnBb[pidx(g, i, j, k, n)], sthb * (c2b * v2pl));
        aadd(&qb[pidx(g, i, j, k, n)], sthb * (c2a * (v1ph - v1pl) + c2b * (v2ph - v2pl)));
    }
    aadd(&macbs[D][pidx(g, i, j, k)], ub);
    aadd(&macbs[T1][pidx(g, qi, qj, qk)], v1qlb);
    aadd(&macbs[T1][pidx(g, q1i, q1j, q1k)], v1qhb);
    aadd(&macbs[T1][pidx(g, i, j, k)], v1plb);
    aadd(&macbs[T1][pidx(g, p1i, p1j, p1k)], v1phb);
    aadd(&macbs[T2][pidx(g, qi, qj, qk)], v2qlb);
    aadd(&macbs[T2][pidx(g, q2i, q2j, q2k)], v2qhb);
    aadd(&macbs[T2][pidx(g, i, j, k)], v2plb);
    aadd(&macbs[T2][pidx(g, p2i, p2j, p2k)], v2phb);
}

// corner coupling of faces normal to D along T: crn = upwind(lov, hiv, u_D)
template <int D, int T>
__global__ void __launch_bounds__(256) k_adj_corner(PGeom g, int ncomp, const double* __restrict__ crnb, FV vT, FV vD,
                                                    FV q, FV edT, double c3, double* __restrict__ lob,
                                                    double* __restrict__ hib, double* __restrict__ edTb,
                                                    double* __restrict__ vTb, double* __restrict__ qb)
{
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    int qi = i, qj = j, qk = k;
    shift<D>(qi, qj, qk, -1);
    int qti = qi, qtj = qj, qtk = qk, pti = i, ptj = j, ptk = k;
    shift<T>(qti, qtj, qtk, 1);
    shift<T>(pti, ptj, ptk, 1);
    const double u = rd(vD, g, i, j, k);
    const double wql = rd(vT, g, qi, qj, qk), wqh = rd(vT, g, qti, qtj, qtk);
    const double wpl = rd(vT, g, i, j, k), wph = rd(vT, g, pti, ptj, ptk);
    double dlo, dhi;
    upwind_partials(u, dlo, dhi);
    double wqlb = 0, wqhb = 0, wplb = 0, wphb = 0;
    for (int n = 0; n < ncomp; ++n) {
        const int f = pidx(g, i, j, k, n);
        const double cb = crnb[f];
        if (cb == 0.0) continue;
        const double lovb = dlo * cb, hivb = dhi * cb;
        lob[f] += lovb;
        hib[f] += hivb;
        const double eql = rd(edT, g, qi, qj, qk, n), eqh = rd(edT, g, qti, qtj, qtk, n);
        const double epl = rd(edT, g, i, j, k, n), eph = rd(edT, g, pti, ptj, ptk, n);
        const double ql = rd(q, g, qi, qj, qk, n), qh = rd(q, g, i, j, k, n);
        // lov = lo - c3 (wqh eqh - wql eql) + c3 ql (wqh - wql)
        wqhb += lovb * (-c3 * eqh + c3 * ql);
        wqlb += lovb * (c3 * eql - c3 * ql);
        aadd(&edTb[pidx(g, qti, qtj, qtk, n)], lovb * (-c3 * wqh));
        aadd(&edTb[pidx(g, qi, qj, qk, n)], lovb * (c3 * wql));
        aadd(&qb[pidx(g, qi, qj, qk, n)], lovb * (c3 * (wqh - wql)));
        wphb += hivb * (-c3 * eph + c3 * qh);
        wplb += hivb * (c3 * epl - c3 * qh);
        aadd(&edTb[pidx(g, pti, ptj, ptk, n)], hivb * (-c3 * wph));
        aadd(&edTb[pidx(g, i, j, k, n)], hivb * (c3 * wpl));
        aadd(&qb[pidx(g, i, j, k, n)], hivb * (c3 * (wph - wpl)));
    }
    aadd(&vTb[pidx(g, qi, qj, qk)], wqlb);
    aadd(&vTb[pidx(g, qti, qtj, qtk)], wqhb);
    aadd(&vTb[pidx(g, i, j, k)], wplb);
    aadd(&vTb[pidx(g, pti, ptj, ptk)], wphb);
}

// normal traces of direction D: lo = Ip(low cell), hi = Im(high cell),
// ed = upwind(lo, hi, u_D)
template <int D>
__global__ void __launch_bounds__(256) k_adj_faceA(PGeom g, int ncomp, FV q, FV vD, double dtdx, int lim,
                                                   const double* __restrict__ lob, const double* __restrict__ hib,
                                                   const double* __restrict__ edb, double* __restrict__ vDb,
                                                   double* __restrict__ qb)
{
    using D7 = Dual<7>;
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const int nD = D == 0 ? g.n0 : (D == 1 ? g.n1 : g.n2);
    const int x = D == 0 ? i : (D == 1 ? j : k);
    int ci[6], cj[6], ck[6];
    for (int m = 0; m < 6; ++m) {
        ci[m] = i;
        cj[m] = j;
        ck[m] = k;
        shift<D>(ci[m], cj[m], ck[m], m - 3);
    }
    int ni = i, nj = j, nk = k;
    shift<D>(ni, nj, nk, 1);
    int qi = i, qj = j, qk = k;
    shift<D>(qi, qj, qk, -1);
    const double u = rd(vD, g, i, j, k);
    const D7 uf = dual_seed<7>(u, 6);
    const D7 uq = dual_const<7>(rd(vD, g, qi, qj, qk));
    const D7 un = dual_const<7>(rd(vD, g, ni, nj, nk));
    double dlo, dhi;
    upwind_partials(u, dlo, dhi);
    double ub = 0.0;
    for (int n = 0; n < ncomp; ++n) {
        const int f = pidx(g, i, j, k, n);
        const double lovb = lob[f] + dlo * edb[f];
        const double hivb = hib[f] + dhi * edb[f];
        if (lovb == 0.0 && hivb == 0.0) continue;
        D7 s[6];
        for (int m = 0; m < 6; ++m) s[m] = dual_seed<7>(rd(q, g, ci[m], cj[m], ck[m], n), m);
        D7 dmy, lov, hiv;
        // low cell: stencil x-3 .. x+1, advected by u(x-1) (low) and u(x) (high)
        ppm_trace(s[0], s[1], s[2], s[3], s[4], lim, KSGF_BC_INT_DIR, KSGF_BC_INT_DIR, x - 1, 0, nD - 1, uq, uf, dtdx,
                  dmy, lov);
        // high cell: stencil x-2 .. x+2, advected by u(x) (low) and u(x+1) (high)
        ppm_trace(s[1], s[2], s[3], s[4], s[5], lim, KSGF_BC_INT_DIR, KSGF_BC_INT_DIR, x, 0, nD - 1, uf, un, dtdx, hiv,
                  dmy);
        for (int m = 0; m < 6; ++m) aadd(&qb[pidx(g, ci[m], cj[m], ck[m], n)], lovb * lov.d[m] + hivb * hiv.d[m]);
        ub += lovb * lov.d[6] + hivb * hiv.d[6];
    }
    aadd(&vDb[pidx(g, i, j, k)], ub);
}

// ---------------------------------------------------------------------------
// velocity mode (ExtrapVelToFaces): advective velocities from Burgers
// ---------------------------------------------------------------------------

// finals: u_mac_D = burgers(stl, sth) of the normal component
template <int D>
__global__ void __launch_bounds__(256) k_adj_final_vel(PGeom g, const double* __restrict__ macb, FV loD, FV hiD,
                                                       FV uad1, FV uad2, FV crnA, FV crnB, double c4a, double c4b,
                                                       double* __restrict__ lob, double* __restrict__ hib,
                                                       double* __restrict__ uad1b, double* __restrict__ uad2b,
                                                       double* __restrict__ crnAb, double* __restrict__ crnBb)
{
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const double mb = macb[pidx(g, i, j, k)];
    if (mb == 0.0) return;
    int qi = i, qj = j, qk = k;
    shift<D>(qi, qj, qk, -1);
    int q1i = qi, q1j = qj, q1k = qk, p1i = i, p1j = j, p1k = k;
    shift<T1>(q1i, q1j, q1k, 1);
    shift<T1>(p1i, p1j, p1k, 1);
    int q2i = qi, q2j = qj, q2k = qk, p2i = i, p2j = j, p2k = k;
    shift<T2>(q2i, q2j, q2k, 1);
    shift<T2>(p2i, p2j, p2k, 1);
    const double v1ql = rd(uad1, g, qi, qj, qk), v1qh = rd(uad1, g, q1i, q1j, q1k);
    const double v1pl = rd(uad1, g, i, j, k), v1ph = rd(uad1, g, p1i, p1j, p1k);
    const double v2ql = rd(uad2, g, qi, qj, qk), v2qh = rd(uad2, g, q2i, q2j, q2k);
    const double v2pl = rd(uad2, g, i, j, k), v2ph = rd(uad2, g, p2i, p2j, p2k);
    const double A_q = rd(crnA, g, qi, qj, qk), A_q1 = rd(crnA, g, q1i, q1j, q1k);
    const double A_p = rd(crnA, g, i, j, k), A_p1 = rd(crnA, g, p1i, p1j, p1k);
    const double B_q = rd(crnB, g, qi, qj, qk), B_q2 = rd(crnB, g, q2i, q2j, q2k);
    const double B_p = rd(crnB, g, i, j, k), B_p2 = rd(crnB, g, p2i, p2j, p2k);
    // the forward expressions (same operand order), for the Burgers branch
    const double stl = rd(loD, g, i, j, k, D) - c4a * (v1qh + v1ql) * (A_q1 - A_q) - c4b * (v2qh + v2ql) * (B_q2 - B_q);
    const double sth = rd(hiD, g, i, j, k, D) - c4a * (v1ph + v1pl) * (A_p1 - A_p) - c4b * (v2ph + v2pl) * (B_p2 - B_p);
    double blo, bhi;
    burgers_partials(stl, sth, blo, bhi);
    const double stlb = blo * mb, sthb = bhi * mb;
    const int f = pidx(g, i, j, k, D);
    lob[f] += stlb;
    hib[f] += sthb;
    aadd(&uad1b[pidx(g, qi, qj, qk)], stlb * (-c4a * (A_q1 - A_q)));
    aadd(&uad1b[pidx(g, q1i, q1j, q1k)], stlb * (-c4a * (A_q1 - A_q)));
    aadd(&crnAb[pidx(g, q1i, q1j, q1k)], stlb * (-c4a * (v1qh + v1ql)));
    aadd(&crnAb[pidx(g, qi, qj, qk)], stlb * (c4a * (v1qh + v1ql)));
    aadd(&uad2b[pidx(g, qi, qj, qk)], stlb * (-c4b * (B_q2 - B_q)));
    aadd(&uad2b[pidx(g, q2i, q2j, q2k)], stlb * (-c4b * (B_q2 - B_q)));
    aadd(&crnBb[pidx(g, q2i, q2j, q2k)], stlb * (-c4b * (v2qh + v2ql)));
    aadd(&crnBb[pidx(g, qi, qj, qk)], stlb * (c4b * (v2qh + v2ql)));
    aadd(&uad1b[pidx(g, i, j, k)], sthb * (-c4a * (A_p1 - A_p)));
    aadd(&uad1b[pidx(g, p1i, p1j, p1k)], sthb * (-c4a * (A_p1 - A_p)));
    aadd(&crnAb[pidx(g, p1i, p1j, p1k)], sthb * (-c4a * (v1ph + v1pl)));
    aadd(&crnAb[pidx(g, i, j, k)], sthb * (c4a * (v1ph + v1pl)));
    aadd(&uad2b[pidx(g, i, j, k)], sthb * (-c4b * (B_p2 - B_p)));
    aadd(&uad2b[pidx(g, p2i, p2j, p2k)], sthb * (-c4b * (B_p2 - B_p)));
    aadd(&crnBb[pidx(g, p2i, p2j, p2k)], sthb * (-c4b * (v2ph + v2pl)));
    aadd(&crnBb[pidx(g, i, j, k)], sthb * (c4b * (v2ph + v2pl)));
}

// corner coupling of component E = 3-D-T on faces normal to D along T
template <int D, int T>
__global__ void __launch_bounds__(256) k_adj_corner_vel(PGeom g, const double* __restrict__ crnb, FV uadT, FV uadD,
                                                        FV edT, double c6, double* __restrict__ lob,
                                                        double* __restrict__ hib, double* __restrict__ edTb,
                                                        double* __restrict__ uadTb)
{
    constexpr int E = 3 - D - T;
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const double cb = crnb[pidx(g, i, j, k)];
    if (cb == 0.0) return;
    int qi = i, qj = j, qk = k;
    shift<D>(qi, qj, qk, -1);
    int qti = qi, qtj = qj, qtk = qk, pti = i, ptj = j, ptk = k;
    shift<T>(qti, qtj, qtk, 1);
    shift<T>(pti, ptj, ptk, 1);
    double dlo, dhi;
    upwind_partials(rd(uadD, g, i, j, k), dlo, dhi);
    const double lovb = dlo * cb, hivb = dhi * cb;
    const int f = pidx(g, i, j, k, E);
    lob[f] += lovb;
    hib[f] += hivb;
    const double wql = rd(uadT, g, qi, qj, qk), wqh = rd(uadT, g, qti, qtj, qtk);
    const double wpl = rd(uadT, g, i, j, k), wph = rd(uadT, g, pti, ptj, ptk);
    const double eql = rd(edT, g, qi, qj, qk, E), eqh = rd(edT, g, qti, qtj, qtk, E);
    const double epl = rd(edT, g, i, j, k, E), eph = rd(edT, g, pti, ptj, ptk, E);
    // lov = lo - c6 (wqh + wql) (eqh - eql)
    aadd(&uadTb[pidx(g, qi, qj, qk)], lovb * (-c6 * (eqh - eql)));
    aadd(&uadTb[pidx(g, qti, qtj, qtk)], lovb * (-c6 * (eqh - eql)));
    aadd(&edTb[pidx(g, qti, qtj, qtk, E)], lovb * (-c6 * (wqh + wql)));
    aadd(&edTb[pidx(g, qi, qj, qk, E)], lovb * (c6 * (wqh + wql)));
    aadd(&uadTb[pidx(g, i, j, k)], hivb * (-c6 * (eph - epl)));
    aadd(&uadTb[pidx(g, pti, ptj, ptk)], hivb * (-c6 * (eph - epl)));
    aadd(&edTb[pidx(g, pti, ptj, ptk, E)], hivb * (-c6 * (wph + wpl)));
    aadd(&edTb[pidx(g, i, j, k, E)], hivb * (c6 * (wph + wpl)));
}

// normal traces of direction D for all three components, advected by the
// normal component of each cell; uad_D = burgers(lo_D, hi_D);
// ed_n = upwind(lo_n, hi_n, uad_D)
template <int D>
__global__ void __launch_bounds__(256) k_adj_faceA_vel(PGeom g, FV vel, FV loD, FV hiD, double dtdx, int lim,
                                                       const double* __restrict__ lob, const double* __restrict__ hib,
                                                       const double* __restrict__ edb,
                                                       const double* __restrict__ uadb, double* __restrict__ velb)
{
    using D8 = Dual<8>;
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const int nD = D == 0 ? g.n0 : (D == 1 ? g.n1 : g.n2);
    const int x = D == 0 ? i : (D == 1 ? j : k);
    int ci[6], cj[6], ck[6];
    for (int m = 0; m < 6; ++m) {
        ci[m] = i;
        cj[m] = j;
        ck[m] = k;
        shift<D>(ci[m], cj[m], ck[m], m - 3);
    }
    // advecting velocities: normal component of the low (c[2]) and high (c[3]) cells
    const D8 aq = dual_seed<8>(rd(vel, g, ci[2], cj[2], ck[2], D), 6);
    const D8 ap = dual_seed<8>(rd(vel, g, ci[3], cj[3], ck[3], D), 7);
    const double loDD = rd(loD, g, i, j, k, D), hiDD = rd(hiD, g, i, j, k, D);
    const double uad = burgers(loDD, hiDD);
    double blo, bhi, dlo, dhi;
    burgers_partials(loDD, hiDD, blo, bhi);
    upwind_partials(uad, dlo, dhi);
    const double ub = uadb[pidx(g, i, j, k)];
    double aqb = 0.0, apb = 0.0;
    for (int n = 0; n < 3; ++n) {
        const int f = pidx(g, i, j, k, n);
        double lovb = lob[f] + dlo * edb[f];
        double hivb = hib[f] + dhi * edb[f];
        if (n == D) {
            lovb += blo * ub;
            hivb += bhi * ub;
        }
        if (lovb == 0.0 && hivb == 0.0) continue;
        D8 s[6];
        for (int m = 0; m < 6; ++m) s[m] = dual_seed<8>(rd(vel, g, ci[m], cj[m], ck[m], n), m);
        D8 dmy, lov, hiv;
        ppm_trace(s[0], s[1], s[2], s[3], s[4], lim, KSGF_BC_INT_DIR, KSGF_BC_INT_DIR, x - 1, 0, nD - 1, aq, aq, dtdx,
                  dmy, lov);
        ppm_trace(s[1], s[2], s[3], s[4], s[5], lim, KSGF_BC_INT_DIR, KSGF_BC_INT_DIR, x, 0, nD - 1, ap, ap, dtdx, hiv,
                  dmy);
        for (int m = 0; m < 6; ++m) aadd(&velb[pidx(g, ci[m], cj[m], ck[m], n)], lovb * lov.d[m] + hivb * hiv.d[m]);
        aqb += lovb * lov.d[6] + hivb * hiv.d[6];
        apb += lovb * lov.d[7] + hivb * hiv.d[7];
    }
    aadd(&velb[pidx(g, ci[2], cj[2], ck[2], D)], aqb);
    aadd(&velb[pidx(g, ci[3], cj[3], ck[3], D)], apb);
}

// ---------------------------------------------------------------------------
// layout moves between periodic arrays and one-box MultiFabs
// ---------------------------------------------------------------------------
// periodic array (ncomp) -> FAB with ghosts (every FAB point, by wrapping)
__global__ void __launch_bounds__(256) k_periodic_to_fab(const double* __restrict__ a, double* __restrict__ fab,
                                                         PGeom g, int ng, int nd, int ncomp)
{
    const int m0 = g.n0 + 2 * ng + (nd == 0), m1 = g.n1 + 2 * ng + (nd == 1), m2 = g.n2 + 2 * ng + (nd == 2);
    const int cs = m0 * m1 * m2;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= cs) return;
    const int q = t / m0;
    const int i = t - q * m0 - ng;
    const int k = q / m1 - ng;
    const int j = q - (k + ng) * m1 - ng;
    for (int n = 0; n < ncomp; ++n) fab[t + n * cs] = a[pidx(g, i, j, k, n)];
}

// FAB valid points (cells, or the n_d distinct faces) -> periodic array
__global__ void __launch_bounds__(256) k_fab_to_periodic(const double* __restrict__ fab, double* __restrict__ a,
                                                         PGeom g, int ng, int nd, int ncomp)
{
    int i, j, k;
    if (!point_of(g, i, j, k)) return;
    const int m0 = g.n0 + 2 * ng + (nd == 0), m1 = g.n1 + 2 * ng + (nd == 1), m2 = g.n2 + 2 * ng + (nd == 2);
    const int cs = m0 * m1 * m2;
    for (int n = 0; n < ncomp; ++n)
        a[pidx(g, i, j, k, n)] = fab[(i + ng) + m0 * ((j + ng) + m1 * (k + ng)) + n * cs];
}

__global__ void __launch_bounds__(256) k_axpy(double* __restrict__ y, const double* __restrict__ x, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] += x[t];
}

}  // namespace ksgf

using namespace ksgf;

struct ksgf_adjoint {
    PGeom g{};
    double dx[3];
    int lim = KSGF_PPM_VANLEER;
    double rtol = 1e-11, atol = 1e-14;
    ksgf_boxarray* ba = nullptr;
    ksgf_macproj* mp = nullptr;
    ksgf_bcrec bc3[3], bc1[1];
    int iconserv[3] = {1, 1, 1};
    // one-box MultiFabs of the forward step
    double *vel = nullptr, *vel_new = nullptr, *s = nullptr, *s_new = nullptr;
    double *f3 = nullptr, *f1 = nullptr, *rho = nullptr, *src3 = nullptr, *src1 = nullptr;
    double* mac[3] = {nullptr, nullptr, nullptr};
    double *flux[3] = {nullptr, nullptr, nullptr}, *face[3] = {nullptr, nullptr, nullptr};
    double *flux_s[3] = {nullptr, nullptr, nullptr}, *face_s[3] = {nullptr, nullptr, nullptr};
    double *conv = nullptr, *conv_s = nullptr;
    // checkpoints and the states of the segment being reversed (periodic arrays)
    int nsteps = 0, every = 1;
    double dt = 0.0;
    std::vector<double*> ck;   // 4N each: u (3N), s (N)
    std::vector<double*> seg;  // every entries, 4N each
    // adjoint arrays (periodic)
    double *ub = nullptr, *ub2 = nullptr, *sb = nullptr, *sb2 = nullptr;  // 3N, 3N, N, N
    double *macb = nullptr, *macpb = nullptr;                             // 3N each (dir-major)
    double* fb = nullptr;                                                 // 3 dirs x 3 comps x N
    double* stage = nullptr;  // lo, hi, ed (3 dirs x 3 comps), crn (3x3 x 3 comps), uad (3): 48 N
    size_t stage_len = 0;
    std::vector<void*> owned;
    int last_iters = 0;
};

namespace ksgf {

static int adj_alloc(ksgf_adjoint* a, double** p, size_t n)
{
    if (hipMalloc((void**)p, sizeof(double) * n) != hipSuccess) {
        set_error("ksgf_adjoint: hipMalloc");
        return KSGF_ERR_NOMEM;
    }
    a->owned.push_back(*p);
    return KSGF_OK;
}

static long long mfsize(const ksgf_boxarray* ba, int itype, int ncomp, int ngrow)
{
    return ksgf_multifab_size(ba, itype, ncomp, ngrow);
}

static dim3 grid_n(long long n)
{
    return dim3((unsigned)grid_for(n, 256));
}

// periodic array -> one-box MultiFab (all points, ghosts included)
static void to_fab(ksgf_adjoint* a, const double* src, double* fab, int ng, int nd, int ncomp, hipStream_t st)
{
    const PGeom& g = a->g;
    const long long cs = (long long)(g.n0 + 2 * ng + (nd == 0)) * (g.n1 + 2 * ng + (nd == 1)) * (g.n2 + 2 * ng + (nd == 2));
    hipLaunchKernelGGL(k_periodic_to_fab, grid_n(cs), dim3(256), 0, st, src, fab, g, ng, nd, ncomp);
}

static void from_fab(ksgf_adjoint* a, const double* fab, double* dst, int ng, int nd, int ncomp, hipStream_t st)
{
    hipLaunchKernelGGL(k_fab_to_periodic, grid_n(a->g.N), dim3(256), 0, st, fab, dst, a->g, ng, nd, ncomp);
}

// u_mac = P ExtrapVelToFaces(vel), ghosts filled; vel holds the state
static int adj_mac(ksgf_adjoint* a, hipStream_t st, GodunovStages* keep)
{
    int e;
    const ksgf_stream_t s = (ksgf_stream_t)st;
    if ((e = extrap_vel_to_faces_keep(a->ba, a->vel, a->f3, a->mac[0], a->mac[1], a->mac[2], a->bc3, a->dt, 1, 0,
                                      a->lim, 0, s, keep)))
        return e;
    if (keep) return KSGF_OK;  // the caller only needs the stages
    ksgf_mlmg_info info;
    e = ksgf_macproj_project(a->mp, a->mac[0], a->mac[1], a->mac[2], a->rtol, a->atol, &info, s);
    if (e < 0) return e;
    a->last_iters = info.iters;
    for (int d = 0; d < 3; ++d)
        if ((e = ksgf_fill_boundary(a->ba, a->mac[d], KSGF_XFACE + d, 1, 1, s))) return e;
    return KSGF_OK;
}

// one forward step of (vel, s) -> (vel_new, s_new), swapped into (vel, s)
static int adj_step(ksgf_adjoint* a, hipStream_t st)
{
    int e;
    const ksgf_stream_t s = (ksgf_stream_t)st;
    if ((e = adj_mac(a, st, nullptr))) return e;
    // passive scalar (density 1: rho s = s)
    if ((e = ksgf_compute_fluxes_from_state(a->ba, 1, a->s, a->s, a->flux_s[0], a->flux_s[1], a->flux_s[2],
                                            a->face_s[0], a->face_s[1], a->face_s[2], 0, a->mac[0], a->mac[1],
                                            a->mac[2], nullptr, a->f1, a->dt, a->bc1, a->iconserv, 1, 0, 0, 0, a->lim,
                                            0, s)))
        return e;
    if ((e = ksgf_compute_divergence(a->ba, a->conv_s, a->flux_s[0], a->flux_s[1], a->flux_s[2], 1, -1.0, 0, s)))
        return e;
    if ((e = ksgf_predictor_update(a->ba, a->s_new, 3, a->s, 3, a->rho, 3, a->rho, 3, a->conv_s, a->src1, nullptr, 0.0,
                                   a->dt, 1, s)))
        return e;
    if ((e = ksgf_fill_boundary(a->ba, a->s_new, KSGF_CELL, 1, 3, s))) return e;
    // momentum (density 1: rho u = u)
    if ((e = ksgf_compute_fluxes_from_state(a->ba, 3, a->vel, a->vel, a->flux[0], a->flux[1], a->flux[2], a->face[0],
                                            a->face[1], a->face[2], 0, a->mac[0], a->mac[1], a->mac[2], nullptr,
                                            a->f3, a->dt, a->bc3, a->iconserv, 1, 0, 1, 0, a->lim, 0, s)))
        return e;
    if ((e = ksgf_compute_divergence(a->ba, a->conv, a->flux[0], a->flux[1], a->flux[2], 3, -1.0, 0, s))) return e;
    if ((e = ksgf_predictor_update(a->ba, a->vel_new, 3, a->vel, 3, a->rho, 3, a->rho, 3, a->conv, a->src3, nullptr,
                                   0.0, a->dt, 3, s)))
        return e;
    if ((e = ksgf_fill_boundary(a->ba, a->vel_new, KSGF_CELL, 3, 3, s))) return e;
    std::swap(a->vel, a->vel_new);
    std::swap(a->s, a->s_new);
    return KSGF_OK;
}

// state (periodic, 4N: u then s) <-> MultiFabs
static int load_state(ksgf_adjoint* a, const double* st4, hipStream_t st)
{
    to_fab(a, st4, a->vel, 3, -1, 3, st);
    to_fab(a, st4 + 3LL * a->g.N, a->s, 3, -1, 1, st);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}
static int save_state(ksgf_adjoint* a, double* st4, hipStream_t st)
{
    from_fab(a, a->vel, st4, 3, -1, 3, st);
    from_fab(a, a->s, st4 + 3LL * a->g.N, 3, -1, 1, st);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

struct StageAdj {
    double *lo[3], *hi[3], *ed[3], *crn[3][3], *uad[3];
};

// stage adjoints carved from a->stage: lo, hi, ed (3 directions x 3
// components), crn (the 6 normal/transverse pairs x 3 components), uad (3
// directions): 27 + 18 + 3 = kStageWords N doubles
constexpr int kStageWords = 48;

static int stage_adj(ksgf_adjoint* a, StageAdj& S)
{
    const long long N = a->g.N;
    double* p = a->stage;
    for (int d = 0; d < 3; ++d) {
        S.lo[d] = p;
        p += 3 * N;
        S.hi[d] = p;
        p += 3 * N;
        S.ed[d] = p;
        p += 3 * N;
    }
    for (int d = 0; d < 3; ++d)
        for (int t = 0; t < 3; ++t) {
            S.crn[d][t] = nullptr;
            if (t == d) continue;
            S.crn[d][t] = p;
            p += 3 * N;
        }
    for (int d = 0; d < 3; ++d) {
        S.uad[d] = p;
        p += N;
    }
    if ((size_t)(p - a->stage) > a->stage_len) {
        set_error("ksgf_adjoint: stage adjoint storage too small");
        return KSGF_ERR_INVALID;
    }
    return KSGF_OK;
}

static double dxi(const ksgf_adjoint* a, int d)
{
    return 1.0 / a->ba->dx[d];
}

// reverse of conv = -div F and x' = x + dt conv: fb from xb (ncomp)
static void adj_update(ksgf_adjoint* a, const double* xb, int ncomp, hipStream_t st)
{
    const PGeom& g = a->g;
    const long long N = g.N;
    // dx'/dF_d(f) = dt * (-1) * dxi * (+1 for the cell below f, -1 above) -> c = -dt dxi
    hipLaunchKernelGGL(k_adj_flux_from_update<0>, grid_n(N), dim3(256), 0, st, xb, a->fb, ncomp, g, -a->dt * dxi(a, 0));
    hipLaunchKernelGGL(k_adj_flux_from_update<1>, grid_n(N), dim3(256), 0, st, xb, a->fb + 3 * N, ncomp, g,
                       -a->dt * dxi(a, 1));
    hipLaunchKernelGGL(k_adj_flux_from_update<2>, grid_n(N), dim3(256), 0, st, xb, a->fb + 6 * N, ncomp, g,
                       -a->dt * dxi(a, 2));
}

// reverse of the conservative PPM fluxes of q (ncomp) by the MAC velocities:
// fb (3 dirs x ncomp) -> qb (+=), macb (+=)
static int adj_fluxes(ksgf_adjoint* a, const GodunovStages& S, const double* qfab, double* const* facep, int ncomp,
                      double* qb, hipStream_t st)
{
    const PGeom& g = a->g;
    const long long N = g.N;
    StageAdj B;
    if (int e = stage_adj(a, B)) return e;
    KSGF_HIP_CHECK(hipMemsetAsync(a->stage, 0, sizeof(double) * a->stage_len, st));
    FV macv[3], lov[3], hiv[3], edv[3], facev[3], crnv[3][3];
    for (int d = 0; d < 3; ++d) {
        macv[d] = make_fv(a->mac[d], g, 1, d);
        lov[d] = make_fv(S.lo[d], g);
        hiv[d] = make_fv(S.hi[d], g);
        edv[d] = make_fv(S.ed[d], g);
        facev[d] = make_fv(facep[d], g, 0, d);
        for (int t = 0; t < 3; ++t)
            if (t != d) crnv[d][t] = make_fv(S.crn[d][t], g);
    }
    const FV qv = make_fv(qfab, g, 3, -1);
    double c2[3], c3[3], dtdx[3];
    for (int d = 0; d < 3; ++d) {
        c2[d] = a->dt / (2.0 * a->ba->dx[d]);
        c3[d] = a->dt / (3.0 * a->ba->dx[d]);
        dtdx[d] = a->dt / a->ba->dx[d];
    }
    double* mb[3] = {a->macb, a->macb + N, a->macb + 2 * N};
    const dim3 gr = grid_n(N), bl(256);
    // finals
    hipLaunchKernelGGL(k_adj_final<0>, gr, bl, 0, st, g, ncomp, a->fb, macv[0], macv[1], macv[2], qv, facev[0],
                       crnv[1][2], crnv[2][1], c2[1], c2[2], B.lo[0], B.hi[0], B.crn[1][2], B.crn[2][1], mb[0], mb[1],
                       mb[2], qb);
    hipLaunchKernelGGL(k_adj_final<1>, gr, bl, 0, st, g, ncomp, a->fb + 3 * N, macv[0], macv[1], macv[2], qv, facev[1],
                       crnv[2][0], crnv[0][2], c2[2], c2[0], B.lo[1], B.hi[1], B.crn[2][0], B.crn[0][2], mb[0], mb[1],
                       mb[2], qb);
    hipLaunchKernelGGL(k_adj_final<2>, gr, bl, 0, st, g, ncomp, a->fb + 6 * N, macv[0], macv[1], macv[2], qv, facev[2],
                       crnv[0][1], crnv[1][0], c2[0], c2[1], B.lo[2], B.hi[2], B.crn[0][1], B.crn[1][0], mb[0], mb[1],
                       mb[2], qb);
    // corner couplings
#define KSGF_ADJ_CORNER(D, T)                                                                                         \
    hipLaunchKernelGGL((k_adj_corner<D, T>), gr, bl, 0, st, g, ncomp, B.crn[D][T], macv[T], macv[D], qv, edv[T],   \
                       c3[T], B.lo[D], B.hi[D], B.ed[T], mb[T], qb)
    KSGF_ADJ_CORNER(0, 1);
    KSGF_ADJ_CORNER(0, 2);
    KSGF_ADJ_CORNER(1, 0);
    KSGF_ADJ_CORNER(1, 2);
    KSGF_ADJ_CORNER(2, 0);
    KSGF_ADJ_CORNER(2, 1);
#undef KSGF_ADJ_CORNER
    // normal traces
    hipLaunchKernelGGL(k_adj_faceA<0>, gr, bl, 0, st, g, ncomp, qv, macv[0], dtdx[0], a->lim, B.lo[0], B.hi[0], B.ed[0],
                       mb[0], qb);
    hipLaunchKernelGGL(k_adj_faceA<1>, gr, bl, 0, st, g, ncomp, qv, macv[1], dtdx[1], a->lim, B.lo[1], B.hi[1], B.ed[1],
                       mb[1], qb);
    hipLaunchKernelGGL(k_adj_faceA<2>, gr, bl, 0, st, g, ncomp, qv, macv[2], dtdx[2], a->lim, B.lo[2], B.hi[2], B.ed[2],
                       mb[2], qb);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

// reverse of ExtrapVelToFaces: macpb (3 dirs) -> velb (+=)
static int adj_extrap(ksgf_adjoint* a, const GodunovStages& S, double* velb, hipStream_t st)
{
    const PGeom& g = a->g;
    const long long N = g.N;
    StageAdj B;
    if (int e = stage_adj(a, B)) return e;
    KSGF_HIP_CHECK(hipMemsetAsync(a->stage, 0, sizeof(double) * a->stage_len, st));
    FV lov[3], hiv[3], edv[3], uadv[3], crnv[3][3];
    for (int d = 0; d < 3; ++d) {
        lov[d] = make_fv(S.lo[d], g);
        hiv[d] = make_fv(S.hi[d], g);
        edv[d] = make_fv(S.ed[d], g);
        uadv[d] = make_fv(S.uad[d], g);
        for (int t = 0; t < 3; ++t)
            if (t != d) crnv[d][t] = make_fv(S.crn[d][t], g);
    }
    const FV velv = make_fv(a->vel, g, 3, -1);
    double c4[3], c6[3], dtdx[3];
    for (int d = 0; d < 3; ++d) {
        c4[d] = a->dt / (4.0 * a->ba->dx[d]);
        c6[d] = a->dt / (6.0 * a->ba->dx[d]);
        dtdx[d] = a->dt / a->ba->dx[d];
    }
    const dim3 gr = grid_n(N), bl(256);
    hipLaunchKernelGGL(k_adj_final_vel<0>, gr, bl, 0, st, g, a->macpb, lov[0], hiv[0], uadv[1], uadv[2], crnv[1][2],
                       crnv[2][1], c4[1], c4[2], B.lo[0], B.hi[0], B.uad[1], B.uad[2], B.crn[1][2], B.crn[2][1]);
    hipLaunchKernelGGL(k_adj_final_vel<1>, gr, bl, 0, st, g, a->macpb + N, lov[1], hiv[1], uadv[2], uadv[0], crnv[2][0],
                       crnv[0][2], c4[2], c4[0], B.lo[1], B.hi[1], B.uad[2], B.uad[0], B.crn[2][0], B.crn[0][2]);
    hipLaunchKernelGGL(k_adj_final_vel<2>, gr, bl, 0, st, g, a->macpb + 2 * N, lov[2], hiv[2], uadv[0], uadv[1],
                       crnv[0][1], crnv[1][0], c4[0], c4[1], B.lo[2], B.hi[2], B.uad[0], B.uad[1], B.crn[0][1],
                       B.crn[1][0]);
#define KSGF_ADJ_CORNER_VEL(D, T)                                                                                    \
    hipLaunchKernelGGL((k_adj_corner_vel<D, T>), gr, bl, 0, st, g, B.crn[D][T], uadv[T], uadv[D], edv[T], c6[T],  \
                       B.lo[D], B.hi[D], B.ed[T], B.uad[T])
    KSGF_ADJ_CORNER_VEL(0, 1);
    KSGF_ADJ_CORNER_VEL(0, 2);
    KSGF_ADJ_CORNER_VEL(1, 0);
    KSGF_ADJ_CORNER_VEL(1, 2);
    KSGF_ADJ_CORNER_VEL(2, 0);
    KSGF_ADJ_CORNER_VEL(2, 1);
#undef KSGF_ADJ_CORNER_VEL
    hipLaunchKernelGGL(k_adj_faceA_vel<0>, gr, bl, 0, st, g, velv, lov[0], hiv[0], dtdx[0], a->lim, B.lo[0], B.hi[0],
                       B.ed[0], B.uad[0], velb);
    hipLaunchKernelGGL(k_adj_faceA_vel<1>, gr, bl, 0, st, g, velv, lov[1], hiv[1], dtdx[1], a->lim, B.lo[1], B.hi[1],
                       B.ed[1], B.uad[1], velb);
    hipLaunchKernelGGL(k_adj_faceA_vel<2>, gr, bl, 0, st, g, velv, lov[2], hiv[2], dtdx[2], a->lim, B.lo[2], B.hi[2],
                       B.ed[2], B.uad[2], velb);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

// reverse of one step from the state (vel, s) of its start, loaded in the
// MultiFabs: (ub, sb) = adjoint of the step's output -> (ub2, sb2) =
// adjoint of its input
static int adj_reverse_step(ksgf_adjoint* a, hipStream_t st)
{
    int e;
    const PGeom& g = a->g;
    const long long N = g.N;
    const ksgf_stream_t s = (ksgf_stream_t)st;
    // identity parts of u' = u + dt conv and s' = s + dt conv_s
    KSGF_HIP_CHECK(hipMemcpyAsync(a->ub2, a->ub, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, st));
    KSGF_HIP_CHECK(hipMemcpyAsync(a->sb2, a->sb, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    KSGF_HIP_CHECK(hipMemsetAsync(a->macb, 0, sizeof(double) * 3 * N, st));
    // the projected MAC velocities of this step
    if ((e = adj_mac(a, st, nullptr))) return e;
    // momentum fluxes
    GodunovStages S;
    if ((e = compute_fluxes_keep(a->ba, 3, a->vel, a->vel, a->flux[0], a->flux[1], a->flux[2], a->face[0], a->face[1],
                                 a->face[2], 0, a->mac[0], a->mac[1], a->mac[2], nullptr, a->f3, a->dt, a->bc3,
                                 a->iconserv, 1, 0, 1, 0, a->lim, 0, s, &S)))
        return e;
    adj_update(a, a->ub, 3, st);
    if ((e = adj_fluxes(a, S, a->vel, a->face, 3, a->ub2, st))) return e;
    // scalar fluxes
    if ((e = compute_fluxes_keep(a->ba, 1, a->s, a->s, a->flux_s[0], a->flux_s[1], a->flux_s[2], a->face_s[0],
                                 a->face_s[1], a->face_s[2], 0, a->mac[0], a->mac[1], a->mac[2], nullptr, a->f1, a->dt,
                                 a->bc1, a->iconserv, 1, 0, 0, 0, a->lim, 0, s, &S)))
        return e;
    adj_update(a, a->sb, 1, st);
    if ((e = adj_fluxes(a, S, a->s, a->face_s, 1, a->sb2, st))) return e;
    // projection: macpb = P macb (the projection is symmetric)
    for (int d = 0; d < 3; ++d) to_fab(a, a->macb + d * N, a->mac[d], 1, d, 1, st);
    KSGF_LAUNCH_CHECK();
    // relative tolerance only: the adjoint field has no natural absolute scale
    ksgf_mlmg_info info;
    e = ksgf_macproj_project(a->mp, a->mac[0], a->mac[1], a->mac[2], a->rtol, 0.0, &info, s);
    if (e < 0) return e;
    for (int d = 0; d < 3; ++d) from_fab(a, a->mac[d], a->macpb + d * N, 1, d, 1, st);
    // ExtrapVelToFaces (its stages; the MAC arrays are scratch from here)
    if ((e = adj_mac(a, st, &S))) return e;
    if ((e = adj_extrap(a, S, a->ub2, st))) return e;
    std::swap(a->ub, a->ub2);
    std::swap(a->sb, a->sb2);
    return KSGF_OK;
}

}  // namespace ksgf

extern "C" {

int ksgf_adjoint_destroy(ksgf_adjoint* a)
{
    if (!a) return KSGF_OK;
    for (void* p : a->owned) (void)hipFree(p);
    if (a->mp) ksgf_macproj_destroy(a->mp);
    if (a->ba) ksgf_boxarray_destroy(a->ba);
    delete a;
    return KSGF_OK;
}

int ksgf_adjoint_create(const int* n, const double* prob_hi, int limiter_type, double mg_rtol, double mg_atol,
                        const ksgf_mlmg_options* opts, ksgf_adjoint** out)
{
    if (!n || !prob_hi || !out) {
        set_error("ksgf_adjoint_create: null argument");
        return KSGF_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d)
        if (n[d] < 8 || prob_hi[d] <= 0.0) {
            set_error("ksgf_adjoint_create: every extent must be at least 8 cells and the domain positive");
            return KSGF_ERR_INVALID;
        }
    if (limiter_type != KSGF_PPM_VANLEER && limiter_type != KSGF_PPM_NOLIMITER && limiter_type != KSGF_PPM_UPWIND &&
        limiter_type != KSGF_PPM_MINMOD) {
        set_error("ksgf_adjoint_create: unsupported limiter");
        return KSGF_ERR_UNSUPPORTED;
    }
    if ((long long)n[0] * n[1] * n[2] >= (1LL << 30)) {
        set_error("ksgf_adjoint_create: domain too large");
        return KSGF_ERR_UNSUPPORTED;
    }
    auto* a = new ksgf_adjoint();
    a->g.n0 = n[0];
    a->g.n1 = n[1];
    a->g.n2 = n[2];
    a->g.N = n[0] * n[1] * n[2];
    a->lim = limiter_type;
    a->rtol = mg_rtol;
    a->atol = mg_atol;
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) a->bc3[c].lo[d] = a->bc3[c].hi[d] = KSGF_BC_INT_DIR;
    a->bc1[0] = a->bc3[0];
    ksgf_geometry geo{};
    int box[6];
    for (int d = 0; d < 3; ++d) {
        geo.domain_lo[d] = 0;
        geo.domain_hi[d] = n[d] - 1;
        geo.prob_lo[d] = 0.0;
        geo.prob_hi[d] = prob_hi[d];
        geo.is_periodic[d] = 1;
        box[d] = 0;
        box[3 + d] = n[d] - 1;
    }
    int e;
    if ((e = ksgf_boxarray_create(&geo, box, 1, &a->ba))) {
        delete a;
        return e;
    }
    const int per[3] = {KSGF_LO_PERIODIC, KSGF_LO_PERIODIC, KSGF_LO_PERIODIC};
    ksgf_mlmg_options o;
    if (opts)
        o = *opts;
    else
        ksgf_mlmg_default_options(&o);
    if ((e = ksgf_macproj_create(a->ba, per, per, 1.0, &o, &a->mp))) {
        ksgf_adjoint_destroy(a);
        return e;
    }
    const ksgf_boxarray* ba = a->ba;
    const size_t N = (size_t)a->g.N;
    bool ok = !adj_alloc(a, &a->vel, mfsize(ba, KSGF_CELL, 3, 3)) && !adj_alloc(a, &a->vel_new, mfsize(ba, KSGF_CELL, 3, 3)) &&
              !adj_alloc(a, &a->s, mfsize(ba, KSGF_CELL, 1, 3)) && !adj_alloc(a, &a->s_new, mfsize(ba, KSGF_CELL, 1, 3)) &&
              !adj_alloc(a, &a->f3, mfsize(ba, KSGF_CELL, 3, 1)) && !adj_alloc(a, &a->f1, mfsize(ba, KSGF_CELL, 1, 1)) &&
              !adj_alloc(a, &a->rho, mfsize(ba, KSGF_CELL, 1, 3)) && !adj_alloc(a, &a->src3, mfsize(ba, KSGF_CELL, 3, 0)) &&
              !adj_alloc(a, &a->src1, mfsize(ba, KSGF_CELL, 1, 0)) && !adj_alloc(a, &a->conv, mfsize(ba, KSGF_CELL, 3, 0)) &&
              !adj_alloc(a, &a->conv_s, mfsize(ba, KSGF_CELL, 1, 0));
    for (int d = 0; ok && d < 3; ++d)
        ok = !adj_alloc(a, &a->mac[d], mfsize(ba, KSGF_XFACE + d, 1, 1)) &&
             !adj_alloc(a, &a->flux[d], mfsize(ba, KSGF_XFACE + d, 3, 0)) &&
             !adj_alloc(a, &a->face[d], mfsize(ba, KSGF_XFACE + d, 3, 0)) &&
             !adj_alloc(a, &a->flux_s[d], mfsize(ba, KSGF_XFACE + d, 1, 0)) &&
             !adj_alloc(a, &a->face_s[d], mfsize(ba, KSGF_XFACE + d, 1, 0));
    a->stage_len = kStageWords * N;
    ok = ok && !adj_alloc(a, &a->ub, 3 * N) && !adj_alloc(a, &a->ub2, 3 * N) && !adj_alloc(a, &a->sb, N) &&
         !adj_alloc(a, &a->sb2, N) && !adj_alloc(a, &a->macb, 3 * N) && !adj_alloc(a, &a->macpb, 3 * N) &&
         !adj_alloc(a, &a->fb, 9 * N) && !adj_alloc(a, &a->stage, a->stage_len);
    if (!ok) {
        ksgf_adjoint_destroy(a);
        return KSGF_ERR_NOMEM;
    }
    // density 1, zero forcing and sources
    const long long nr = mfsize(ba, KSGF_CELL, 1, 3);
    std::vector<double> ones((size_t)nr, 1.0);
    if (hipMemcpy(a->rho, ones.data(), sizeof(double) * nr, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(a->f3, 0, sizeof(double) * mfsize(ba, KSGF_CELL, 3, 1)) != hipSuccess ||
        hipMemset(a->f1, 0, sizeof(double) * mfsize(ba, KSGF_CELL, 1, 1)) != hipSuccess ||
        hipMemset(a->src3, 0, sizeof(double) * mfsize(ba, KSGF_CELL, 3, 0)) != hipSuccess ||
        hipMemset(a->src1, 0, sizeof(double) * mfsize(ba, KSGF_CELL, 1, 0)) != hipSuccess) {
        ksgf_adjoint_destroy(a);
        set_error("ksgf_adjoint_create: initialisation");
        return KSGF_ERR_HIP;
    }
    *out = a;
    return KSGF_OK;
}

int ksgf_adjoint_forward(ksgf_adjoint* a, const double* vel0, const double* s0, int nsteps, double dt,
                         int checkpoint_every, double* vel_out, double* s_out, int* mlmg_iters, ksgf_stream_t stream)
{
    if (!a || !vel0 || !s0 || nsteps < 0 || !(dt > 0.0) || checkpoint_every < 1) {
        set_error("ksgf_adjoint_forward: bad argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    const long long N = a->g.N;
    a->nsteps = nsteps;
    a->dt = dt;
    a->every = checkpoint_every;
    const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every + 1;
    while ((int)a->ck.size() < nck) {
        double* p;
        if (adj_alloc(a, &p, 4 * (size_t)N)) return KSGF_ERR_NOMEM;
        a->ck.push_back(p);
    }
    while ((int)a->seg.size() < checkpoint_every) {
        double* p;
        if (adj_alloc(a, &p, 4 * (size_t)N)) return KSGF_ERR_NOMEM;
        a->seg.push_back(p);
    }
    int e;
    KSGF_HIP_CHECK(hipMemcpyAsync(a->ck[0], vel0, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, st));
    KSGF_HIP_CHECK(hipMemcpyAsync(a->ck[0] + 3 * N, s0, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    if ((e = load_state(a, a->ck[0], st))) return e;
    for (int n = 0; n < nsteps; ++n) {
        if ((e = adj_step(a, st))) return e;
        if (mlmg_iters) mlmg_iters[n] = a->last_iters;
        if ((n + 1) % checkpoint_every == 0 && n + 1 < nsteps) {
            if ((e = save_state(a, a->ck[(n + 1) / checkpoint_every], st))) return e;
        }
    }
    if (vel_out) from_fab(a, a->vel, vel_out, 3, -1, 3, st);
    if (s_out) from_fab(a, a->s, s_out, 3, -1, 1, st);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_adjoint_backward(ksgf_adjoint* a, const double* vel_bar, const double* s_bar, double* vel0_bar,
                          double* s0_bar, ksgf_stream_t stream)
{
    if (!a || !vel_bar || !s_bar || !vel0_bar || !s0_bar) {
        set_error("ksgf_adjoint_backward: null argument");
        return KSGF_ERR_INVALID;
    }
    if (a->dt <= 0.0) {
        set_error("ksgf_adjoint_backward: run ksgf_adjoint_forward first");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    const long long N = a->g.N;
    int e;
    KSGF_HIP_CHECK(hipMemcpyAsync(a->ub, vel_bar, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, st));
    KSGF_HIP_CHECK(hipMemcpyAsync(a->sb, s_bar, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    const int K = a->every;
    for (int c0 = ((a->nsteps - 1) / K) * K; c0 >= 0; c0 -= K) {
        const int c1 = std::min(c0 + K, a->nsteps);
        // recompute the states of steps c0 .. c1-1 from checkpoint c0 / K
        if ((e = load_state(a, a->ck[c0 / K], st))) return e;
        for (int n = c0; n < c1; ++n) {
            if ((e = save_state(a, a->seg[n - c0], st))) return e;
            if (n + 1 < c1 && (e = adj_step(a, st))) return e;
        }
        for (int n = c1 - 1; n >= c0; --n) {
            if ((e = load_state(a, a->seg[n - c0], st))) return e;
            if ((e = adj_reverse_step(a, st))) return e;
        }
    }
    KSGF_HIP_CHECK(hipMemcpyAsync(vel0_bar, a->ub, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, st));
    KSGF_HIP_CHECK(hipMemcpyAsync(s0_bar, a->sb, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    return KSGF_OK;
}

}  // extern "C"
