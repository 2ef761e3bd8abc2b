// Godunov PPM advection on a ragged BoxArray:
//   HydroUtils::ExtrapVelToFaces            (call site icns_advection.H:211-216)
//   HydroUtils::ComputeFluxesOnBoxFromState (icns_advection.H:359-373)
//   HydroUtils::ComputeDivergence           (icns_advection.H:418-422)
//   PPM::PredictStateOn{X,Y,Z}Face          (test_mflux_schemes.cpp:87-151)
//
// Two launches per face direction over every box of the BoxArray (flat
// index space, binary search of the box from a prefix table):
//   stage A  PPM traces of the two cells of each face, normal face states,
//            advective velocity and upwinded normal edge states, on faces
//            grown by one cell transversely (kept in scratch MultiFabs);
//   stage B  corner-coupled (transverse-corrected) face states, one
//            launch per (normal, transverse) direction pair;
//   stage C  final face states (and fluxes).
#include "ksgf_internal.h"
#include "ksgf_ppm.h"

namespace ksgf {

constexpr int kMaxComp = 8;

struct BCs {
    int lo[kMaxComp][3];
    int hi[kMaxComp][3];
};
struct GeomDev {
    int dlo[3], dhi[3];
};
struct MF3 {
    MFDev m[3];
};
struct CMF3 {
    CMFDev m[3];
};
struct Coef {
    double v[3];
};

__device__ __forceinline__ CMFDev cview(const MFDev& m)
{
    CMFDev c;
    c.p = m.p;
    c.pre = m.pre;
    c.itype = m.itype;
    c.ngrow = m.ngrow;
    c.ncomp = m.ncomp;
    return c;
}

template <int D>
__device__ __forceinline__ void unit(int& a0, int& a1, int& a2, int s)
{
    if (D == 0) a0 += s;
    if (D == 1) a1 += s;
    if (D == 2) a2 += s;
}
template <int D>
__device__ __forceinline__ int pick(int a0, int a1, int a2)
{
    return D == 0 ? a0 : (D == 1 ? a1 : a2);
}

// ---------------------------------------------------------------------------
// stage A: normal face states on faces of direction D grown by one
// transversely.  The PPM traces of the two cells sharing the face (Ip of
// the low cell, Im of the high cell) are evaluated here from a 6-point
// stencil along D.
//   vel_mode: ExtrapVelToFaces -- the advecting velocity of a cell is its
//             own normal component; the advective velocity uad solves the
//             Burgers problem of the normal component; Dirichlet sets both
//             sides for every component.
//   otherwise: ComputeEdgeState -- the MAC velocity advects and upwinds;
//             is_velocity selects both-side Dirichlet for the normal
//             component only.
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256) k_faceA(CMFDev q, CMFDev qbc, CMFDev f, int ncomp, CMFDev macD, MFDev flo,
                                               MFDev fhi, MFDev fed, MFDev uad, RegDev r,
                                               const int* __restrict__ box, BCs bc, GeomDev g, double dtdx,
                                               double hdt, int uft, int vel_mode, int is_velocity, int lim)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    int qi = i, qj = j, qk = k;
    unit<D>(qi, qj, qk, -1);
    CFab s = fab_at(q, box, b);
    const long long st = (D == 0) ? 1 : (D == 1 ? (long long)s.n0 : (long long)s.n0 * s.n1);
    const long long cp = s.idx(i, j, k);
    CFab ff = fab_at(f, box, b);
    CFab fqb = fab_at(qbc, box, b);
    Fab flo_ = fab_at(flo, box, b);
    Fab fhi_ = fab_at(fhi, box, b);
    Fab fed_ = fab_at(fed, box, b);
    const int x = pick<D>(i, j, k);
    int gli = i, glj = j, glk = k, ghi_ = i, ghj = j, ghk = k;
    if (D == 0) { gli = g.dlo[0] - 1; ghi_ = g.dhi[0] + 1; }
    if (D == 1) { glj = g.dlo[1] - 1; ghj = g.dhi[1] + 1; }
    if (D == 2) { glk = g.dlo[2] - 1; ghk = g.dhi[2] + 1; }
    // advecting velocities of the low cell q and the high cell p
    double uqm, uqp, upm, upp;
    if (vel_mode) {
        uqm = uqp = s.p[cp - st + D * s.cs];
        upm = upp = s.p[cp + D * s.cs];
    } else {
        CFab m = fab_at(macD, box, b);
        int i1 = i, j1 = j, k1 = k;
        unit<D>(i1, j1, k1, 1);
        uqm = m(qi, qj, qk);
        uqp = m(i, j, k);
        upm = uqp;
        upp = m(i1, j1, k1);
    }
    // Ip of the low cell q (stencil x-3 .. x+1) and Im of the high cell p
    // (stencil x-2 .. x+2), sharing their common slopes and edge
    auto face_states = [&](int n, double& lov, double& hiv) {
        const double* pn = s.p + cp + n * s.cs;
        const double w[6] = {pn[-3 * st], pn[-2 * st], pn[-st], pn[0], pn[st], pn[2 * st]};
        ppm_trace_pair(w, lim, bc.lo[n][D], bc.hi[n][D], x, g.dlo[D], g.dhi[D], uqm, uqp, upm, upp, dtdx, lov, hiv);
        if (uft) {
            lov += hdt * ff(qi, qj, qk, n);
            hiv += hdt * ff(i, j, k, n);
        }
    };
    double u;
    double lovD = 0.0, hivD = 0.0;  // velocity mode: the normal component's states, reused below
    if (vel_mode) {
        face_states(D, lovD, hivD);
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = fqb(gli, glj, glk, D);
        if (x == g.dhi[D] + 1) ghi = fqb(ghi_, ghj, ghk, D);
        edge_bcs(x, g.dlo[D], g.dhi[D], bc.lo[D][D], bc.hi[D][D], glo, ghi, true, lovD, hivD);
        u = burgers(lovD, hivD);
        Fab fu = fab_at(uad, box, b);
        fu(i, j, k) = u;
    } else {
        u = uqp;
    }
    for (int n = 0; n < ncomp; ++n) {
        double lov, hiv;
        if (vel_mode && n == D) {
            lov = lovD;
            hiv = hivD;
            flo_(i, j, k, n) = lov;
            fhi_(i, j, k, n) = hiv;
            if (fed.p) fed_(i, j, k, n) = upwind(lov, hiv, u);
            continue;
        }
        face_states(n, lov, hiv);
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = fqb(gli, glj, glk, n);
        if (x == g.dhi[D] + 1) ghi = fqb(ghi_, ghj, ghk, n);
        bool both = vel_mode ? true : (is_velocity && n == D);
        edge_bcs(x, g.dlo[D], g.dhi[D], bc.lo[n][D], bc.hi[n][D], glo, ghi, both, lov, hiv);
        flo_(i, j, k, n) = lov;
        fhi_(i, j, k, n) = hiv;
        if (fed.p) fed_(i, j, k, n) = upwind(lov, hiv, u);
    }
}

// ---------------------------------------------------------------------------
// stage A, marching form (van Leer-limited and unlimited PPM; faces of y and
// z).  One thread owns a segment of kMarch consecutive faces along D of one
// (transverse) face column and walks it: every cell's van Leer slope, interior edge and
// parabola is formed once and serves the two faces of the cell (Ip for the
// face above, Im for the face below), where k_faceA forms them once per
// face side.  The slopes, edges, parabolas and traces are the same
// ppm_* expressions on the same values, so the states are k_faceA's bit for
// bit.
// ---------------------------------------------------------------------------
constexpr int kMarch = 8;

// launch table of the marching kernels: the faces of direction D of every
// box grown by glo / ghi, with the D extent cut into segments of kMarch
struct SegDev {
    const int* chunks;
    int glo0, glo1, glo2, ghi0, ghi1, ghi2;
};

template <int D>
__device__ __forceinline__ bool seg_locate(const SegDev& r, const int* box, int& b, int& i, int& j, int& k, int& len)
{
    b = r.chunks[2 * blockIdx.x];
    const unsigned loc = (unsigned)r.chunks[2 * blockIdx.x + 1] + threadIdx.x;
    const int* bx = box + 6 * b;
    const int l0 = bx[0] - r.glo0, l1 = bx[1] - r.glo1, l2 = bx[2] - r.glo2;
    const int e0 = bx[3] + r.ghi0 - l0 + 1, e1 = bx[4] + r.ghi1 - l1 + 1, e2 = bx[5] + r.ghi2 - l2 + 1;
    const int eD = pick<D>(e0, e1, e2);
    const int ns = (eD + kMarch - 1) / kMarch;
    const unsigned n0 = (D == 0) ? ns : e0, n1 = (D == 1) ? ns : e1, n2 = (D == 2) ? ns : e2;
    if (loc >= n0 * n1 * n2) return false;
    const unsigned q = loc / n0;
    const int a0 = (int)(loc - q * n0), a1 = (int)(q % n1), a2 = (int)(q / n1);
    const int seg = pick<D>(a0, a1, a2);
    i = l0 + ((D == 0) ? seg * kMarch : a0);
    j = l1 + ((D == 1) ? seg * kMarch : a1);
    k = l2 + ((D == 2) ? seg * kMarch : a2);
    len = min(kMarch, eD - seg * kMarch);
    return true;
}

template <int D>
__global__ void __launch_bounds__(256) k_faceA_march(CMFDev q, CMFDev qbc, CMFDev f, int ncomp, CMFDev macD,
                                                     MFDev flo, MFDev fhi, MFDev fed, MFDev uad, SegDev r,
                                                     const int* __restrict__ box, BCs bc, GeomDev g, double dtdx,
                                                     double hdt, int uft, int vel_mode, int is_velocity, int lim)
{
    int b, i, j, k, len;
    if (!seg_locate<D>(r, box, b, i, j, k, len)) return;
    CFab s = fab_at(q, box, b);
    const int st = (D == 0) ? 1 : (D == 1 ? s.n0 : s.n0 * s.n1);
    const int cp = s.idx(i, j, k);  // cell p of the segment's first face
    CFab ff = fab_at(f, box, b);
    const int fst = (D == 0) ? 1 : (D == 1 ? ff.n0 : ff.n0 * ff.n1);
    const int fp = ff.idx(i, j, k);
    CFab fqb = fab_at(qbc, box, b);
    Fab flo_ = fab_at(flo, box, b);
    Fab fhi_ = fab_at(fhi, box, b);
    Fab fed_ = fab_at(fed, box, b);
    const int ost = (D == 0) ? 1 : (D == 1 ? flo_.n0 : flo_.n0 * flo_.n1);
    const int op = flo_.idx(i, j, k);
    const int x0 = pick<D>(i, j, k);
    const bool vl = (lim == KSGF_PPM_VANLEER);
    int gli = i, glj = j, glk = k, ghi_ = i, ghj = j, ghk = k;
    if (D == 0) { gli = g.dlo[0] - 1; ghi_ = g.dhi[0] + 1; }
    if (D == 1) { glj = g.dlo[1] - 1; ghj = g.dhi[1] + 1; }
    if (D == 2) { glk = g.dlo[2] - 1; ghk = g.dhi[2] + 1; }
    // advecting velocity at each face: the MAC velocity, or (velocity mode)
    // the Burgers solution of the normal component, found first
    double uf[kMarch];
    const double* pD = s.p + cp + D * s.cs;  // normal component (velocity mode)
    if (!vel_mode) {
        CFab m = fab_at(macD, box, b);
        const int mst = (D == 0) ? 1 : (D == 1 ? m.n0 : m.n0 * m.n1);
        const int mp = m.idx(i, j, k);
#pragma unroll
        for (int t = 0; t < kMarch; ++t)
            if (t < len) uf[t] = m.p[mp + t * mst];
    }
    for (int tc = 0; tc < ncomp; ++tc) {
        const int n = vel_mode ? ((tc == 0) ? D : ((tc - 1 < D) ? tc - 1 : tc)) : tc;
        const bool normal_first = vel_mode && tc == 0;
        const bool both = vel_mode ? true : (is_velocity && n == D);
        const double* pn = s.p + cp + n * s.cs;
        // cells x0-3 .. x0+1: slopes of x0-2 .. x0, the parabola of q = x0-1
        double s_m3 = pn[-3 * st], s_m2 = pn[-2 * st], s_m1 = pn[-st], s_0 = pn[0], s_p1 = pn[st];
        // edges e(x0-2|x0-1) and e(x0-1|x0); d_0 is the slope of cell x0
        double d_0 = 0.0, e_lo, e_hi;
        if (vl) {
            const double d_m2 = vanleer(s_m3, s_m2, s_m1);
            const double d_m1 = vanleer(s_m2, s_m1, s_0);
            d_0 = vanleer(s_m1, s_0, s_p1);
            e_lo = ppm_edge_vl(s_m2, s_m1, d_m2, d_m1);
            e_hi = ppm_edge_vl(s_m1, s_0, d_m1, d_0);
        } else {
            e_lo = ppm_edge4(s_m3, s_m2, s_m1, s_0);
            e_hi = ppm_edge4(s_m2, s_m1, s_0, s_p1);
        }
        double smq, spq;
        ppm_edges_finish(e_lo, e_hi, s_m3, s_m2, s_m1, s_0, s_p1, lim, bc.lo[n][D], bc.hi[n][D], x0 - 1, g.dlo[D],
                         g.dhi[D], smq, spq);
        double s_q = s_m1;
        double uq = vel_mode ? pD[-st] : 0.0;  // velocity mode: the low cell's normal velocity
        double fq_q = uft ? ff.p[fp - fst + n * ff.cs] : 0.0;
        // window for face x: cells x-2 .. x+1 in (w_m2, w_m1, w_0, w_p1)
        double w_m2 = s_m2, w_m1 = s_m1, w_0 = s_0, w_p1 = s_p1;
        double e_q_hi = e_hi;  // edge x-1|x
#pragma unroll
        for (int t = 0; t < kMarch; ++t) {
            if (t >= len) break;
            const int x = x0 + t;
            const double w_p2 = pn[(t + 2) * st];
            double e_p_hi;
            if (vl) {
                const double d_p1 = vanleer(w_0, w_p1, w_p2);
                e_p_hi = ppm_edge_vl(w_0, w_p1, d_0, d_p1);
                d_0 = d_p1;
            } else {
                e_p_hi = ppm_edge4(w_m1, w_0, w_p1, w_p2);
            }
            double smp, spp;
            ppm_edges_finish(e_q_hi, e_p_hi, w_m2, w_m1, w_0, w_p1, w_p2, lim, bc.lo[n][D], bc.hi[n][D], x, g.dlo[D],
                             g.dhi[D], smp, spp);
            const double up = vel_mode ? pD[t * st] : 0.0;
            const double u_q = vel_mode ? uq : uf[t];
            const double u_p = vel_mode ? up : uf[t];
            double lov = ppm_ip(smq, s_q, spq, u_q, dtdx);
            double hiv = ppm_im(smp, w_0, spp, u_p, dtdx);
            const double fq_p = uft ? ff.p[fp + t * fst + n * ff.cs] : 0.0;
            if (uft) {
                lov += hdt * fq_q;
                hiv += hdt * fq_p;
            }
            double glo = 0.0, ghi = 0.0;
            int gi = i, gj = j, gk = k;
            unit<D>(gi, gj, gk, t);
            if (x == g.dlo[D]) {
                int a = gi, c = gj, e = gk;
                if (D == 0) a = gli;
                if (D == 1) c = glj;
                if (D == 2) e = glk;
                glo = fqb(a, c, e, n);
            }
            if (x == g.dhi[D] + 1) {
                int a = gi, c = gj, e = gk;
                if (D == 0) a = ghi_;
                if (D == 1) c = ghj;
                if (D == 2) e = ghk;
                ghi = fqb(a, c, e, n);
            }
            edge_bcs(x, g.dlo[D], g.dhi[D], bc.lo[n][D], bc.hi[n][D], glo, ghi, both, lov, hiv);
            double u;
            if (normal_first) {
                u = burgers(lov, hiv);
                uf[t] = u;
                Fab fu = fab_at(uad, box, b);
                fu.p[fu.idx(i, j, k) + t * ost] = u;
            } else {
                u = uf[t];
            }
            const int o = op + t * ost;
            flo_.p[o + n * flo_.cs] = lov;
            fhi_.p[o + n * fhi_.cs] = hiv;
            if (fed.p) fed_.p[o + n * fed_.cs] = upwind(lov, hiv, u);
            // shift: p becomes the next face's q
            smq = smp;
            spq = spp;
            s_q = w_0;
            uq = up;
            fq_q = fq_p;
            e_q_hi = e_p_hi;
            w_m2 = w_m1;
            w_m1 = w_0;
            w_0 = w_p1;
            w_p1 = w_p2;
        }
    }
}

// ---------------------------------------------------------------------------
// stage B: corner-coupled states on faces normal to D corrected along T.
// vel_mode: only component E = 3-D-T, advective form with the predicted
// advective velocities (dt/6 factor);  otherwise all components with the
// MAC velocities, conservative (dt/3, flux difference + q * transverse
// divergence) or advective (dt/6) per component.
// ---------------------------------------------------------------------------
template <int D, int T>
__global__ void __launch_bounds__(256) k_corner(CMFDev q, CMFDev qbc, int ncomp, CMFDev flo, CMFDev fhi,
                                                CMFDev fedT, CMFDev velT, CMFDev velD, MFDev crn, RegDev r,
                                                const int* __restrict__ box, BCs bc, GeomDev g, double c3,
                                                double c6, int vel_mode, int is_velocity,
                                                const int* __restrict__ iconserv_unused, int cons_mask)
{
    constexpr int E = 3 - D - T;
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    int qi = i, qj = j, qk = k;
    unit<D>(qi, qj, qk, -1);
    int qti = qi, qtj = qj, qtk = qk;
    unit<T>(qti, qtj, qtk, 1);
    int pti = i, ptj = j, ptk = k;
    unit<T>(pti, ptj, ptk, 1);
    CFab vt = fab_at(velT, box, b);
    CFab et = fab_at(fedT, box, b);
    CFab lo_ = fab_at(flo, box, b);
    CFab hi_ = fab_at(fhi, box, b);
    CFab vd = fab_at(velD, box, b);
    CFab fqb = fab_at(qbc, box, b);
    Fab out = fab_at(crn, box, b);
    const int x = pick<D>(i, j, k);
    int gli = i, glj = j, glk = k, ghi_ = i, ghj = j, ghk = k;
    if (D == 0) { gli = g.dlo[0] - 1; ghi_ = g.dhi[0] + 1; }
    if (D == 1) { glj = g.dlo[1] - 1; ghj = g.dhi[1] + 1; }
    if (D == 2) { glk = g.dlo[2] - 1; ghk = g.dhi[2] + 1; }
    const double wql = vt(qi, qj, qk), wqh = vt(qti, qtj, qtk);
    const double wpl = vt(i, j, k), wph = vt(pti, ptj, ptk);
    const double u = vd(i, j, k);
    if (vel_mode) {
        double lov = lo_(i, j, k, E) - c6 * (wqh + wql) * (et(qti, qtj, qtk, E) - et(qi, qj, qk, E));
        double hiv = hi_(i, j, k, E) - c6 * (wph + wpl) * (et(pti, ptj, ptk, E) - et(i, j, k, E));
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = fqb(gli, glj, glk, E);
        if (x == g.dhi[D] + 1) ghi = fqb(ghi_, ghj, ghk, E);
        edge_bcs(x, g.dlo[D], g.dhi[D], bc.lo[E][D], bc.hi[E][D], glo, ghi, true, lov, hiv);
        out(i, j, k) = upwind(lov, hiv, u);
        return;
    }
    CFab fq = fab_at(q, box, b);
    for (int n = 0; n < ncomp; ++n) {
        double eql = et(qi, qj, qk, n), eqh = et(qti, qtj, qtk, n);
        double epl = et(i, j, k, n), eph = et(pti, ptj, ptk, n);
        double lov, hiv;
        if ((cons_mask >> n) & 1) {
            double ql = fq(qi, qj, qk, n), qh = fq(i, j, k, n);
            lov = lo_(i, j, k, n) - c3 * (wqh * eqh - wql * eql) + c3 * ql * (wqh - wql);
            hiv = hi_(i, j, k, n) - c3 * (wph * eph - wpl * epl) + c3 * qh * (wph - wpl);
        } else {
            lov = lo_(i, j, k, n) - c6 * (wqh + wql) * (eqh - eql);
            hiv = hi_(i, j, k, n) - c6 * (wph + wpl) * (eph - epl);
        }
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = fqb(gli, glj, glk, n);
        if (x == g.dhi[D] + 1) ghi = fqb(ghi_, ghj, ghk, n);
        edge_bcs(x, g.dlo[D], g.dhi[D], bc.lo[n][D], bc.hi[n][D], glo, ghi, is_velocity && n == D, lov, hiv);
        out(i, j, k, n) = upwind(lov, hiv, u);
    }
}

// ---------------------------------------------------------------------------
// stage C: final states on the valid faces of direction D.
// crnA: corner states normal to T1 = (D+1)%3 corrected along T2;
// crnB: normal to T2 corrected along T1.
// vel_mode: normal velocity only -> u_mac (Burgers);  otherwise all
// components -> face states and fluxes upwinded by the MAC velocity.
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256) k_final(CMFDev q, CMFDev qbc, CMFDev f, CMFDev divu, int has_divu,
                                               int ncomp, CMFDev flo, CMFDev fhi, CMFDev crnA, CMFDev crnB,
                                               CMFDev vel1, CMFDev vel2, CMFDev velD, MFDev out_state,
                                               MFDev out_flux, RegDev r, const int* __restrict__ box, BCs bc,
                                               GeomDev g, double c2a, double c2b, double c4a, double c4b,
                                               double hdt, double area, int uft, int vel_mode, int is_velocity,
                                               int area_weighted, int allow_inflow, int cons_mask)
{
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    int qi = i, qj = j, qk = k;
    unit<D>(qi, qj, qk, -1);
    int q1i = qi, q1j = qj, q1k = qk, p1i = i, p1j = j, p1k = k;
    unit<T1>(q1i, q1j, q1k, 1);
    unit<T1>(p1i, p1j, p1k, 1);
    int q2i = qi, q2j = qj, q2k = qk, p2i = i, p2j = j, p2k = k;
    unit<T2>(q2i, q2j, q2k, 1);
    unit<T2>(p2i, p2j, p2k, 1);
    CFab v1 = fab_at(vel1, box, b);
    CFab v2 = fab_at(vel2, box, b);
    CFab A = fab_at(crnA, box, b);
    CFab B = fab_at(crnB, box, b);
    CFab lo_ = fab_at(flo, box, b);
    CFab hi_ = fab_at(fhi, box, b);
    CFab ff = fab_at(f, box, b);
    CFab fqb = fab_at(qbc, box, b);
    const int x = pick<D>(i, j, k);
    int gli = i, glj = j, glk = k, ghi_ = i, ghj = j, ghk = k;
    if (D == 0) { gli = g.dlo[0] - 1; ghi_ = g.dhi[0] + 1; }
    if (D == 1) { glj = g.dlo[1] - 1; ghj = g.dhi[1] + 1; }
    if (D == 2) { glk = g.dlo[2] - 1; ghk = g.dhi[2] + 1; }
    const double v1ql = v1(qi, qj, qk), v1qh = v1(q1i, q1j, q1k);
    const double v1pl = v1(i, j, k), v1ph = v1(p1i, p1j, p1k);
    const double v2ql = v2(qi, qj, qk), v2qh = v2(q2i, q2j, q2k);
    const double v2pl = v2(i, j, k), v2ph = v2(p2i, p2j, p2k);
    if (vel_mode) {
        double stl = lo_(i, j, k, D) - c4a * (v1qh + v1ql) * (A(q1i, q1j, q1k) - A(qi, qj, qk)) -
                     c4b * (v2qh + v2ql) * (B(q2i, q2j, q2k) - B(qi, qj, qk));
        double sth = hi_(i, j, k, D) - c4a * (v1ph + v1pl) * (A(p1i, p1j, p1k) - A(i, j, k)) -
                     c4b * (v2ph + v2pl) * (B(p2i, p2j, p2k) - B(i, j, k));
        if (!uft) {
            stl += hdt * ff(qi, qj, qk, D);
            sth += hdt * ff(i, j, k, D);
        }
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = fqb(gli, glj, glk, D);
        if (x == g.dhi[D] + 1) ghi = fqb(ghi_, ghj, ghk, D);
        edge_bcs(x, g.dlo[D], g.dhi[D], bc.lo[D][D], bc.hi[D][D], glo, ghi, true, stl, sth);
        if (!allow_inflow) {
            int tl = bc.lo[D][D], th = bc.hi[D][D];
            if (x == g.dlo[D] && (tl == KSGF_BC_FOEXTRAP || tl == KSGF_BC_HOEXTRAP)) {
                sth = fmin(sth, 0.0);
                stl = sth;
            }
            if (x == g.dhi[D] + 1 && (th == KSGF_BC_FOEXTRAP || th == KSGF_BC_HOEXTRAP)) {
                stl = fmax(stl, 0.0);
                sth = stl;
            }
        }
        Fab o = fab_at(out_state, box, b);
        o(i, j, k) = burgers(stl, sth);
        return;
    }
    CFab fq = fab_at(q, box, b);
    CFab vd = fab_at(velD, box, b);
    const double u = vd(i, j, k);
    Fab os = fab_at(out_state, box, b);
    Fab of = fab_at(out_flux, box, b);
    double dql = 0.0, dqh = 0.0;
    if (has_divu) {
        CFab dv = fab_at(divu, box, b);
        dql = dv(qi, qj, qk);
        dqh = dv(i, j, k);
    }
    for (int n = 0; n < ncomp; ++n) {
        double a_ql = A(qi, qj, qk, n), a_qh = A(q1i, q1j, q1k, n);
        double a_pl = A(i, j, k, n), a_ph = A(p1i, p1j, p1k, n);
        double b_ql = B(qi, qj, qk, n), b_qh = B(q2i, q2j, q2k, n);
        double b_pl = B(i, j, k, n), b_ph = B(p2i, p2j, p2k, n);
        double stl, sth;
        if ((cons_mask >> n) & 1) {
            double ql = fq(qi, qj, qk, n), qh = fq(i, j, k, n);
            stl = lo_(i, j, k, n) - c2a * (v1qh * a_qh - v1ql * a_ql) - c2b * (v2qh * b_qh - v2ql * b_ql) +
                  c2a * ql * (v1qh - v1ql) + c2b * ql * (v2qh - v2ql);
            sth = hi_(i, j, k, n) - c2a * (v1ph * a_ph - v1pl * a_pl) - c2b * (v2ph * b_ph - v2pl * b_pl) +
                  c2a * qh * (v1ph - v1pl) + c2b * qh * (v2ph - v2pl);
            if (has_divu) {
                stl -= hdt * ql * dql;
                sth -= hdt * qh * dqh;
            }
        } else {
            stl = lo_(i, j, k, n) - c4a * (v1qh + v1ql) * (a_qh - a_ql) - c4b * (v2qh + v2ql) * (b_qh - b_ql);
            sth = hi_(i, j, k, n) - c4a * (v1ph + v1pl) * (a_ph - a_pl) - c4b * (v2ph + v2pl) * (b_ph - b_pl);
        }
        if (!uft) {
            stl += hdt * ff(qi, qj, qk, n);
            sth += hdt * ff(i, j, k, n);
        }
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = fqb(gli, glj, glk, n);
        if (x == g.dhi[D] + 1) ghi = fqb(ghi_, ghj, ghk, n);
        edge_bcs(x, g.dlo[D], g.dhi[D], bc.lo[n][D], bc.hi[n][D], glo, ghi, is_velocity && n == D, stl, sth);
        if (is_velocity && n == D && !allow_inflow) {
            int tl = bc.lo[n][D], th = bc.hi[n][D];
            if (x == g.dlo[D] && (tl == KSGF_BC_FOEXTRAP || tl == KSGF_BC_HOEXTRAP)) {
                sth = fmin(sth, 0.0);
                stl = sth;
            }
            if (x == g.dhi[D] + 1 && (th == KSGF_BC_FOEXTRAP || th == KSGF_BC_HOEXTRAP)) {
                stl = fmax(stl, 0.0);
                sth = stl;
            }
        }
        double e = upwind(stl, sth, u);
        os(i, j, k, n) = e;
        of(i, j, k, n) = area_weighted ? (u * area) * e : u * e;
    }
}

// ---------------------------------------------------------------------------
// stages B and C fused: corner-coupled states and final states of one
// component in one kernel.  A workgroup owns a tile of (x, y) positions of
// one box -- an interior of output positions and a ring of one (sizes
// below) -- and marches up z.  At each plane every thread forms the six
// corner states at its position (the expressions of k_corner, reading the
// stage-A arrays) into LDS; interior threads then form the final states of
// the plane below (x and y faces) and of the current z face (the expressions
// of k_final) from their neighbours' corner states.  The corner arrays never
// reach HBM.
//   VEL: ExtrapVelToFaces (corner states of the third component, advected
//        by the Burgers velocities; finals: the normal velocity -> u_mac);
//   otherwise ComputeFluxesOnBoxFromState, one component per workgroup
//        (blockIdx.y): face states and fluxes.
// A position (i, j, k) owns the low x, y and z faces of cell (i, j, k); the
// box's top faces are owned by the positions one past its high end.
// ---------------------------------------------------------------------------
// A tile is kCfTx x kCfTy positions, one thread each (247 of the block's 256
// threads; the rest idle), of which the interior kCfInX x kCfInY write
// outputs.  A box of n cells has n + 1 face positions per direction; for the
// box edges 32 / 48 / 64 the 17 x 11 interior wastes fewer threads on partial
// tiles than a square 14 x 14 (0.61 against 0.52 output positions per
// thread-plane on the 256^3 ABL layout).
constexpr int kCfTx = 19, kCfTy = 13;
constexpr int kCfInX = kCfTx - 2, kCfInY = kCfTy - 2;
constexpr int kCfThreads = 256;
static_assert(kCfTx * kCfTy <= kCfThreads, "tile larger than the block");

struct CfTile {
    const int* tiles;  // 3 ints per tile: box, x0, y0 (first interior position)
};

struct CfArgs {
    CMFDev q, qbc, f, divu;
    int has_divu, ncomp;
    CMF3 lo, hi;       // stage A: face states (faces grown by 1)
    CMF3 vel;          // advecting velocities: MAC (fluxes) or Burgers (VEL), ngrow 1
    MF3 state, flux;   // outputs (fluxes); state[d] = u_mac[d] (VEL)
    double c3[3], c6[3], c2[3], c4[3], area[3];
    double hdt;
    int uft, is_velocity, area_weighted, allow_inflow, cons_mask;
};

// the corner arithmetic of k_corner: low / high side states of the D face
// corrected along T (eq*/ep*: upwinded T states on the T faces of the low
// and high cells, w*: their advecting velocities, ql/qh: the cells' values)
template <bool VEL>
__device__ __forceinline__ void cf_corner_states(double lo, double hi, double wql, double wqh, double wpl, double wph,
                                                 double eql, double eqh, double epl, double eph, double ql, double qh,
                                                 bool cons, double c3, double c6, double& lov, double& hiv)
{
    if (!VEL && cons) {
        lov = lo - c3 * (wqh * eqh - wql * eql) + c3 * ql * (wqh - wql);
        hiv = hi - c3 * (wph * eph - wpl * epl) + c3 * qh * (wph - wpl);
    } else {
        lov = lo - c6 * (wqh + wql) * (eqh - eql);
        hiv = hi - c6 * (wph + wpl) * (eph - epl);
    }
}

// BC codes of one component: lo/hi per direction
struct Bc1 {
    int lo[3], hi[3];
};
__device__ __forceinline__ Bc1 bc_of(const BCs& bc, int c)
{
    // a select per entry (constant indices): a runtime index into the
    // by-value argument would copy it to private memory
    Bc1 r;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = bc.lo[0][d];
        r.hi[d] = bc.hi[0][d];
    }
#pragma unroll
    for (int m = 1; m < kMaxComp; ++m) {
        const bool sel = (m == c);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            r.lo[d] = sel ? bc.lo[m][d] : r.lo[d];
            r.hi[d] = sel ? bc.hi[m][d] : r.hi[d];
        }
    }
    return r;
}

// the value of component c of qbc in the ghost cell beyond the domain
// boundary face of direction D at (i, j, k) (low side when lowside)
template <int D>
__device__ __noinline__ double cf_ghost(const double* qp, const long long* pre, int ncomp, const int* box, int b,
                                        int i, int j, int k, int c, int xg)
{
    CMFDev qbc;
    qbc.p = qp;
    qbc.pre = pre;
    qbc.itype = KSGF_CELL;
    qbc.ngrow = 3;
    qbc.ncomp = ncomp;
    int gi = i, gj = j, gk = k;
    const int x = xg;
    if (D == 0) gi = x;
    if (D == 1) gj = x;
    if (D == 2) gk = x;
    return fab_at(qbc, box, b)(gi, gj, gk, c);
}

// boundary treatment + upwinding of a corner state on the D face at (i, j, k)
template <int D, bool VEL>
__device__ __forceinline__ double cf_corner_finish(const CfArgs& a, const int* box, int b, int i, int j, int k, int c,
                                                   const Bc1& bcc, bool both, double lov, double hiv, double u,
                                                   const GeomDev& g)
{
    const int x = pick<D>(i, j, k);
    if (x == g.dlo[D] || x == g.dhi[D] + 1) {
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = cf_ghost<D>(a.qbc.p, a.qbc.pre, a.qbc.ncomp, box, b, i, j, k, c, g.dlo[D] - 1);
        if (x == g.dhi[D] + 1) ghi = cf_ghost<D>(a.qbc.p, a.qbc.pre, a.qbc.ncomp, box, b, i, j, k, c, g.dhi[D] + 1);
        edge_bcs(x, g.dlo[D], g.dhi[D], bcc.lo[D], bcc.hi[D], glo, ghi, both, lov, hiv);
    }
    return upwind(lov, hiv, u);
}

// final state of the D face at (i, j, k) (the arithmetic of k_final<D>):
// lo/hi: stage-A states of the face; A_*: corner states normal to
// T1 = (D+1)%3 corrected along T2 at q = p - e_D, q + e_T1, p, p + e_T1;
// B_*: normal to T2 corrected along T1 at q, q + e_T2, p, p + e_T2; v1*, v2*:
// the T1 / T2 advecting velocities at the same faces; fl/fh: the forcing of
// the low / high cell.  Returns the upwinded state (Burgers for VEL).
template <int D, bool VEL>
__device__ __forceinline__ double cf_final(const CfArgs& a, const int* box, int b, int i, int j, int k, int c,
                                           const Bc1& bcc, bool normal, bool cons, double lo, double hi, double a_ql,
                                           double a_qh, double a_pl, double a_ph, double b_ql, double b_qh,
                                           double b_pl, double b_ph, double v1ql, double v1qh, double v1pl,
                                           double v1ph, double v2ql, double v2qh, double v2pl, double v2ph, double ql,
                                           double qh, double dql, double dqh, double fl, double fh, double u,
                                           const GeomDev& g)
{
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    const int x = pick<D>(i, j, k);
    double stl, sth;
    if (!VEL && cons) {
        stl = lo - a.c2[T1] * (v1qh * a_qh - v1ql * a_ql) - a.c2[T2] * (v2qh * b_qh - v2ql * b_ql) +
              a.c2[T1] * ql * (v1qh - v1ql) + a.c2[T2] * ql * (v2qh - v2ql);
        sth = hi - a.c2[T1] * (v1ph * a_ph - v1pl * a_pl) - a.c2[T2] * (v2ph * b_ph - v2pl * b_pl) +
              a.c2[T1] * qh * (v1ph - v1pl) + a.c2[T2] * qh * (v2ph - v2pl);
        if (a.has_divu) {
            stl -= a.hdt * ql * dql;
            sth -= a.hdt * qh * dqh;
        }
    } else {
        stl = lo - a.c4[T1] * (v1qh + v1ql) * (a_qh - a_ql) - a.c4[T2] * (v2qh + v2ql) * (b_qh - b_ql);
        sth = hi - a.c4[T1] * (v1ph + v1pl) * (a_ph - a_pl) - a.c4[T2] * (v2ph + v2pl) * (b_ph - b_pl);
    }
    if (!a.uft) {
        stl += a.hdt * fl;
        sth += a.hdt * fh;
    }
    if (x == g.dlo[D] || x == g.dhi[D] + 1) {
        double glo = 0.0, ghi = 0.0;
        if (x == g.dlo[D]) glo = cf_ghost<D>(a.qbc.p, a.qbc.pre, a.qbc.ncomp, box, b, i, j, k, c, g.dlo[D] - 1);
        if (x == g.dhi[D] + 1) ghi = cf_ghost<D>(a.qbc.p, a.qbc.pre, a.qbc.ncomp, box, b, i, j, k, c, g.dhi[D] + 1);
        edge_bcs(x, g.dlo[D], g.dhi[D], bcc.lo[D], bcc.hi[D], glo, ghi, normal, stl, sth);
        if (normal && !a.allow_inflow) {
            const int tl = bcc.lo[D], th = bcc.hi[D];
            if (x == g.dlo[D] && (tl == KSGF_BC_FOEXTRAP || tl == KSGF_BC_HOEXTRAP)) {
                sth = fmin(sth, 0.0);
                stl = sth;
            }
            if (x == g.dhi[D] + 1 && (th == KSGF_BC_FOEXTRAP || th == KSGF_BC_HOEXTRAP)) {
                stl = fmax(stl, 0.0);
                sth = stl;
            }
        }
    }
    return VEL ? burgers(stl, sth) : upwind(stl, sth, u);
}

template <bool VEL>
__global__ void __launch_bounds__(256) k_corner_final(CfArgs a, CfTile tl, const int* __restrict__ box, BCs bc,
                                                      GeomDev g)
{
    constexpr int NS = VEL ? 2 : 1;  // components of each upwinded state the corners read
    constexpr int NC = VEL ? 3 : 1;  // components the workgroup forms states of
    constexpr int NP = kCfThreads;  // LDS slots: one per thread
    // inputs of plane k & 1: upwinded T states, advecting velocities, cell values
    __shared__ double s_ed[3][NS][2][NP];
    __shared__ double s_vel[3][2][NP];
    __shared__ double s_q[VEL ? 1 : 2][NP];
    // corner states: xy, yx, zx, zy at planes k & 1; xz, yz at plane k - 1
    __shared__ double s_c[4][2][NP];
    __shared__ double s_cz[2][NP];
    // forcing of planes k & 1 (components x, y, z for VEL); the finals read
    // their cells and the low neighbours from here
    __shared__ double s_f[NC][2][NP];
    // consecutive tiles (x fastest within a box row) on one XCD, so the rows
    // tiles share (halos, cache lines across tile edges) meet in its L2
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int b = tl.tiles[3 * t];
    const int x0 = tl.tiles[3 * t + 1], y0 = tl.tiles[3 * t + 2];
    const int tx = threadIdx.x % kCfTx, ty = threadIdx.x / kCfTx;  // ty == kCfTy: idle thread
    const int i = x0 - 1 + tx, j = y0 - 1 + ty;
    const int n = VEL ? 0 : (int)blockIdx.y;
    const int* bx = box + 6 * b;
    const int l0 = bx[0], l1 = bx[1], l2 = bx[2], h0 = bx[3], h1 = bx[4], h2 = bx[5];
    const int P = threadIdx.x;
    const int PXm = P - 1, PXp = P + 1, PYm = P - kCfTx, PYp = P + kCfTx;
    // storage of the face arrays with one ghost layer (stage A, advecting
    // velocities; u_mac for VEL): per direction, plane stride and the
    // thread's offset at plane l2 - 1; component stride = whole FAB
    const int e0 = h0 - l0 + 3, e1 = h1 - l1 + 3, e2 = h2 - l2 + 3;
    int ps1[3], cs1[3], o1[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int m0 = e0 + (d == 0), m1 = e1 + (d == 1), m2 = e2 + (d == 2);
        ps1[d] = m0 * m1;
        cs1[d] = ps1[d] * m2;
        o1[d] = (i - l0 + 1) + m0 * (j - l1 + 1);  // + ps1 * (k - l2 + 1)
    }
    // stage-A face states and advecting velocities
    const double *lo_p[3], *hi_p[3], *vel_p[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo_p[d] = a.lo.m[d].p + (long long)a.lo.m[d].ncomp * a.lo.m[d].pre[b];
        hi_p[d] = a.hi.m[d].p + (long long)a.hi.m[d].ncomp * a.hi.m[d].pre[b];
        vel_p[d] = a.vel.m[d].p + a.vel.m[d].pre[b];
    }
    // cell arrays: q (ngrow 3), f and divu (ngrow 1)
    const int q0 = h0 - l0 + 7, q1 = h1 - l1 + 7;
    const int psq = q0 * q1, csq = psq * (h2 - l2 + 7);
    const int oq = (i - l0 + 3) + q0 * (j - l1 + 3);  // + psq * (k - l2 + 3)
    const double* q_p = a.q.p + (long long)a.q.ncomp * a.q.pre[b];
    const int f0 = h0 - l0 + 3, f1 = h1 - l1 + 3;
    const int psf = f0 * f1, csf = psf * (h2 - l2 + 3);
    const int of = (i - l0 + 1) + f0 * (j - l1 + 1);  // + psf * (k - l2 + 1)
    const double* f_p = a.f.p + (long long)a.f.ncomp * a.f.pre[b];
    const double* dv_p = a.has_divu ? a.divu.p + a.divu.pre[b] : nullptr;
    // outputs: faces with no ghost layer (fluxes) or u_mac (ngrow 1, VEL)
    int ps0[3], cs0[3], o0[3];
    double *st_p[3], *fl_p[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (VEL) {
            ps0[d] = ps1[d];
            cs0[d] = cs1[d];
            o0[d] = o1[d];  // u_mac: the same storage as the stage arrays
        } else {
            const int m0 = e0 - 2 + (d == 0), m1 = e1 - 2 + (d == 1), m2 = e2 - 2 + (d == 2);
            ps0[d] = m0 * m1;
            cs0[d] = ps0[d] * m2;
            o0[d] = (i - l0) + m0 * (j - l1);  // + ps0 * (k - l2)
        }
        st_p[d] = a.state.m[d].p + (long long)a.state.m[d].ncomp * a.state.m[d].pre[b];
        fl_p[d] = VEL ? nullptr : a.flux.m[d].p + (long long)a.flux.m[d].ncomp * a.flux.m[d].pre[b];
    }
    // box ranges of the corner states (faces of D, E = 3-D-T grown by one)
    const bool ri_x = (i >= l0 && i <= h0 + 1), ri_c = (i >= l0 && i <= h0), ri_g = (i >= l0 - 1 && i <= h0 + 1);
    const bool rj_y = (j >= l1 && j <= h1 + 1), rj_c = (j >= l1 && j <= h1), rj_g = (j >= l1 - 1 && j <= h1 + 1);
    // tile positions whose corner states an interior final reads (their
    // inputs then lie inside the tile)
    const bool live = ty < kCfTy;
    const bool tx_in = tx >= 1 && tx <= kCfInX, ty_in = ty >= 1 && ty <= kCfInY;
    const bool m_xy = tx >= 1 && ty_in && ri_x && rj_c;
    const bool m_xz = tx >= 1 && ty <= kCfInY && ri_x && rj_g;
    const bool m_yx = tx_in && ty >= 1 && ri_c && rj_y;
    const bool m_yz = live && tx <= kCfInX && ty >= 1 && ri_g && rj_y;
    const bool m_zx = tx_in && ty <= kCfInY && ri_c && rj_g;
    const bool m_zy = tx <= kCfInX && ty_in && ri_g && rj_c;
    const bool inner = tx_in && ty_in;
    // stage-loadable positions: face storage per direction (x, y extents)
    bool okf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        okf[d] = i >= l0 - 1 && i <= h0 + 1 + (d == 0) && j >= l1 - 1 && j <= h1 + 1 + (d == 1);
    const bool q_ok = i >= l0 - 3 && i <= h0 + 3 && j >= l1 - 3 && j <= h1 + 3;
    const bool f_ok = i >= l0 - 1 && i <= h0 + 1 && j >= l1 - 1 && j <= h1 + 1;
    const bool cons = !VEL && ((a.cons_mask >> n) & 1);
    const Bc1 bcn = bc_of(bc, n);
    Bc1 bcv[3];
    if (VEL)
#pragma unroll
        for (int d = 0; d < 3; ++d) bcv[d] = bc_of(bc, d);
    // own-position inputs of plane kk, prefetched one plane ahead: the
    // stage-A states of the x, y and z faces of every component the
    // workgroup forms (the upwinded states are formed from them here),
    // advecting velocities, forcing and (momentum) the cell value
    double p_lx[NC], p_hx[NC], p_ly[NC], p_hy[NC], p_lz[NC], p_hz[NC], p_vel[3], p_f[NC], p_q = 0.0;
    auto load_plane = [&](int kk) {
        const int kz = kk - l2 + 1;
        const bool okx = okf[0] && kk <= h2 + 1, oky = okf[1] && kk <= h2 + 1;
        const bool okz = okf[2] && kk >= l2 && kk <= h2 + 1;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int cc = VEL ? c : n;
            const int ox = o1[0] + ps1[0] * kz + cc * cs1[0], oy = o1[1] + ps1[1] * kz + cc * cs1[1];
            const int oz = o1[2] + ps1[2] * kz + cc * cs1[2];
            p_lx[c] = okx ? lo_p[0][ox] : 0.0;
            p_hx[c] = okx ? hi_p[0][ox] : 0.0;
            p_ly[c] = oky ? lo_p[1][oy] : 0.0;
            p_hy[c] = oky ? hi_p[1][oy] : 0.0;
            p_lz[c] = okz ? lo_p[2][oz] : 0.0;
            p_hz[c] = okz ? hi_p[2][oz] : 0.0;
            p_f[c] = (!a.uft && f_ok) ? f_p[of + psf * kz + cc * csf] : 0.0;
        }
        if (!VEL) p_q = q_ok ? q_p[oq + psq * (kk - l2 + 3) + n * csq] : 0.0;
#pragma unroll
        for (int T = 0; T < 3; ++T) {
            const bool ok = okf[T] && (kk <= h2 + 1 + (T == 2));
            p_vel[T] = ok ? vel_p[T][o1[T] + ps1[T] * kz] : 0.0;
        }
    };
    // x and y face states of plane k - 1: the finals and the corners along z
    // read them one plane after they are loaded
    double cxl[NC], cxh[NC], cyl[NC], cyh[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) cxl[c] = cxh[c] = cyl[c] = cyh[c] = 0.0;
    load_plane(l2 - 1);
    for (int k = l2 - 1; k <= h2 + 1; ++k) {
        const int s = k & 1, s1 = s ^ 1;
        const int kz1 = k - l2 + 1;  // plane index in storage with one ghost layer
        const bool rk_z = (k >= l2 && k <= h2 + 1), rk_c1 = (k - 1 >= l2 && k - 1 <= h2);
        // the own-position inputs of plane k
        double lx[NC], hx[NC], ly[NC], hy[NC], lz[NC], hz[NC], fz[NC], vel[3];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            lx[c] = p_lx[c];
            hx[c] = p_hx[c];
            ly[c] = p_ly[c];
            hy[c] = p_hy[c];
            lz[c] = p_lz[c];
            hz[c] = p_hz[c];
            fz[c] = p_f[c];
        }
        const double qk = p_q;
#pragma unroll
        for (int T = 0; T < 3; ++T) vel[T] = p_vel[T];
        // prefetch plane k + 1
        if (k + 1 <= h2 + 1) load_plane(k + 1);
        // stage plane k: upwinded states (VEL: the transverse components,
        // x faces z, y; y faces x, z; z faces y, x), velocities, cells, forcing
        if (VEL) {
            s_ed[0][0][s][P] = upwind(lx[NC - 1], hx[NC - 1], vel[0]);
            s_ed[0][NS - 1][s][P] = upwind(lx[NC > 1 ? 1 : 0], hx[NC > 1 ? 1 : 0], vel[0]);
            s_ed[1][0][s][P] = upwind(ly[0], hy[0], vel[1]);
            s_ed[1][NS - 1][s][P] = upwind(ly[NC - 1], hy[NC - 1], vel[1]);
            s_ed[2][0][s][P] = upwind(lz[NC > 1 ? 1 : 0], hz[NC > 1 ? 1 : 0], vel[2]);
            s_ed[2][NS - 1][s][P] = upwind(lz[0], hz[0], vel[2]);
        } else {
            s_ed[0][0][s][P] = upwind(lx[0], hx[0], vel[0]);
            s_ed[1][0][s][P] = upwind(ly[0], hy[0], vel[1]);
            s_ed[2][0][s][P] = upwind(lz[0], hz[0], vel[2]);
            s_q[s][P] = qk;
        }
#pragma unroll
        for (int T = 0; T < 3; ++T) s_vel[T][s][P] = vel[T];
#pragma unroll
        for (int c = 0; c < NC; ++c) s_f[c][s][P] = fz[c];
        const bool d_xy = m_xy, d_yx = m_yx, d_zx = m_zx && rk_z, d_zy = m_zy && rk_z, d_xz = m_xz && rk_c1,
                   d_yz = m_yz && rk_c1;
        const bool f_x = inner && rk_c1 && ri_x && rj_c, f_y = inner && rk_c1 && ri_c && rj_y,
                   f_z = inner && rk_z && ri_c && rj_c;
        // the face states the corners and finals read: own faces of plane k
        // (z face k), the x and y faces of plane k - 1 carried over
        constexpr int cxy = VEL ? 2 : 0, cyx = VEL ? 2 : 0, czx = VEL ? 1 : 0, czy = 0, cxz = VEL ? 1 : 0, cyz = 0,
                      cfx = 0, cfy = VEL ? 1 : 0, cfz = VEL ? 2 : 0;
        const double lxy = lx[cxy], hxy = hx[cxy], lyx = ly[cyx], hyx = hy[cyx], lzx = lz[czx], hzx = hz[czx];
        const double lzy = lz[czy], hzy = hz[czy], lfz = lz[cfz], hfz = hz[cfz];
        const double lxz = cxl[cxz], hxz = cxh[cxz], lfx = cxl[cfx], hfx = cxh[cfx];
        const double lyz = cyl[cyz], hyz = cyh[cyz], lfy = cyl[cfy], hfy = cyh[cfy];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            cxl[c] = lx[c];
            cxh[c] = hx[c];
            cyl[c] = ly[c];
            cyh[c] = hy[c];
        }
        const int kf = k - l2 + 1;  // plane k in the ngrow-1 cell storage of f / divu
        __syncthreads();
        double lov, hiv;
        // xy (x face, corrected along y, plane k); VEL: component z (slot 1)
        if (d_xy) {
            const int sl = VEL ? 1 : 0;
            cf_corner_states<VEL>(lxy, hxy, s_vel[1][s][PXm], s_vel[1][s][PXm + kCfTx], s_vel[1][s][P],
                                  s_vel[1][s][PYp], s_ed[1][sl][s][PXm], s_ed[1][sl][s][PXm + kCfTx],
                                  s_ed[1][sl][s][P], s_ed[1][sl][s][PYp], VEL ? 0.0 : s_q[s][PXm],
                                  VEL ? 0.0 : s_q[s][P], cons, a.c3[1], a.c6[1], lov, hiv);
            s_c[0][s][P] = cf_corner_finish<0, VEL>(a, box, b, i, j, k, VEL ? cxy : n, VEL ? bcv[cxy] : bcn,
                                                    VEL ? true : (a.is_velocity && n == 0), lov, hiv,
                                                    s_vel[0][s][P], g);
        }
        // yx (y face, corrected along x, plane k); VEL: component z (slot 0)
        if (d_yx) {
            cf_corner_states<VEL>(lyx, hyx, s_vel[0][s][PYm], s_vel[0][s][PYm + 1], s_vel[0][s][P], s_vel[0][s][PXp],
                                  s_ed[0][0][s][PYm], s_ed[0][0][s][PYm + 1], s_ed[0][0][s][P], s_ed[0][0][s][PXp],
                                  VEL ? 0.0 : s_q[s][PYm], VEL ? 0.0 : s_q[s][P], cons, a.c3[0], a.c6[0], lov, hiv);
            s_c[1][s][P] = cf_corner_finish<1, VEL>(a, box, b, i, j, k, VEL ? cyx : n, VEL ? bcv[cyx] : bcn,
                                                    VEL ? true : (a.is_velocity && n == 1), lov, hiv,
                                                    s_vel[1][s][P], g);
        }
        // zx (z face k, corrected along x); VEL: component y (slot 1)
        if (d_zx) {
            const int sl = VEL ? 1 : 0;
            cf_corner_states<VEL>(lzx, hzx, s_vel[0][s1][P], s_vel[0][s1][PXp], s_vel[0][s][P], s_vel[0][s][PXp],
                                  s_ed[0][sl][s1][P], s_ed[0][sl][s1][PXp], s_ed[0][sl][s][P], s_ed[0][sl][s][PXp],
                                  VEL ? 0.0 : s_q[s1][P], VEL ? 0.0 : s_q[s][P], cons, a.c3[0], a.c6[0], lov, hiv);
            s_c[2][s][P] = cf_corner_finish<2, VEL>(a, box, b, i, j, k, VEL ? czx : n, VEL ? bcv[czx] : bcn,
                                                    VEL ? true : (a.is_velocity && n == 2), lov, hiv,
                                                    s_vel[2][s][P], g);
        }
        // zy (z face k, corrected along y); VEL: component x (slot 0)
        if (d_zy) {
            cf_corner_states<VEL>(lzy, hzy, s_vel[1][s1][P], s_vel[1][s1][PYp], s_vel[1][s][P], s_vel[1][s][PYp],
                                  s_ed[1][0][s1][P], s_ed[1][0][s1][PYp], s_ed[1][0][s][P], s_ed[1][0][s][PYp],
                                  VEL ? 0.0 : s_q[s1][P], VEL ? 0.0 : s_q[s][P], cons, a.c3[1], a.c6[1], lov, hiv);
            s_c[3][s][P] = cf_corner_finish<2, VEL>(a, box, b, i, j, k, VEL ? czy : n, VEL ? bcv[czy] : bcn,
                                                    VEL ? true : (a.is_velocity && n == 2), lov, hiv,
                                                    s_vel[2][s][P], g);
        }
        // xz (x face, corrected along z, plane k - 1); VEL: component y (slot 0)
        if (d_xz) {
            cf_corner_states<VEL>(lxz, hxz, s_vel[2][s1][PXm], s_vel[2][s][PXm], s_vel[2][s1][P], s_vel[2][s][P],
                                  s_ed[2][0][s1][PXm], s_ed[2][0][s][PXm], s_ed[2][0][s1][P], s_ed[2][0][s][P],
                                  VEL ? 0.0 : s_q[s1][PXm], VEL ? 0.0 : s_q[s1][P], cons, a.c3[2], a.c6[2], lov, hiv);
            s_cz[0][P] = cf_corner_finish<0, VEL>(a, box, b, i, j, k - 1, VEL ? cxz : n, VEL ? bcv[cxz] : bcn,
                                                  VEL ? true : (a.is_velocity && n == 0), lov, hiv, s_vel[0][s1][P],
                                                  g);
        }
        // yz (y face, corrected along z, plane k - 1); VEL: component x (slot 1)
        if (d_yz) {
            const int sl = VEL ? 1 : 0;
            cf_corner_states<VEL>(lyz, hyz, s_vel[2][s1][PYm], s_vel[2][s][PYm], s_vel[2][s1][P], s_vel[2][s][P],
                                  s_ed[2][sl][s1][PYm], s_ed[2][sl][s][PYm], s_ed[2][sl][s1][P], s_ed[2][sl][s][P],
                                  VEL ? 0.0 : s_q[s1][PYm], VEL ? 0.0 : s_q[s1][P], cons, a.c3[2], a.c6[2], lov,
                                  hiv);
            s_cz[1][P] = cf_corner_finish<1, VEL>(a, box, b, i, j, k - 1, VEL ? cyz : n, VEL ? bcv[cyz] : bcn,
                                                  VEL ? true : (a.is_velocity && n == 1), lov, hiv, s_vel[1][s1][P],
                                                  g);
        }
        __syncthreads();
        // forcing of the finals' cells (0 with forces in the transverse terms)
        constexpr int ffx = VEL ? 0 : 0, ffy = VEL ? 1 : 0, ffz = VEL ? 2 : 0;
        const double fxl = f_x ? s_f[ffx][s1][PXm] : 0.0, fxh = f_x ? s_f[ffx][s1][P] : 0.0;
        const double fyl = f_y ? s_f[ffy][s1][PYm] : 0.0, fyh = f_y ? s_f[ffy][s1][P] : 0.0;
        const double fzl = f_z ? s_f[ffz][s1][P] : 0.0, fzh = f_z ? s_f[ffz][s][P] : 0.0;
        // x face of cell (i, j, k-1): A = yz (plane k-1), B = zy (planes k-1, k)
        if (f_x) {
            const double dql = (!VEL && a.has_divu) ? dv_p[of + psf * (kf - 1) - 1] : 0.0;
            const double dqh = (!VEL && a.has_divu) ? dv_p[of + psf * (kf - 1)] : 0.0;
            const double e = cf_final<0, VEL>(
                a, box, b, i, j, k - 1, VEL ? cfx : n, VEL ? bcv[cfx] : bcn, VEL ? true : (a.is_velocity && n == 0),
                cons, lfx, hfx, s_cz[1][PXm], s_cz[1][PXm + kCfTx], s_cz[1][P], s_cz[1][PYp], s_c[3][s1][PXm],
                s_c[3][s][PXm], s_c[3][s1][P], s_c[3][s][P], s_vel[1][s1][PXm], s_vel[1][s1][PXm + kCfTx],
                s_vel[1][s1][P], s_vel[1][s1][PYp], s_vel[2][s1][PXm], s_vel[2][s][PXm], s_vel[2][s1][P],
                s_vel[2][s][P], VEL ? 0.0 : s_q[s1][PXm], VEL ? 0.0 : s_q[s1][P], dql, dqh, fxl, fxh,
                s_vel[0][s1][P], g);
            const double u = s_vel[0][s1][P];
            const int oo = o0[0] + ps0[0] * (VEL ? kz1 - 1 : k - 1 - l2);
            if (VEL) {
                st_p[0][oo] = e;
            } else {
                st_p[0][oo + n * cs0[0]] = e;
                fl_p[0][oo + n * cs0[0]] = a.area_weighted ? (u * a.area[0]) * e : u * e;
            }
        }
        // y face of cell (i, j, k-1): A = zx (planes k-1, k), B = xz (plane k-1)
        if (f_y) {
            const double dql = (!VEL && a.has_divu) ? dv_p[of + psf * (kf - 1) - f0] : 0.0;
            const double dqh = (!VEL && a.has_divu) ? dv_p[of + psf * (kf - 1)] : 0.0;
            const double e = cf_final<1, VEL>(
                a, box, b, i, j, k - 1, VEL ? cfy : n, VEL ? bcv[cfy] : bcn, VEL ? true : (a.is_velocity && n == 1),
                cons, lfy, hfy, s_c[2][s1][PYm], s_c[2][s][PYm], s_c[2][s1][P], s_c[2][s][P], s_cz[0][PYm],
                s_cz[0][PYm + 1], s_cz[0][P], s_cz[0][PXp], s_vel[2][s1][PYm], s_vel[2][s][PYm], s_vel[2][s1][P],
                s_vel[2][s][P], s_vel[0][s1][PYm], s_vel[0][s1][PYm + 1], s_vel[0][s1][P], s_vel[0][s1][PXp],
                VEL ? 0.0 : s_q[s1][PYm], VEL ? 0.0 : s_q[s1][P], dql, dqh, fyl, fyh, s_vel[1][s1][P], g);
            const double u = s_vel[1][s1][P];
            const int oo = o0[1] + ps0[1] * (VEL ? kz1 - 1 : k - 1 - l2);
            if (VEL) {
                st_p[1][oo] = e;
            } else {
                st_p[1][oo + n * cs0[1]] = e;
                fl_p[1][oo + n * cs0[1]] = a.area_weighted ? (u * a.area[1]) * e : u * e;
            }
        }
        // z face k: A = xy (planes k-1, k), B = yx (planes k-1, k)
        if (f_z) {
            const double dql = (!VEL && a.has_divu) ? dv_p[of + psf * (kf - 1)] : 0.0;
            const double dqh = (!VEL && a.has_divu) ? dv_p[of + psf * kf] : 0.0;
            const double e = cf_final<2, VEL>(
                a, box, b, i, j, k, VEL ? cfz : n, VEL ? bcv[cfz] : bcn, VEL ? true : (a.is_velocity && n == 2),
                cons, lfz, hfz, s_c[0][s1][P], s_c[0][s1][PXp], s_c[0][s][P], s_c[0][s][PXp], s_c[1][s1][P],
                s_c[1][s1][PYp], s_c[1][s][P], s_c[1][s][PYp], s_vel[0][s1][P], s_vel[0][s1][PXp], s_vel[0][s][P],
                s_vel[0][s][PXp], s_vel[1][s1][P], s_vel[1][s1][PYp], s_vel[1][s][P], s_vel[1][s][PYp],
                VEL ? 0.0 : s_q[s1][P], VEL ? 0.0 : s_q[s][P], dql, dqh, fzl, fzh, s_vel[2][s][P], g);
            const double u = s_vel[2][s][P];
            const int oo = o0[2] + ps0[2] * (VEL ? kz1 : k - l2);
            if (VEL) {
                st_p[2][oo] = e;
            } else {
                st_p[2][oo + n * cs0[2]] = e;
                fl_p[2][oo + n * cs0[2]] = a.area_weighted ? (u * a.area[2]) * e : u * e;
            }
        }
        __syncthreads();
    }
}

// known edge state: flux = u * face
template <int D>
__global__ void __launch_bounds__(256) k_known_flux(CMFDev face, CMFDev velD, MFDev flux, int ncomp, RegDev r,
                                                    const int* __restrict__ box, double area, int area_weighted)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab fc = fab_at(face, box, b);
    CFab vd = fab_at(velD, box, b);
    Fab fl = fab_at(flux, box, b);
    double u = vd(i, j, k);
    for (int n = 0; n < ncomp; ++n) fl(i, j, k, n) = area_weighted ? (u * area) * fc(i, j, k, n) : u * fc(i, j, k, n);
}

__global__ void __launch_bounds__(256) k_divergence(MFDev div, CMF3 fl, int ncomp, RegDev r,
                                                    const int* __restrict__ box, double mult, double dxi,
                                                    double dyi, double dzi, double vol, int area_weighted)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    Fab o = fab_at(div, box, b);
    CFab fx = fab_at(fl.m[0], box, b);
    CFab fy = fab_at(fl.m[1], box, b);
    CFab fz = fab_at(fl.m[2], box, b);
    for (int n = 0; n < ncomp; ++n) {
        double r_;
        if (area_weighted)
            r_ = mult * ((fx(i + 1, j, k, n) - fx(i, j, k, n)) + (fy(i, j + 1, k, n) - fy(i, j, k, n)) +
                         (fz(i, j, k + 1, n) - fz(i, j, k, n))) /
                 vol;
        else
            r_ = mult * ((fx(i + 1, j, k, n) - fx(i, j, k, n)) * dxi + (fy(i, j + 1, k, n) - fy(i, j, k, n)) * dyi +
                         (fz(i, j, k + 1, n) - fz(i, j, k, n)) * dzi);
        o(i, j, k, n) = r_;
    }
}

// PredictStateOn{X,Y,Z}Face for one component on valid cells
template <int D>
__global__ void __launch_bounds__(256) k_ppm_predict(CMFDev s, int comp, CMFDev vel, MFDev Im, MFDev Ip,
                                                     RegDev r, const int* __restrict__ box, int bclo, int bchi,
                                                     int dlo, int dhi, double dtdx, int lim)
{
    int b, i, j, k;
    if (!region_locate(r, box, b, i, j, k)) return;
    CFab sf = fab_at(s, box, b);
    CFab vf = fab_at(vel, box, b);
    const long long st = (D == 0) ? 1 : (D == 1 ? (long long)sf.n0 : (long long)sf.n0 * sf.n1);
    const double* pn = sf.p + sf.idx(i, j, k) + comp * sf.cs;
    int i1 = i, j1 = j, k1 = k;
    unit<D>(i1, j1, k1, 1);
    double a, bb;
    ppm_trace(pn[-2 * st], pn[-st], pn[0], pn[st], pn[2 * st], lim, bclo, bchi, pick<D>(i, j, k), dlo, dhi,
              vf(i, j, k), vf(i1, j1, k1), dtdx, a, bb);
    Fab fm = fab_at(Im, box, b);
    Fab fp = fab_at(Ip, box, b);
    fm(i, j, k) = a;
    fp(i, j, k) = bb;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int check_lim(int lim)
{
    if (lim == KSGF_PPM_WENOZ || lim == KSGF_PPM_WENO_JS) {
        set_error("WENO limiters are not provided by this build");
        return KSGF_ERR_UNSUPPORTED;
    }
    if (lim != KSGF_PPM_NOLIMITER && lim != KSGF_PPM_VANLEER && lim != KSGF_PPM_UPWIND && lim != KSGF_PPM_MINMOD) {
        set_error("unknown limiter_type");
        return KSGF_ERR_INVALID;
    }
    return KSGF_OK;
}

static int check_bc(const ksgf_bcrec* bc, int ncomp, const ksgf_boxarray* ba)
{
    for (int n = 0; n < ncomp; ++n)
        for (int d = 0; d < 3; ++d) {
            int t[2] = {bc[n].lo[d], bc[n].hi[d]};
            for (int s = 0; s < 2; ++s) {
                if (t[s] != KSGF_BC_INT_DIR && t[s] != KSGF_BC_FOEXTRAP && t[s] != KSGF_BC_EXT_DIR &&
                    t[s] != KSGF_BC_HOEXTRAP) {
                    set_error("unsupported BCType (only int_dir, foextrap, ext_dir, hoextrap)");
                    return KSGF_ERR_UNSUPPORTED;
                }
                if (ba->periodic[d] && t[s] != KSGF_BC_INT_DIR) {
                    set_error("periodic direction requires int_dir BCs");
                    return KSGF_ERR_INVALID;
                }
            }
        }
    return KSGF_OK;
}

static BCs make_bcs(const ksgf_bcrec* bc, int ncomp)
{
    BCs b;
    for (int n = 0; n < kMaxComp; ++n)
        for (int d = 0; d < 3; ++d) {
            b.lo[n][d] = (n < ncomp) ? bc[n].lo[d] : 0;
            b.hi[n][d] = (n < ncomp) ? bc[n].hi[d] : 0;
        }
    return b;
}

static GeomDev make_geom(const ksgf_boxarray* ba)
{
    GeomDev g;
    for (int d = 0; d < 3; ++d) {
        g.dlo[d] = ba->dlo[d];
        g.dhi[d] = ba->dhi[d];
    }
    return g;
}

static CMFDev ccast(const MFDev& m)
{
    CMFDev c;
    c.p = m.p;
    c.pre = m.pre;
    c.itype = m.itype;
    c.ngrow = m.ngrow;
    c.ncomp = m.ncomp;
    return c;
}

// scratch slot map
enum {
    S_LO = 6,    // 3 slots
    S_HI = 9,    // 3
    S_ED = 12,   // 3
    S_UAD = 15,  // 3
    S_CRN = 18,  // 9 (d*3+t)
};

using Stages = GodunovStages;

static int alloc_stages(const ksgf_boxarray* ba, int ncomp, int ncrn, Stages& S)
{
    int err = KSGF_OK;
    for (int d = 0; d < 3; ++d) {
        const Region* rf = fab_region(ba, KSGF_XFACE + d, 1, &err);
        if (!rf) return err;
        double* p;
        if (!(p = get_scratch(ba, S_LO + d, (size_t)rf->total * ncomp, &err))) return err;
        S.lo[d] = {p, rf->pre, KSGF_XFACE + d, 1, ncomp};
        if (!(p = get_scratch(ba, S_HI + d, (size_t)rf->total * ncomp, &err))) return err;
        S.hi[d] = {p, rf->pre, KSGF_XFACE + d, 1, ncomp};
        if (!(p = get_scratch(ba, S_ED + d, (size_t)rf->total * ncomp, &err))) return err;
        S.ed[d] = {p, rf->pre, KSGF_XFACE + d, 1, ncomp};
        if (!(p = get_scratch(ba, S_UAD + d, (size_t)rf->total, &err))) return err;
        S.uad[d] = {p, rf->pre, KSGF_XFACE + d, 1, 1};
        for (int t = 0; t < 3; ++t) {
            if (t == d) continue;
            if (!(p = get_scratch(ba, S_CRN + 3 * d + t, (size_t)rf->total * ncrn, &err))) return err;
            S.crn[d][t] = {p, rf->pre, KSGF_XFACE + d, 1, ncrn};
        }
    }
    return KSGF_OK;
}

// chunk table of the marching stage-A launch for direction D (faces of D,
// grown by one transversely, the D extent in segments of kMarch), kept with
// the BoxArray's other launch tables under a key outside their range
static const Region* get_seg_region(const ksgf_boxarray* ba, int D, int* err)
{
    RegionKey key;
    for (int d = 0; d < 3; ++d) key.glo[d] = key.ghi[d] = 0;
    key.glo[0] = 100 + D;
    key.glo[1] = kMarch;
    std::lock_guard<std::mutex> lk(ba->mu);
    auto it = ba->regions.find(key);
    if (it != ba->regions.end()) return it->second;
    Region* r = new Region();
    r->key = key;
    r->hpre.resize(ba->nbox + 1);
    std::vector<int> ch;
    long long acc = 0;
    for (int b = 0; b < ba->nbox; ++b) {
        r->hpre[b] = acc;
        long long n = 1;
        for (int d = 0; d < 3; ++d) {
            const int e = ba->hbox[6 * b + 3 + d] - ba->hbox[6 * b + d] + 1 + ((d == D) ? 1 : 2);
            n *= (d == D) ? (e + kMarch - 1) / kMarch : e;
        }
        for (long long s0 = 0; s0 < n; s0 += kChunk) {
            ch.push_back(b);
            ch.push_back((int)s0);
        }
        acc += n;
    }
    r->hpre[ba->nbox] = acc;
    r->total = acc;
    r->nchunks = (int)(ch.size() / 2);
    if (hipMalloc(&r->chunks, sizeof(int) * ch.size()) != hipSuccess ||
        hipMemcpy(r->chunks, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice) != hipSuccess) {
        if (r->chunks) (void)hipFree(r->chunks);
        delete r;
        set_error("segment table upload");
        *err = KSGF_ERR_NOMEM;
        return nullptr;
    }
    ba->regions[key] = r;
    return r;
}

// tile table of the fused corner + final kernel: per box, kCfInX x kCfInY tiles of
// output positions (x in lo..hi+1, y in lo..hi+1), kept with the BoxArray's
// launch tables under a key outside their range
static const Region* get_cf_tiles(const ksgf_boxarray* ba, int* err)
{
    RegionKey key;
    for (int d = 0; d < 3; ++d) key.glo[d] = key.ghi[d] = 0;
    key.glo[0] = 300;
    std::lock_guard<std::mutex> lk(ba->mu);
    auto it = ba->regions.find(key);
    if (it != ba->regions.end()) return it->second;
    std::vector<int> t;
    for (int b = 0; b < ba->nbox; ++b) {
        const int* bx = &ba->hbox[6 * b];
        for (int y0 = bx[1]; y0 <= bx[4] + 1; y0 += kCfInY)
            for (int x0 = bx[0]; x0 <= bx[3] + 1; x0 += kCfInX) {
                t.push_back(b);
                t.push_back(x0);
                t.push_back(y0);
            }
    }
    Region* r = new Region();
    r->key = key;
    r->nchunks = (int)(t.size() / 3);
    if (hipMalloc(&r->chunks, sizeof(int) * t.size()) != hipSuccess ||
        hipMemcpy(r->chunks, t.data(), sizeof(int) * t.size(), hipMemcpyHostToDevice) != hipSuccess) {
        if (r->chunks) (void)hipFree(r->chunks);
        delete r;
        set_error("corner/final tile table upload");
        *err = KSGF_ERR_NOMEM;
        return nullptr;
    }
    ba->regions[key] = r;
    return r;
}

template <bool VEL>
static int launch_corner_final(const ksgf_boxarray* ba, const CfArgs& a, const BCs& bc, const GeomDev& g, int ncomp,
                               hipStream_t st)
{
    int err = KSGF_OK;
    const Region* r = get_cf_tiles(ba, &err);
    if (!r) return err;
    CfTile tl;
    tl.tiles = r->chunks;
    hipLaunchKernelGGL(k_corner_final<VEL>, dim3(r->nchunks, VEL ? 1 : ncomp), dim3(kCfThreads), 0, st, a, tl,
                       ba->dbox, bc, g);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

template <int D>
static int launch_faceA(const ksgf_boxarray* ba, CMFDev q, CMFDev qbc, CMFDev f, int ncomp, CMFDev macD,
                        const Stages& S, const BCs& bc, const GeomDev& g, double dtdx, double hdt, int uft,
                        int vel_mode, int is_velocity, int lim, bool write_ed, hipStream_t st)
{
    int err = KSGF_OK;
    MFDev ed = S.ed[D];
    if (!write_ed) ed.p = nullptr;
    // marching along y and z: consecutive lanes are consecutive x columns, so
    // loads and stores stay coalesced.  Along x the lanes of a march would sit
    // kMarch faces apart (measured 3.5x slower), so x keeps one face per thread.
    if (D > 0 && (lim == KSGF_PPM_VANLEER || lim == KSGF_PPM_NOLIMITER)) {
        const Region* r = get_seg_region(ba, D, &err);
        if (!r) return err;
        SegDev sd;
        sd.chunks = r->chunks;
        sd.glo0 = (D == 0) ? 0 : 1;
        sd.glo1 = (D == 1) ? 0 : 1;
        sd.glo2 = (D == 2) ? 0 : 1;
        sd.ghi0 = sd.ghi1 = sd.ghi2 = 1;
        hipLaunchKernelGGL(k_faceA_march<D>, dim3(r->nchunks), dim3(256), 0, st, q, qbc, f, ncomp, macD, S.lo[D],
                           S.hi[D], ed, S.uad[D], sd, ba->dbox, bc, g, dtdx, hdt, uft, vel_mode, is_velocity,
                           lim);
        KSGF_LAUNCH_CHECK();
        return KSGF_OK;
    }
    int glo[3], ghi[3];
    for (int e = 0; e < 3; ++e) {
        glo[e] = (e == D) ? 0 : 1;
        ghi[e] = 1;
    }
    const Region* r = get_region(ba, glo, ghi, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_faceA<D>, dim3(r->nchunks), dim3(256), 0, st, q, qbc, f, ncomp, macD, S.lo[D],
                       S.hi[D], ed, S.uad[D], regdev(r, ba->nbox), ba->dbox, bc, g, dtdx, hdt, uft, vel_mode,
                       is_velocity, lim);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

template <int D, int T>
static int launch_corner(const ksgf_boxarray* ba, CMFDev q, CMFDev qbc, int ncomp, CMFDev velT, CMFDev velD,
                         const Stages& S, const BCs& bc, const GeomDev& g, double c3, double c6, int vel_mode,
                         int is_velocity, int cons_mask, hipStream_t st)
{
    constexpr int E = 3 - D - T;
    int err = KSGF_OK;
    int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
    ghi[D] = 1;
    glo[E] = 1;
    ghi[E] += 1;
    const Region* r = get_region(ba, glo, ghi, &err);
    if (!r) return err;
    hipLaunchKernelGGL((k_corner<D, T>), dim3(r->nchunks), dim3(256), 0, st, q, qbc, ncomp,
                       ccast(S.lo[D]), ccast(S.hi[D]), ccast(S.ed[T]), velT, velD, S.crn[D][T],
                       regdev(r, ba->nbox), ba->dbox, bc, g, c3, c6, vel_mode, is_velocity, (const int*)nullptr,
                       cons_mask);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

template <int D>
static int launch_final(const ksgf_boxarray* ba, CMFDev q, CMFDev qbc, CMFDev f, CMFDev divu, int has_divu,
                        int ncomp, CMFDev vel1, CMFDev vel2, CMFDev velD, MFDev out_state, MFDev out_flux,
                        const Stages& S, const BCs& bc, const GeomDev& g, const double* c2, const double* c4,
                        double hdt, double area, int uft, int vel_mode, int is_velocity, int area_weighted,
                        int allow_inflow, int cons_mask, hipStream_t st)
{
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    int err = KSGF_OK;
    int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
    ghi[D] = 1;
    const Region* r = get_region(ba, glo, ghi, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_final<D>, dim3(r->nchunks), dim3(256), 0, st, q, qbc, f, divu, has_divu,
                       ncomp, ccast(S.lo[D]), ccast(S.hi[D]), ccast(S.crn[T1][T2]), ccast(S.crn[T2][T1]), vel1,
                       vel2, velD, out_state, out_flux, regdev(r, ba->nbox), ba->dbox, bc, g, c2[T1], c2[T2],
                       c4[T1], c4[T2], hdt, area, uft, vel_mode, is_velocity, area_weighted, allow_inflow,
                       cons_mask);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

#define KSGF_TRY(x)                 \
    do {                            \
        int e__ = (x);              \
        if (e__ != KSGF_OK) return e__; \
    } while (0)

}  // namespace ksgf

using namespace ksgf;

extern "C" {

int ksgf_ppm_predict_state_on_face(const ksgf_boxarray* ba, int dir, const double* s, int ncomp_s, int ng_s,
                                   int comp, const double* vel_edge, int ng_v, const ksgf_bcrec* bc, double dt,
                                   int limiter_type, double* Im, double* Ip, ksgf_stream_t stream)
{
    if (!ba || !s || !vel_edge || !bc || !Im || !Ip || dir < 0 || dir > 2 || comp < 0 || comp >= ncomp_s ||
        ng_s < 2 || ng_v < 0) {
        set_error("ksgf_ppm_predict_state_on_face: bad argument");
        return KSGF_ERR_INVALID;
    }
    KSGF_TRY(check_lim(limiter_type));
    KSGF_TRY(check_bc(bc, 1, ba));
    CMFDev sv, vv;
    MFDev im, ip;
    KSGF_TRY(cmfdev(ba, s, KSGF_CELL, ncomp_s, ng_s, &sv));
    KSGF_TRY(cmfdev(ba, vel_edge, KSGF_XFACE + dir, 1, ng_v, &vv));
    KSGF_TRY(mfdev(ba, Im, KSGF_CELL, 1, 0, &im));
    KSGF_TRY(mfdev(ba, Ip, KSGF_CELL, 1, 0, &ip));
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    double dtdx = dt / ba->dx[dir];
    hipStream_t st = (hipStream_t)stream;
    dim3 gr(r->nchunks);
    if (dir == 0)
        hipLaunchKernelGGL(k_ppm_predict<0>, gr, dim3(256), 0, st, sv, comp, vv, im, ip, regdev(r, ba->nbox),
                           ba->dbox, bc->lo[0], bc->hi[0], ba->dlo[0], ba->dhi[0], dtdx, limiter_type);
    else if (dir == 1)
        hipLaunchKernelGGL(k_ppm_predict<1>, gr, dim3(256), 0, st, sv, comp, vv, im, ip, regdev(r, ba->nbox),
                           ba->dbox, bc->lo[1], bc->hi[1], ba->dlo[1], ba->dhi[1], dtdx, limiter_type);
    else
        hipLaunchKernelGGL(k_ppm_predict<2>, gr, dim3(256), 0, st, sv, comp, vv, im, ip, regdev(r, ba->nbox),
                           ba->dbox, bc->lo[2], bc->hi[2], ba->dlo[2], ba->dhi[2], dtdx, limiter_type);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

int ksgf_extrap_vel_to_faces(const ksgf_boxarray* ba, const double* vel, const double* vel_forces, double* u_mac,
                             double* v_mac, double* w_mac, const ksgf_bcrec* bcrec, double dt, int use_ppm,
                             int use_forces_in_trans, int limiter_type, int allow_inflow_on_outflow,
                             ksgf_stream_t stream)
{
    return extrap_vel_to_faces_keep(ba, vel, vel_forces, u_mac, v_mac, w_mac, bcrec, dt, use_ppm, use_forces_in_trans,
                                    limiter_type, allow_inflow_on_outflow, stream, nullptr);
}

int ksgf_compute_fluxes_from_state(const ksgf_boxarray* ba, int ncomp, const double* q, const double* q_on_face_bc,
                                   double* flux_x, double* flux_y, double* flux_z, double* face_x, double* face_y,
                                   double* face_z, int known_edge_state, const double* u_mac, const double* v_mac,
                                   const double* w_mac, const double* divu, const double* fq, double dt,
                                   const ksgf_bcrec* bcrec, const int* iconserv, int use_ppm,
                                   int use_forces_in_trans, int is_velocity, int fluxes_are_area_weighted,
                                   int limiter_type, int allow_inflow_on_outflow, ksgf_stream_t stream)
{
    return compute_fluxes_keep(ba, ncomp, q, q_on_face_bc, flux_x, flux_y, flux_z, face_x, face_y, face_z,
                               known_edge_state, u_mac, v_mac, w_mac, divu, fq, dt, bcrec, iconserv, use_ppm,
                               use_forces_in_trans, is_velocity, fluxes_are_area_weighted, limiter_type,
                               allow_inflow_on_outflow, stream, nullptr);
}

}  // extern "C"

namespace ksgf {

int extrap_vel_to_faces_keep(const ksgf_boxarray* ba, const double* vel, const double* vel_forces, double* u_mac,
                             double* v_mac, double* w_mac, const ksgf_bcrec* bcrec, double dt, int use_ppm,
                             int use_forces_in_trans, int limiter_type, int allow_inflow_on_outflow,
                             ksgf_stream_t stream, GodunovStages* keep)
{
    if (!ba || !vel || !vel_forces || !u_mac || !v_mac || !w_mac || !bcrec) {
        set_error("ksgf_extrap_vel_to_faces: null argument");
        return KSGF_ERR_INVALID;
    }
    if (!use_ppm) {
        set_error("ksgf_extrap_vel_to_faces: PLM (use_ppm = false) is not provided by this build");
        return KSGF_ERR_UNSUPPORTED;
    }
    KSGF_TRY(check_lim(limiter_type));
    KSGF_TRY(check_bc(bcrec, 3, ba));
    hipStream_t st = (hipStream_t)stream;
    CMFDev q, f;
    KSGF_TRY(cmfdev(ba, vel, KSGF_CELL, 3, 3, &q));
    KSGF_TRY(cmfdev(ba, vel_forces, KSGF_CELL, 3, 1, &f));
    MFDev mac[3];
    double* macp[3] = {u_mac, v_mac, w_mac};
    for (int d = 0; d < 3; ++d) KSGF_TRY(mfdev(ba, macp[d], KSGF_XFACE + d, 1, 1, &mac[d]));
    Stages S;
    KSGF_TRY(alloc_stages(ba, 3, 1, S));
    BCs bc = make_bcs(bcrec, 3);
    GeomDev g = make_geom(ba);
    double dtdx[3], c6[3], c4[3], c2[3];
    for (int d = 0; d < 3; ++d) {
        dtdx[d] = dt / ba->dx[d];
        c6[d] = dt / (6.0 * ba->dx[d]);
        c4[d] = dt / (4.0 * ba->dx[d]);
        c2[d] = dt / (2.0 * ba->dx[d]);
    }
    const double hdt = 0.5 * dt;
    // stage A of the x and y faces (without the upwinded states when the
    // fused corner + final kernel forms them, and the z faces, from the
    // states); every stage of every direction for the staged kernels
    KSGF_TRY(launch_faceA<0>(ba, q, q, f, 3, q, S, bc, g, dtdx[0], hdt, use_forces_in_trans, 1, 1, limiter_type,
                             keep != nullptr, st));
    KSGF_TRY(launch_faceA<1>(ba, q, q, f, 3, q, S, bc, g, dtdx[1], hdt, use_forces_in_trans, 1, 1, limiter_type,
                             keep != nullptr, st));
    KSGF_TRY(launch_faceA<2>(ba, q, q, f, 3, q, S, bc, g, dtdx[2], hdt, use_forces_in_trans, 1, 1, limiter_type,
                             keep != nullptr, st));
    CMFDev uad[3];
    for (int d = 0; d < 3; ++d) uad[d] = ccast(S.uad[d]);
    if (!keep) {
        // corner and final states in one pass, the corner states kept in LDS
        CfArgs a{};
        a.q = q;
        a.qbc = q;
        a.f = f;
        a.divu = q;
        a.has_divu = 0;
        a.ncomp = 3;
        for (int d = 0; d < 3; ++d) {
            a.lo.m[d] = ccast(S.lo[d]);
            a.hi.m[d] = ccast(S.hi[d]);
            a.vel.m[d] = uad[d];
            a.state.m[d] = mac[d];
            a.flux.m[d] = mac[d];
            a.c6[d] = c6[d];
            a.c4[d] = c4[d];
            a.c2[d] = c2[d];
            a.c3[d] = 0.0;
            a.area[d] = 0.0;
        }
        a.hdt = hdt;
        a.uft = use_forces_in_trans;
        a.is_velocity = 1;
        a.area_weighted = 0;
        a.allow_inflow = allow_inflow_on_outflow;
        a.cons_mask = 0;
        return launch_corner_final<true>(ba, a, bc, g, 3, st);
    }
    KSGF_TRY((launch_corner<0, 1>(ba, q, q, 3, uad[1], uad[0], S, bc, g, 0.0, c6[1], 1, 1, 0, st)));
    KSGF_TRY((launch_corner<0, 2>(ba, q, q, 3, uad[2], uad[0], S, bc, g, 0.0, c6[2], 1, 1, 0, st)));
    KSGF_TRY((launch_corner<1, 0>(ba, q, q, 3, uad[0], uad[1], S, bc, g, 0.0, c6[0], 1, 1, 0, st)));
    KSGF_TRY((launch_corner<1, 2>(ba, q, q, 3, uad[2], uad[1], S, bc, g, 0.0, c6[2], 1, 1, 0, st)));
    KSGF_TRY((launch_corner<2, 0>(ba, q, q, 3, uad[0], uad[2], S, bc, g, 0.0, c6[0], 1, 1, 0, st)));
    KSGF_TRY((launch_corner<2, 1>(ba, q, q, 3, uad[1], uad[2], S, bc, g, 0.0, c6[1], 1, 1, 0, st)));
    KSGF_TRY(launch_final<0>(ba, q, q, f, q, 0, 3, uad[1], uad[2], uad[0], mac[0], mac[0], S, bc, g, c2, c4, hdt,
                             0.0, use_forces_in_trans, 1, 1, 0, allow_inflow_on_outflow, 0, st));
    KSGF_TRY(launch_final<1>(ba, q, q, f, q, 0, 3, uad[2], uad[0], uad[1], mac[1], mac[1], S, bc, g, c2, c4, hdt,
                             0.0, use_forces_in_trans, 1, 1, 0, allow_inflow_on_outflow, 0, st));
    KSGF_TRY(launch_final<2>(ba, q, q, f, q, 0, 3, uad[0], uad[1], uad[2], mac[2], mac[2], S, bc, g, c2, c4, hdt,
                             0.0, use_forces_in_trans, 1, 1, 0, allow_inflow_on_outflow, 0, st));
    if (keep) *keep = S;
    return KSGF_OK;
}

int compute_fluxes_keep(const ksgf_boxarray* ba, int ncomp, const double* q, const double* q_on_face_bc,
                        double* flux_x, double* flux_y, double* flux_z, double* face_x, double* face_y,
                        double* face_z, int known_edge_state, const double* u_mac, const double* v_mac,
                        const double* w_mac, const double* divu, const double* fq, double dt,
                        const ksgf_bcrec* bcrec, const int* iconserv, int use_ppm, int use_forces_in_trans,
                        int is_velocity, int fluxes_are_area_weighted, int limiter_type,
                        int allow_inflow_on_outflow, ksgf_stream_t stream, GodunovStages* keep)
{
    if (!ba || ncomp <= 0 || ncomp > kMaxComp || !flux_x || !flux_y || !flux_z || !face_x || !face_y || !face_z ||
        !u_mac || !v_mac || !w_mac) {
        set_error("ksgf_compute_fluxes_from_state: bad argument");
        return KSGF_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    const double area[3] = {ba->dx[1] * ba->dx[2], ba->dx[0] * ba->dx[2], ba->dx[0] * ba->dx[1]};
    CMFDev mac[3];
    const double* macp[3] = {u_mac, v_mac, w_mac};
    for (int d = 0; d < 3; ++d) KSGF_TRY(cmfdev(ba, macp[d], KSGF_XFACE + d, 1, 1, &mac[d]));
    MFDev flux[3], face[3];
    double* fluxp[3] = {flux_x, flux_y, flux_z};
    double* facep[3] = {face_x, face_y, face_z};
    for (int d = 0; d < 3; ++d) {
        KSGF_TRY(mfdev(ba, fluxp[d], KSGF_XFACE + d, ncomp, 0, &flux[d]));
        KSGF_TRY(mfdev(ba, facep[d], KSGF_XFACE + d, ncomp, 0, &face[d]));
    }
    int err = KSGF_OK;
    if (known_edge_state) {
        for (int d = 0; d < 3; ++d) {
            int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
            ghi[d] = 1;
            const Region* r = get_region(ba, glo, ghi, &err);
            if (!r) return err;
            dim3 gr(r->nchunks);
            if (d == 0)
                hipLaunchKernelGGL(k_known_flux<0>, gr, dim3(256), 0, st, ccast(face[0]), mac[0], flux[0], ncomp,
                                   regdev(r, ba->nbox), ba->dbox, area[0], fluxes_are_area_weighted);
            else if (d == 1)
                hipLaunchKernelGGL(k_known_flux<1>, gr, dim3(256), 0, st, ccast(face[1]), mac[1], flux[1], ncomp,
                                   regdev(r, ba->nbox), ba->dbox, area[1], fluxes_are_area_weighted);
            else
                hipLaunchKernelGGL(k_known_flux<2>, gr, dim3(256), 0, st, ccast(face[2]), mac[2], flux[2], ncomp,
                                   regdev(r, ba->nbox), ba->dbox, area[2], fluxes_are_area_weighted);
            KSGF_LAUNCH_CHECK();
        }
        return KSGF_OK;
    }
    if (!q || !q_on_face_bc || !fq || !bcrec || !iconserv) {
        set_error("ksgf_compute_fluxes_from_state: null state/forcing/bc argument");
        return KSGF_ERR_INVALID;
    }
    if (!use_ppm) {
        set_error("ksgf_compute_fluxes_from_state: PLM (use_ppm = false) is not provided by this build");
        return KSGF_ERR_UNSUPPORTED;
    }
    KSGF_TRY(check_lim(limiter_type));
    KSGF_TRY(check_bc(bcrec, ncomp, ba));
    CMFDev qv, qb, fv, dv;
    KSGF_TRY(cmfdev(ba, q, KSGF_CELL, ncomp, 3, &qv));
    KSGF_TRY(cmfdev(ba, q_on_face_bc, KSGF_CELL, ncomp, 3, &qb));
    KSGF_TRY(cmfdev(ba, fq, KSGF_CELL, ncomp, 1, &fv));
    int has_divu = divu ? 1 : 0;
    if (divu)
        KSGF_TRY(cmfdev(ba, divu, KSGF_CELL, 1, 1, &dv));
    else
        dv = qv;
    int cons_mask = 0;
    for (int n = 0; n < ncomp; ++n)
        if (iconserv[n]) cons_mask |= (1 << n);
    Stages S;
    KSGF_TRY(alloc_stages(ba, ncomp, ncomp, S));
    BCs bc = make_bcs(bcrec, ncomp);
    GeomDev g = make_geom(ba);
    double dtdx[3], c3[3], c6[3], c2[3], c4[3];
    for (int d = 0; d < 3; ++d) {
        dtdx[d] = dt / ba->dx[d];
        c3[d] = dt / (3.0 * ba->dx[d]);
        c6[d] = dt / (6.0 * ba->dx[d]);
        c2[d] = dt / (2.0 * ba->dx[d]);
        c4[d] = dt / (4.0 * ba->dx[d]);
    }
    const double hdt = 0.5 * dt;
    KSGF_TRY(launch_faceA<0>(ba, qv, qb, fv, ncomp, mac[0], S, bc, g, dtdx[0], hdt, use_forces_in_trans, 0,
                             is_velocity, limiter_type, keep != nullptr, st));
    KSGF_TRY(launch_faceA<1>(ba, qv, qb, fv, ncomp, mac[1], S, bc, g, dtdx[1], hdt, use_forces_in_trans, 0,
                             is_velocity, limiter_type, keep != nullptr, st));
    KSGF_TRY(launch_faceA<2>(ba, qv, qb, fv, ncomp, mac[2], S, bc, g, dtdx[2], hdt, use_forces_in_trans, 0,
                             is_velocity, limiter_type, keep != nullptr, st));
    if (!keep) {
        // corner and final states in one pass per component, the corner
        // states kept in LDS
        CfArgs a{};
        a.q = qv;
        a.qbc = qb;
        a.f = fv;
        a.divu = dv;
        a.has_divu = has_divu;
        a.ncomp = ncomp;
        for (int d = 0; d < 3; ++d) {
            a.lo.m[d] = ccast(S.lo[d]);
            a.hi.m[d] = ccast(S.hi[d]);
            a.vel.m[d] = mac[d];
            a.state.m[d] = face[d];
            a.flux.m[d] = flux[d];
            a.c3[d] = c3[d];
            a.c6[d] = c6[d];
            a.c4[d] = c4[d];
            a.c2[d] = c2[d];
            a.area[d] = area[d];
        }
        a.hdt = hdt;
        a.uft = use_forces_in_trans;
        a.is_velocity = is_velocity;
        a.area_weighted = fluxes_are_area_weighted;
        a.allow_inflow = allow_inflow_on_outflow;
        a.cons_mask = cons_mask;
        return launch_corner_final<false>(ba, a, bc, g, ncomp, st);
    }
    KSGF_TRY((launch_corner<0, 1>(ba, qv, qb, ncomp, mac[1], mac[0], S, bc, g, c3[1], c6[1], 0, is_velocity, cons_mask, st)));
    KSGF_TRY((launch_corner<0, 2>(ba, qv, qb, ncomp, mac[2], mac[0], S, bc, g, c3[2], c6[2], 0, is_velocity, cons_mask, st)));
    KSGF_TRY((launch_corner<1, 0>(ba, qv, qb, ncomp, mac[0], mac[1], S, bc, g, c3[0], c6[0], 0, is_velocity, cons_mask, st)));
    KSGF_TRY((launch_corner<1, 2>(ba, qv, qb, ncomp, mac[2], mac[1], S, bc, g, c3[2], c6[2], 0, is_velocity, cons_mask, st)));
    KSGF_TRY((launch_corner<2, 0>(ba, qv, qb, ncomp, mac[0], mac[2], S, bc, g, c3[0], c6[0], 0, is_velocity, cons_mask, st)));
    KSGF_TRY((launch_corner<2, 1>(ba, qv, qb, ncomp, mac[1], mac[2], S, bc, g, c3[1], c6[1], 0, is_velocity, cons_mask, st)));
    KSGF_TRY(launch_final<0>(ba, qv, qb, fv, dv, has_divu, ncomp, mac[1], mac[2], mac[0], face[0], flux[0], S, bc, g,
                             c2, c4, hdt, area[0], use_forces_in_trans, 0, is_velocity, fluxes_are_area_weighted,
                             allow_inflow_on_outflow, cons_mask, st));
    KSGF_TRY(launch_final<1>(ba, qv, qb, fv, dv, has_divu, ncomp, mac[2], mac[0], mac[1], face[1], flux[1], S, bc, g,
                             c2, c4, hdt, area[1], use_forces_in_trans, 0, is_velocity, fluxes_are_area_weighted,
                             allow_inflow_on_outflow, cons_mask, st));
    KSGF_TRY(launch_final<2>(ba, qv, qb, fv, dv, has_divu, ncomp, mac[0], mac[1], mac[2], face[2], flux[2], S, bc, g,
                             c2, c4, hdt, area[2], use_forces_in_trans, 0, is_velocity, fluxes_are_area_weighted,
                             allow_inflow_on_outflow, cons_mask, st));
    if (keep) *keep = S;
    return KSGF_OK;
}

}  // namespace ksgf

extern "C" {

int ksgf_compute_divergence(const ksgf_boxarray* ba, double* div, const double* flux_x, const double* flux_y,
                            const double* flux_z, int ncomp, double mult, int fluxes_are_area_weighted,
                            ksgf_stream_t stream)
{
    if (!ba || !div || !flux_x || !flux_y || !flux_z || ncomp <= 0) {
        set_error("ksgf_compute_divergence: bad argument");
        return KSGF_ERR_INVALID;
    }
    MFDev dv;
    CMF3 fl;
    KSGF_TRY(mfdev(ba, div, KSGF_CELL, ncomp, 0, &dv));
    const double* fp[3] = {flux_x, flux_y, flux_z};
    for (int d = 0; d < 3; ++d) KSGF_TRY(cmfdev(ba, fp[d], KSGF_XFACE + d, ncomp, 0, &fl.m[d]));
    int err = KSGF_OK;
    int z[3] = {0, 0, 0};
    const Region* r = get_region(ba, z, z, &err);
    if (!r) return err;
    hipLaunchKernelGGL(k_divergence, dim3(r->nchunks), dim3(256), 0, (hipStream_t)stream, dv, fl, ncomp,
                       regdev(r, ba->nbox), ba->dbox, mult, 1.0 / ba->dx[0], 1.0 / ba->dx[1], 1.0 / ba->dx[2],
                       ba->dx[0] * ba->dx[1] * ba->dx[2], fluxes_are_area_weighted);
    KSGF_LAUNCH_CHECK();
    return KSGF_OK;
}

}  // extern "C"
