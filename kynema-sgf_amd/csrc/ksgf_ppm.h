// Device-side PPM edge reconstruction, traces, Riemann/upwind selectors and
// edge boundary conditions used by the Godunov kernels.
//
// Method: Colella & Woodward (1984) piecewise-parabolic reconstruction with
// the monotonised-central (van Leer) limited fourth-order edge interpolant and
// the CW84 monotonicity constraints; unsplit corner transport after Colella
// (1990) and Saltzman (1994).  Boundary cells of ext_dir/hoextrap faces use a
// cubic through the ghost (face) value and three interior cell values.
// small_vel = 1e-8 decides "no upwind direction" everywhere.
//
// The expressions are written in a fixed operand order and the library is
// compiled without floating-point contraction, so results are reproducible
// bit for bit against a CPU evaluation of the same formulas.
//
// The reconstruction is written once for a number type T: double in the
// forward kernels, Dual<N> (ksgf_dual.h) where the adjoint needs the local
// Jacobian of a trace.  Branches compare values only, so a Dual evaluation
// takes the branches of the double evaluation at the same point.
#pragma once

#include "../../include/ksgf.h"
#include "ksgf_dual.h"
#include "ksgf_hd.h"

namespace ksgf {

constexpr double kSmallVel = 1.0e-8;

template <typename T>
KSGF_HD T vanleer(const T& a, const T& b, const T& c)
{
    T dsc = 0.5 * (c - a);
    T dsl = 2.0 * (b - a);
    T dsr = 2.0 * (c - b);
    if (dsl * dsr > 0.0) {
        T m = vmin(vabs(dsc), vmin(vabs(dsl), vabs(dsr)));
        return vcopysign(m, dsc);
    }
    return zero_like(dsc);
}

template <typename T>
KSGF_HD T clip2(T e, const T& a, const T& b)
{
    e = vmax(e, vmin(a, b));
    return vmin(e, vmax(a, b));
}

template <typename T>
KSGF_HD void mono(const T& s0, T& sm, T& sp)
{
    if ((sp - s0) * (s0 - sm) <= 0.0) {
        sp = s0;
        sm = s0;
    } else if (vabs(sp - s0) >= 2.0 * vabs(sm - s0)) {
        sp = 3.0 * s0 - 2.0 * sm;
    } else if (vabs(sm - s0) >= 2.0 * vabs(sp - s0)) {
        sm = 3.0 * s0 - 2.0 * sp;
    }
}

template <typename T>
KSGF_HD T bnd_edge(const T& sb, const T& s0, const T& s1, const T& s2)
{
    return -0.2 * sb + 0.75 * s0 + 0.5 * s1 - 0.05 * s2;
}

// interior value of the edge between cells a and b (stencil aa, a, b, bb):
// fourth-order interpolant, or with the van Leer slopes da, db of a and b
// (limited) the slope-corrected average clipped to [a, b]
template <typename T>
KSGF_HD T ppm_edge4(const T& aa, const T& a, const T& b, const T& bb)
{
    return (7.0 / 12.0) * (a + b) - (1.0 / 12.0) * (aa + bb);
}
template <typename T>
KSGF_HD T ppm_edge_vl(const T& a, const T& b, const T& da, const T& db)
{
    return clip2(0.5 * (a + b) - (db - da) / 6.0, a, b);
}

// boundary treatment and monotonisation of the cell's interior edge values
// (e1 low, e2 high) into (sm, sp)
template <typename T>
KSGF_HD void ppm_edges_finish(const T& e1, const T& e2, const T& sm2, const T& sm1, const T& s0, const T& sp1,
                              const T& sp2, int lim, int bclo, int bchi, int x, int dlo, int dhi, T& sm, T& sp)
{
    const bool limited = (lim != KSGF_PPM_NOLIMITER);
    const bool lw = (bclo == KSGF_BC_EXT_DIR || bclo == KSGF_BC_HOEXTRAP);
    const bool hw = (bchi == KSGF_BC_EXT_DIR || bchi == KSGF_BC_HOEXTRAP);
    if (lw && x == dlo) {
        sm = sm1;
        sp = bnd_edge(sm1, s0, sp1, sp2);
        if (limited) sp = clip2(sp, s0, sp1);
    } else if (lw && x == dlo + 1) {
        sm = bnd_edge(sm2, sm1, s0, sp1);
        if (limited) sm = clip2(sm, sm1, s0);
        sp = e2;
        if (limited) mono(s0, sm, sp);
    } else if (hw && x == dhi) {
        sp = sp1;
        sm = bnd_edge(sp1, s0, sm1, sm2);
        if (limited) sm = clip2(sm, sm1, s0);
    } else if (hw && x == dhi - 1) {
        sp = bnd_edge(sp2, sp1, s0, sm1);
        if (limited) sp = clip2(sp, s0, sp1);
        sm = e1;
        if (limited) mono(s0, sm, sp);
    } else {
        sm = e1;
        sp = e2;
        if (limited) mono(s0, sm, sp);
    }
}

// edge values of the cell whose 5-point stencil is (sm2 .. sp2)
template <typename T>
KSGF_HD void ppm_edges(const T& sm2, const T& sm1, const T& s0, const T& sp1, const T& sp2, int lim, int bclo,
                       int bchi, int x, int dlo, int dhi, T& sm, T& sp)
{
    T e1, e2;
    if (lim == KSGF_PPM_NOLIMITER) {
        e1 = ppm_edge4(sm2, sm1, s0, sp1);
        e2 = ppm_edge4(sm1, s0, sp1, sp2);
    } else {
        T d0 = vanleer(sm2, sm1, s0);
        T d1 = vanleer(sm1, s0, sp1);
        T d2 = vanleer(s0, sp1, sp2);
        e1 = ppm_edge_vl(sm1, s0, d0, d1);
        e2 = ppm_edge_vl(s0, sp1, d1, d2);
    }
    ppm_edges_finish(e1, e2, sm2, sm1, s0, sp1, sp2, lim, bclo, bchi, x, dlo, dhi, sm, sp);
}

template <typename T>
KSGF_HD T minmod_slope(const T& sm1, const T& s0, const T& sp1, int bclo, int bchi, int x, int dlo, int dhi)
{
    const bool lw = (bclo == KSGF_BC_EXT_DIR || bclo == KSGF_BC_HOEXTRAP);
    const bool hw = (bchi == KSGF_BC_EXT_DIR || bchi == KSGF_BC_HOEXTRAP);
    if (lw && x == dlo) return s0 - sm1;
    if (hw && x == dhi) return sp1 - s0;
    T a = s0 - sm1, b = sp1 - s0;
    if (a * b <= 0.0) return zero_like(a);
    return (vabs(a) < vabs(b)) ? a : b;
}

// time-centred state at the high face (advected by up > 0) and at the low
// face (um < 0) of a parabola (sm, s0, sp)
template <typename T>
KSGF_HD T ppm_ip(const T& sm, const T& s0, const T& sp, const T& up, double dtdx)
{
    if (up > kSmallVel) {
        const T s6 = 6.0 * s0 - 3.0 * (sm + sp);
        const T sig = vabs(up) * dtdx;
        return sp - (0.5 * sig) * ((sp - sm) - (1.0 - (2.0 / 3.0) * sig) * s6);
    }
    return s0;
}
template <typename T>
KSGF_HD T ppm_im(const T& sm, const T& s0, const T& sp, const T& um, double dtdx)
{
    if (um < -kSmallVel) {
        const T s6 = 6.0 * s0 - 3.0 * (sm + sp);
        const T sig = vabs(um) * dtdx;
        return sm + (0.5 * sig) * ((sp - sm) + (1.0 - (2.0 / 3.0) * sig) * s6);
    }
    return s0;
}

// time-centred states at the low (Im) and high (Ip) faces of one cell;
// um/up: advecting velocity at the low/high face
template <typename T>
KSGF_HD void ppm_trace(const T& sm2, const T& sm1, const T& s0, const T& sp1, const T& sp2, int lim, int bclo,
                       int bchi, int x, int dlo, int dhi, const T& um, const T& up, double dtdx, T& Im, T& Ip)
{
    if (lim == KSGF_PPM_UPWIND) {
        Im = s0;
        Ip = s0;
        return;
    }
    if (lim == KSGF_PPM_MINMOD) {
        T sl = minmod_slope(sm1, s0, sp1, bclo, bchi, x, dlo, dhi);
        Ip = (up > kSmallVel) ? s0 + 0.5 * (1.0 - up * dtdx) * sl : s0;
        Im = (um < -kSmallVel) ? s0 - 0.5 * (1.0 + um * dtdx) * sl : s0;
        return;
    }
    T sm, sp;
    ppm_edges(sm2, sm1, s0, sp1, sp2, lim, bclo, bchi, x, dlo, dhi, sm, sp);
    Ip = ppm_ip(sm, s0, sp, up, dtdx);
    Im = ppm_im(sm, s0, sp, um, dtdx);
}

// the two traces one face needs from the six cells x-3 .. x+2 around it
// (cell q = x-1 below, cell p = x above): Ip of q and Im of p.  The van
// Leer slopes of the four middle cells and the interior edge between q and
// p are shared by the two parabolas; every value equals the one ppm_trace
// forms for the same cell.
template <typename T>
KSGF_HD void ppm_trace_pair(const T* s, int lim, int bclo, int bchi, int x, int dlo, int dhi, const T& uqm,
                            const T& uqp, const T& upm, const T& upp, double dtdx, T& Ip_q, T& Im_p)
{
    if (lim != KSGF_PPM_VANLEER && lim != KSGF_PPM_NOLIMITER) {
        T dmy;
        ppm_trace(s[0], s[1], s[2], s[3], s[4], lim, bclo, bchi, x - 1, dlo, dhi, uqm, uqp, dtdx, dmy, Ip_q);
        ppm_trace(s[1], s[2], s[3], s[4], s[5], lim, bclo, bchi, x, dlo, dhi, upm, upp, dtdx, Im_p, dmy);
        return;
    }
    T ea, eb, ec;  // interior edges q-1|q, q|p, p|p+1
    if (lim == KSGF_PPM_NOLIMITER) {
        ea = ppm_edge4(s[0], s[1], s[2], s[3]);
        eb = ppm_edge4(s[1], s[2], s[3], s[4]);
        ec = ppm_edge4(s[2], s[3], s[4], s[5]);
    } else {
        const T d1 = vanleer(s[0], s[1], s[2]);
        const T d2 = vanleer(s[1], s[2], s[3]);
        const T d3 = vanleer(s[2], s[3], s[4]);
        const T d4 = vanleer(s[3], s[4], s[5]);
        ea = ppm_edge_vl(s[1], s[2], d1, d2);
        eb = ppm_edge_vl(s[2], s[3], d2, d3);
        ec = ppm_edge_vl(s[3], s[4], d3, d4);
    }
    T sm, sp;
    ppm_edges_finish(ea, eb, s[0], s[1], s[2], s[3], s[4], lim, bclo, bchi, x - 1, dlo, dhi, sm, sp);
    Ip_q = ppm_ip(sm, s[2], sp, uqp, dtdx);
    ppm_edges_finish(eb, ec, s[1], s[2], s[3], s[4], s[5], lim, bclo, bchi, x, dlo, dhi, sm, sp);
    Im_p = ppm_im(sm, s[3], sp, upm, dtdx);
}

KSGF_HD double burgers(double lo, double hi)
{
    double st = ((lo + hi) >= 0.0) ? lo : hi;
    bool ltm = ((lo <= 0.0 && hi >= 0.0) || (fabs(lo + hi) < kSmallVel));
    return ltm ? 0.0 : st;
}

// derivatives of burgers(lo, hi) in the branch it takes
KSGF_HD void burgers_partials(double lo, double hi, double& dlo, double& dhi)
{
    const bool ltm = ((lo <= 0.0 && hi >= 0.0) || (fabs(lo + hi) < kSmallVel));
    const bool left = (lo + hi) >= 0.0;
    dlo = (!ltm && left) ? 1.0 : 0.0;
    dhi = (!ltm && !left) ? 1.0 : 0.0;
}

KSGF_HD double upwind(double lo, double hi, double u)
{
    if (fabs(u) < kSmallVel) return 0.5 * (lo + hi);
    return (u >= 0.0) ? lo : hi;
}

// derivatives of upwind(lo, hi, u) with respect to lo and hi (the selector
// is piecewise constant in u)
KSGF_HD void upwind_partials(double u, double& dlo, double& dhi)
{
    if (fabs(u) < kSmallVel) {
        dlo = 0.5;
        dhi = 0.5;
    } else {
        dlo = (u >= 0.0) ? 1.0 : 0.0;
        dhi = (u >= 0.0) ? 0.0 : 1.0;
    }
}

KSGF_HD void edge_bcs(int x, int dlo, int dhi, int bclo, int bchi, double glo, double ghi, bool both_dir,
                      double& lo, double& hi)
{
    if (x == dlo) {
        if (bclo == KSGF_BC_EXT_DIR) {
            lo = glo;
            if (both_dir) hi = lo;
        } else if (bclo == KSGF_BC_FOEXTRAP || bclo == KSGF_BC_HOEXTRAP) {
            lo = hi;
        }
    }
    if (x == dhi + 1) {
        if (bchi == KSGF_BC_EXT_DIR) {
            hi = ghi;
            if (both_dir) lo = hi;
        } else if (bchi == KSGF_BC_FOEXTRAP || bchi == KSGF_BC_HOEXTRAP) {
            hi = lo;
        }
    }
}

}  // namespace ksgf
