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
#pragma once

#include "../../include/ksgf.h"
#include "ksgf_hd.h"

namespace ksgf {

constexpr double kSmallVel = 1.0e-8;

KSGF_HD double vanleer(double a, double b, double c)
{
    double dsc = 0.5 * (c - a);
    double dsl = 2.0 * (b - a);
    double dsr = 2.0 * (c - b);
    if (dsl * dsr > 0.0) {
        double m = fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
        return copysign(m, dsc);
    }
    return 0.0;
}

KSGF_HD double clip2(double e, double a, double b)
{
    e = fmax(e, fmin(a, b));
    return fmin(e, fmax(a, b));
}

KSGF_HD void mono(double s0, double& sm, double& sp)
{
    if ((sp - s0) * (s0 - sm) <= 0.0) {
        sp = s0;
        sm = s0;
    } else if (fabs(sp - s0) >= 2.0 * fabs(sm - s0)) {
        sp = 3.0 * s0 - 2.0 * sm;
    } else if (fabs(sm - s0) >= 2.0 * fabs(sp - s0)) {
        sm = 3.0 * s0 - 2.0 * sp;
    }
}

KSGF_HD double bnd_edge(double sb, double s0, double s1, double s2)
{
    return -0.2 * sb + 0.75 * s0 + 0.5 * s1 - 0.05 * s2;
}

// edge values of the cell whose 5-point stencil is (sm2 .. sp2)
KSGF_HD void ppm_edges(double sm2, double sm1, double s0, double sp1, double sp2,
                                          int lim, int bclo, int bchi, int x, int dlo, int dhi,
                                          double& sm, double& sp)
{
    double e1, e2;
    const bool limited = (lim != KSGF_PPM_NOLIMITER);
    if (!limited) {
        e1 = (7.0 / 12.0) * (sm1 + s0) - (1.0 / 12.0) * (sm2 + sp1);
        e2 = (7.0 / 12.0) * (s0 + sp1) - (1.0 / 12.0) * (sm1 + sp2);
    } else {
        double d0 = vanleer(sm2, sm1, s0);
        double d1 = vanleer(sm1, s0, sp1);
        double d2 = vanleer(s0, sp1, sp2);
        e1 = 0.5 * (sm1 + s0) - (d1 - d0) / 6.0;
        e1 = clip2(e1, sm1, s0);
        e2 = 0.5 * (s0 + sp1) - (d2 - d1) / 6.0;
        e2 = clip2(e2, s0, sp1);
    }
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

KSGF_HD double minmod_slope(double sm1, double s0, double sp1, int bclo, int bchi,
                                               int x, int dlo, int dhi)
{
    const bool lw = (bclo == KSGF_BC_EXT_DIR || bclo == KSGF_BC_HOEXTRAP);
    const bool hw = (bchi == KSGF_BC_EXT_DIR || bchi == KSGF_BC_HOEXTRAP);
    if (lw && x == dlo) return s0 - sm1;
    if (hw && x == dhi) return sp1 - s0;
    double a = s0 - sm1, b = sp1 - s0;
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}

// time-centred states at the low (Im) and high (Ip) faces of one cell;
// um/up: advecting velocity at the low/high face
KSGF_HD void ppm_trace(double sm2, double sm1, double s0, double sp1, double sp2,
                                          int lim, int bclo, int bchi, int x, int dlo, int dhi,
                                          double um, double up, double dtdx, double& Im, double& Ip)
{
    if (lim == KSGF_PPM_UPWIND) {
        Im = s0;
        Ip = s0;
        return;
    }
    if (lim == KSGF_PPM_MINMOD) {
        double sl = minmod_slope(sm1, s0, sp1, bclo, bchi, x, dlo, dhi);
        Ip = (up > kSmallVel) ? s0 + 0.5 * (1.0 - up * dtdx) * sl : s0;
        Im = (um < -kSmallVel) ? s0 - 0.5 * (1.0 + um * dtdx) * sl : s0;
        return;
    }
    double sm, sp;
    ppm_edges(sm2, sm1, s0, sp1, sp2, lim, bclo, bchi, x, dlo, dhi, sm, sp);
    double s6 = 6.0 * s0 - 3.0 * (sm + sp);
    if (up > kSmallVel) {
        double sig = fabs(up) * dtdx;
        Ip = sp - (0.5 * sig) * ((sp - sm) - (1.0 - (2.0 / 3.0) * sig) * s6);
    } else {
        Ip = s0;
    }
    if (um < -kSmallVel) {
        double sig = fabs(um) * dtdx;
        Im = sm + (0.5 * sig) * ((sp - sm) + (1.0 - (2.0 / 3.0) * sig) * s6);
    } else {
        Im = s0;
    }
}

KSGF_HD double burgers(double lo, double hi)
{
    double st = ((lo + hi) >= 0.0) ? lo : hi;
    bool ltm = ((lo <= 0.0 && hi >= 0.0) || (fabs(lo + hi) < kSmallVel));
    return ltm ? 0.0 : st;
}

KSGF_HD double upwind(double lo, double hi, double u)
{
    if (fabs(u) < kSmallVel) return 0.5 * (lo + hi);
    return (u >= 0.0) ? lo : hi;
}

KSGF_HD void edge_bcs(int x, int dlo, int dhi, int bclo, int bchi, double glo, double ghi,
                                         bool both_dir, double& lo, double& hi)
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
