// Forward-mode dual numbers with N partial derivatives, for the local
// Jacobians of the Godunov reconstruction in the adjoint (vector-Jacobian
// products of one face at a time).  The value part is computed with exactly
// the operations and operand order of plain double code, so every branch a
// limiter or upwind selector takes on the dual value is the branch the
// forward pass took (branch decisions are frozen at the forward state); the
// partials are the derivative of that branch.
//
// vabs / vmin / vmax / vcopysign are the math functions the reconstruction
// uses, overloaded for double (the library functions) and for Dual.  At a
// tie, vmin / vmax follow the first argument (as fmin / fmax return it).
#pragma once

#include "ksgf_hd.h"

namespace ksgf {

KSGF_HD double vabs(double x) { return fabs(x); }
KSGF_HD double vmin(double a, double b) { return fmin(a, b); }
KSGF_HD double vmax(double a, double b) { return fmax(a, b); }
KSGF_HD double vcopysign(double m, double s) { return copysign(m, s); }
KSGF_HD double value_of(double x) { return x; }
KSGF_HD double zero_like(double) { return 0.0; }

template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
KSGF_HD Dual<N> dual_const(double v)
{
    Dual<N> r;
    r.v = v;
    for (int i = 0; i < N; ++i) r.d[i] = 0.0;
    return r;
}

// independent variable number `slot` with value v
template <int N>
KSGF_HD Dual<N> dual_seed(double v, int slot)
{
    Dual<N> r = dual_const<N>(v);
    r.d[slot] = 1.0;
    return r;
}

template <int N>
KSGF_HD Dual<N> zero_like(const Dual<N>&)
{
    return dual_const<N>(0.0);
}

template <int N>
KSGF_HD double value_of(const Dual<N>& x)
{
    return x.v;
}

template <int N>
KSGF_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
KSGF_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
KSGF_HD Dual<N> operator-(const Dual<N>& a)
{
    Dual<N> r;
    r.v = -a.v;
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N>
KSGF_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N>
KSGF_HD Dual<N> operator*(double a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
    return r;
}
template <int N>
KSGF_HD Dual<N> operator*(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v * b;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N>
KSGF_HD Dual<N> operator/(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v / b;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b;
    return r;
}
template <int N>
KSGF_HD Dual<N> operator+(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N>
KSGF_HD Dual<N> operator+(double a, const Dual<N>& b)
{
    Dual<N> r = b;
    r.v = a + b.v;
    return r;
}
template <int N>
KSGF_HD Dual<N> operator-(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N>
KSGF_HD Dual<N> operator-(double a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
    return r;
}
template <int N>
KSGF_HD Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b)
{
    a = a + b;
    return a;
}

// comparisons act on the values (frozen branches)
template <int N>
KSGF_HD bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N>
KSGF_HD bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N>
KSGF_HD bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }
template <int N>
KSGF_HD bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }
template <int N>
KSGF_HD bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
template <int N>
KSGF_HD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }

template <int N>
KSGF_HD Dual<N> vabs(const Dual<N>& x)
{
    Dual<N> r = (x.v < 0.0) ? -x : x;
    r.v = fabs(x.v);
    return r;
}
template <int N>
KSGF_HD Dual<N> vmin(const Dual<N>& a, const Dual<N>& b)
{
    return (b.v < a.v) ? b : a;
}
template <int N>
KSGF_HD Dual<N> vmax(const Dual<N>& a, const Dual<N>& b)
{
    return (b.v > a.v) ? b : a;
}
// |m| with the sign of s (m and s as in copysign)
template <int N>
KSGF_HD Dual<N> vcopysign(const Dual<N>& m, const Dual<N>& s)
{
    Dual<N> r;
    r.v = copysign(m.v, s.v);
    const double f = (copysign(1.0, m.v) != copysign(1.0, s.v)) ? -1.0 : 1.0;
    for (int i = 0; i < N; ++i) r.d[i] = f * m.d[i];
    return r;
}

}  // namespace ksgf
