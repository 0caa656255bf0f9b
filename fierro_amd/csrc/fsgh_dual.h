// fsgh_dual.h -- forward-mode dual numbers (value, tangent) for the tangent-
// linear model of the SGH cycle.  The physics templates (fsgh_physics.h)
// instantiated with Dual compute, alongside every value, its directional
// derivative; the value part performs exactly the double operations, so the
// trajectory is the forward one bit for bit.  Branches (min/max, clamps,
// compression test) follow the values, which is the derivative of the
// branch the forward takes.
#pragma once

#include <cmath>

#ifndef FSGH_HD
#define FSGH_HD __host__ __device__
#endif

namespace fsgh {
// The type and its functions live in their own namespace so that they are
// found only by argument-dependent lookup: plain double math keeps resolving
// to the C library functions.
namespace ad {

struct Dual {
    double v, d;
    FSGH_HD Dual() : v(0.0), d(0.0) {}
    FSGH_HD Dual(double a) : v(a), d(0.0) {}  // NOLINT: constants promote
    FSGH_HD Dual(double a, double b) : v(a), d(b) {}
    FSGH_HD Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
    FSGH_HD Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
};

FSGH_HD inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
FSGH_HD inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
FSGH_HD inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
FSGH_HD inline Dual operator*(const Dual& a, const Dual& b)
{
    return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
FSGH_HD inline Dual operator/(const Dual& a, const Dual& b)
{
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
FSGH_HD inline Dual operator+(const Dual& a, double b) { return Dual(a.v + b, a.d); }
FSGH_HD inline Dual operator+(double a, const Dual& b) { return Dual(a + b.v, b.d); }
FSGH_HD inline Dual operator-(const Dual& a, double b) { return Dual(a.v - b, a.d); }
FSGH_HD inline Dual operator-(double a, const Dual& b) { return Dual(a - b.v, -b.d); }
FSGH_HD inline Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
FSGH_HD inline Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
FSGH_HD inline Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
FSGH_HD inline Dual operator/(double a, const Dual& b)
{
    const double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}

FSGH_HD inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
FSGH_HD inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
FSGH_HD inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
FSGH_HD inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
FSGH_HD inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
FSGH_HD inline bool operator<(const Dual& a, double b) { return a.v < b; }
FSGH_HD inline bool operator>(const Dual& a, double b) { return a.v > b; }
FSGH_HD inline bool operator>=(const Dual& a, double b) { return a.v >= b; }
FSGH_HD inline bool operator<=(const Dual& a, double b) { return a.v <= b; }

FSGH_HD inline Dual sqrt(const Dual& a)
{
    const double s = ::sqrt(a.v);
    return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
FSGH_HD inline Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
// pow with a constant exponent
FSGH_HD inline Dual pow(const Dual& a, double p)
{
    const double r = ::pow(a.v, p);
    return Dual(r, a.v != 0.0 ? p * r / a.v * a.d : 0.0);
}
// min/max take the tangent of the operand whose value they return
FSGH_HD inline Dual fmin(const Dual& a, const Dual& b)
{
    const double m = ::fmin(a.v, b.v);
    return Dual(m, m == a.v ? a.d : b.d);
}
FSGH_HD inline Dual fmax(const Dual& a, const Dual& b)
{
    const double m = ::fmax(a.v, b.v);
    return Dual(m, m == a.v ? a.d : b.d);
}
FSGH_HD inline Dual fmin(const Dual& a, double b) { return fmin(a, Dual(b)); }
FSGH_HD inline Dual fmin(double a, const Dual& b) { return fmin(Dual(a), b); }
FSGH_HD inline Dual fmax(const Dual& a, double b) { return fmax(a, Dual(b)); }
FSGH_HD inline Dual fmax(double a, const Dual& b) { return fmax(Dual(a), b); }

FSGH_HD inline double value(double a) { return a; }
FSGH_HD inline double value(const Dual& a) { return a.v; }

}  // namespace ad
using ad::Dual;
}  // namespace fsgh
