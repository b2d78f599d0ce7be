// Forward-mode automatic differentiation with one tangent direction.  The
// generated right-hand sides are templates on the real type R; instantiated
// with R = cmx_dual they carry d/ds of every quantity along one seeded input
// direction s (the chain rule applied operation by operation: Griewank &
// Walther, Evaluating Derivatives, 2nd ed., SIAM 2008, ch. 3).  The adjoint
// kernels seed one cell input at a time (a state, V or a parameter) to get
// the columns of the cell step's Jacobian.  Branches and comparisons follow
// the value part, so a piecewise rate is differentiated on the branch the
// forward step took.
#pragma once
#include <cfloat>
#include <cmath>
#include <limits>

#include "model_common.h"

struct cmx_dual {
    double v;   // value
    double d;   // derivative along the seeded direction
    CMX_HD constexpr cmx_dual() : v(0.0), d(0.0) {}
    CMX_HD constexpr cmx_dual(double x) : v(x), d(0.0) {}
    CMX_HD constexpr cmx_dual(double x, double dx) : v(x), d(dx) {}
    CMX_HD cmx_dual& operator+=(const cmx_dual& b) { d += b.d; v += b.v; return *this; }
    CMX_HD cmx_dual& operator-=(const cmx_dual& b) { d -= b.d; v -= b.v; return *this; }
    CMX_HD cmx_dual& operator*=(const cmx_dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
};

CMX_HD inline cmx_dual operator+(const cmx_dual& a) { return a; }
CMX_HD inline cmx_dual operator-(const cmx_dual& a) { return cmx_dual(-a.v, -a.d); }
CMX_HD inline cmx_dual operator+(const cmx_dual& a, const cmx_dual& b) { return cmx_dual(a.v + b.v, a.d + b.d); }
CMX_HD inline cmx_dual operator-(const cmx_dual& a, const cmx_dual& b) { return cmx_dual(a.v - b.v, a.d - b.d); }
CMX_HD inline cmx_dual operator*(const cmx_dual& a, const cmx_dual& b)
{
    return cmx_dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
CMX_HD inline cmx_dual operator/(const cmx_dual& a, const cmx_dual& b)
{
    const double q = a.v / b.v;
    return cmx_dual(q, (a.d - q * b.d) / b.v);
}
CMX_HD inline bool operator<(const cmx_dual& a, const cmx_dual& b) { return a.v < b.v; }
CMX_HD inline bool operator>(const cmx_dual& a, const cmx_dual& b) { return a.v > b.v; }
CMX_HD inline bool operator<=(const cmx_dual& a, const cmx_dual& b) { return a.v <= b.v; }
CMX_HD inline bool operator>=(const cmx_dual& a, const cmx_dual& b) { return a.v >= b.v; }
CMX_HD inline bool operator==(const cmx_dual& a, const cmx_dual& b) { return a.v == b.v; }
CMX_HD inline bool operator!=(const cmx_dual& a, const cmx_dual& b) { return a.v != b.v; }

// The elementary functions the generated models call (found by argument-
// dependent lookup next to the double and float overloads).
CMX_HD inline cmx_dual cmx_exp(const cmx_dual& x)
{
    const double e = std::exp(x.v);
    return cmx_dual(e, e * x.d);
}
CMX_HD inline cmx_dual exp(const cmx_dual& x) { return cmx_exp(x); }
CMX_HD inline cmx_dual cmx_expm1(const cmx_dual& x) { return cmx_dual(std::expm1(x.v), std::exp(x.v) * x.d); }
CMX_HD inline cmx_dual cmx_log(const cmx_dual& x) { return cmx_dual(std::log(x.v), x.d / x.v); }
CMX_HD inline cmx_dual log(const cmx_dual& x) { return cmx_log(x); }
CMX_HD inline cmx_dual cmx_div(const cmx_dual& a, const cmx_dual& b) { return a / b; }
CMX_HD inline cmx_dual sqrt(const cmx_dual& x)
{
    const double r = std::sqrt(x.v);
    return cmx_dual(r, x.d / (2.0 * r));
}
CMX_HD inline cmx_dual fabs(const cmx_dual& x) { return x.v < 0.0 ? -x : x; }
CMX_HD inline cmx_dual pow(const cmx_dual& x, double e)
{
    const double p = std::pow(x.v, e);
    return cmx_dual(p, e * std::pow(x.v, e - 1.0) * x.d);
}

namespace std {
template <>
struct numeric_limits<cmx_dual> {
    static constexpr bool is_specialized = true;
    static CMX_HD constexpr double epsilon() { return DBL_EPSILON; }
};
}  // namespace std
