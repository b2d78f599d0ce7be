// Forward-mode automatic differentiation with K tangent directions.  The
// generated right-hand sides are templates on the real type R; instantiated
// with R = cmx_dualn<K> they carry d/ds_j of every quantity along K seeded
// input directions s_j at once (the chain rule applied operation by
// operation: Griewank & Walther, Evaluating Derivatives, 2nd ed., SIAM 2008,
// ch. 3, the vector forward mode of section 3.1).  One evaluation of the
// values (the exponentials, logarithms and divisions) then serves K columns
// of the Jacobian.  The adjoint kernels seed K cell inputs per pass (states,
// V or parameters).  Branches and comparisons follow the value part, so a
// piecewise rate is differentiated on the branch the forward step took.
//
// The operators and elementary functions are friends defined in the class,
// so they are found by argument-dependent lookup and a double operand
// converts to a constant.
#pragma once
#include <cfloat>
#include <cmath>
#include <limits>

#include "model_common.h"

template <int K>
struct cmx_dualn {
    double v;      // value
    double d[K];   // derivatives along the seeded directions
    CMX_HD constexpr cmx_dualn() : v(0.0), d{} {}
    CMX_HD constexpr cmx_dualn(double x) : v(x), d{} {}

    CMX_HD cmx_dualn& operator+=(const cmx_dualn& b)
    {
        for (int k = 0; k < K; ++k) d[k] += b.d[k];
        v += b.v;
        return *this;
    }
    CMX_HD cmx_dualn& operator-=(const cmx_dualn& b)
    {
        for (int k = 0; k < K; ++k) d[k] -= b.d[k];
        v -= b.v;
        return *this;
    }
    CMX_HD cmx_dualn& operator*=(const cmx_dualn& b)
    {
        for (int k = 0; k < K; ++k) d[k] = d[k] * b.v + v * b.d[k];
        v *= b.v;
        return *this;
    }

    // value x with derivative f' x.d: the chain rule for every unary function
    CMX_HD static cmx_dualn chain(const cmx_dualn& x, double fx, double dfx)
    {
        cmx_dualn r(fx);
        for (int k = 0; k < K; ++k) r.d[k] = dfx * x.d[k];
        return r;
    }

    friend CMX_HD cmx_dualn operator+(const cmx_dualn& a) { return a; }
    friend CMX_HD cmx_dualn operator-(const cmx_dualn& a) { return chain(a, -a.v, -1.0); }
    friend CMX_HD cmx_dualn operator+(const cmx_dualn& a, const cmx_dualn& b)
    {
        cmx_dualn r(a.v + b.v);
        for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
        return r;
    }
    friend CMX_HD cmx_dualn operator-(const cmx_dualn& a, const cmx_dualn& b)
    {
        cmx_dualn r(a.v - b.v);
        for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
        return r;
    }
    friend CMX_HD cmx_dualn operator*(const cmx_dualn& a, const cmx_dualn& b)
    {
        cmx_dualn r(a.v * b.v);
        for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
        return r;
    }
    friend CMX_HD cmx_dualn operator/(const cmx_dualn& a, const cmx_dualn& b)
    {
        const double q = a.v / b.v;
        cmx_dualn r(q);
        for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
        return r;
    }
    friend CMX_HD bool operator<(const cmx_dualn& a, const cmx_dualn& b) { return a.v < b.v; }
    friend CMX_HD bool operator>(const cmx_dualn& a, const cmx_dualn& b) { return a.v > b.v; }
    friend CMX_HD bool operator<=(const cmx_dualn& a, const cmx_dualn& b) { return a.v <= b.v; }
    friend CMX_HD bool operator>=(const cmx_dualn& a, const cmx_dualn& b) { return a.v >= b.v; }
    friend CMX_HD bool operator==(const cmx_dualn& a, const cmx_dualn& b) { return a.v == b.v; }
    friend CMX_HD bool operator!=(const cmx_dualn& a, const cmx_dualn& b) { return a.v != b.v; }

    // the elementary functions the generated models call
    friend CMX_HD cmx_dualn cmx_exp(const cmx_dualn& x)
    {
        const double e = std::exp(x.v);
        return chain(x, e, e);
    }
    friend CMX_HD cmx_dualn exp(const cmx_dualn& x) { return cmx_exp(x); }
    friend CMX_HD cmx_dualn cmx_expm1(const cmx_dualn& x) { return chain(x, std::expm1(x.v), std::exp(x.v)); }
    friend CMX_HD cmx_dualn cmx_exp_v(const cmx_dualn& e, const cmx_dualn& V, double k2, double c2)
    {
        (void)V; (void)k2; (void)c2;
        return cmx_exp(e);
    }
    friend CMX_HD cmx_dualn cmx_exp_c(const cmx_dualn& e, const cmx_dualn& x, double k2)
    {
        (void)x; (void)k2;
        return cmx_exp(e);
    }
    friend CMX_HD cmx_dualn cmx_decay(const cmx_dualn& gk, const cmx_dualn& h) { return cmx_exp(gk * h); }
    friend CMX_HD cmx_dualn cmx_log(const cmx_dualn& x)
    {
        cmx_dualn r(std::log(x.v));
        for (int k = 0; k < K; ++k) r.d[k] = x.d[k] / x.v;
        return r;
    }
    friend CMX_HD cmx_dualn log(const cmx_dualn& x) { return cmx_log(x); }
    friend CMX_HD cmx_dualn cmx_div(const cmx_dualn& a, const cmx_dualn& b) { return a / b; }
    friend CMX_HD cmx_dualn sqrt(const cmx_dualn& x)
    {
        const double q = std::sqrt(x.v);
        cmx_dualn r(q);
        for (int k = 0; k < K; ++k) r.d[k] = x.d[k] / (2.0 * q);
        return r;
    }
    friend CMX_HD cmx_dualn fabs(const cmx_dualn& x) { return x.v < 0.0 ? -x : x; }
    friend CMX_HD cmx_dualn pow(const cmx_dualn& x, double e)
    {
        return chain(x, std::pow(x.v, e), e * std::pow(x.v, e - 1.0));
    }
    // x^y with both varying: y x^(y-1) dx + x^y ln(x) dy (the second term
    // only where x > 0, where x^y with a varying y is defined)
    friend CMX_HD cmx_dualn pow(const cmx_dualn& x, const cmx_dualn& y)
    {
        const double v = std::pow(x.v, y.v);
        const double dx = y.v * std::pow(x.v, y.v - 1.0);
        const double dy = x.v > 0.0 ? v * std::log(x.v) : 0.0;
        cmx_dualn r(v);
        for (int k = 0; k < K; ++k) r.d[k] = dx * x.d[k] + dy * y.d[k];
        return r;
    }
};

using cmx_dual = cmx_dualn<1>;

namespace std {
template <int K>
struct numeric_limits<cmx_dualn<K>> {
    static constexpr bool is_specialized = true;
    static CMX_HD constexpr double epsilon() { return DBL_EPSILON; }
};
}  // namespace std
