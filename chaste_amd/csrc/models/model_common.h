// Shared helpers for the generated cell-model headers.  The headers compile as
// HIP device/host code under hipcc and as plain C++ under g++ (the CPU oracle).
#pragma once
#include <cmath>
#include <limits>

#if defined(__HIPCC__)
#define CMX_HD __host__ __device__
#else
#define CMX_HD
#endif

#define CMX_MATH_USING \
    using std::exp; using std::log; using std::log10; using std::pow; \
    using std::sqrt; using std::floor; using std::ceil; using std::fabs;

// exp, log and division of the generated right-hand sides.  Double precision
// keeps the correctly rounded library forms; single precision (the FP32 cell
// path) uses the hardware exp2/log2/reciprocal sequences.
CMX_HD inline double cmx_exp(double x) { return std::exp(x); }
CMX_HD inline double cmx_log(double x) { return std::log(x); }
CMX_HD inline double cmx_div(double a, double b) { return a / b; }
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float cmx_exp(float x) { return __expf(x); }
__device__ inline float cmx_log(float x) { return __logf(x); }
__device__ inline float cmx_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
#else
inline float cmx_exp(float x) { return std::exp(x); }
inline float cmx_log(float x) { return std::log(x); }
inline float cmx_div(float a, float b) { return a / b; }
#endif

// exp(x) - 1 without cancellation for small x (GHK-type currents).
CMX_HD inline double cmx_expm1(double x) { return std::expm1(x); }
CMX_HD inline float cmx_expm1(float x) { return std::expm1(x); }

// f(V) near a removable singularity c (0/0 at V = c): linear interpolation
// between V = c - d and c + d inside that window, f(V) elsewhere.  d is dv,
// widened to a few units in the last place of c where R cannot resolve dv
// (single precision: +-1e-6 mV is below the float spacing near -47 mV).
template <class R, class F>
CMX_HD inline R cmx_remove_singularity(R V, R c, R dv, F f)
{
    const R ulp4 = R(4) * std::fabs(c) * std::numeric_limits<R>::epsilon();
    const R d = dv > ulp4 ? dv : ulp4;
    if (std::fabs(V - c) < d) {
        const R f0 = f(c - d), f1 = f(c + d);
        return f0 + (V - (c - d)) * (f1 - f0) / (R(2) * d);
    }
    return f(V);
}

// Integer powers as products: pow(x, 3.0) in the CellML becomes x*x*x.
template <int N, class R>
CMX_HD inline R cmx_ipow(R x)
{
    if constexpr (N == 1) return x;
    else if constexpr (N == 2) return x * x;
    else if constexpr (N == 3) return x * x * x;
    else return (x * x) * (x * x);
}
