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

#if defined(__clang__)
#define CMX_UNROLL _Pragma("unroll")
#else
#define CMX_UNROLL _Pragma("GCC unroll 16")
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
    using std::fabs;
    const R ulp4 = R(4) * fabs(c) * std::numeric_limits<R>::epsilon();
    const R d = dv > ulp4 ? dv : ulp4;
    if (fabs(V - c) < d) {
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

// Newton's method of the reference's backward-Euler cells
// (heart/src/odes/CardiacNewtonSolver.hpp:66-160): Jacobian solve by Gaussian
// elimination without pivoting, stop when the update's max-norm is <= 1e-6,
// and when the residual grows while the update is not yet small, walk only
// 0.2 of the way in the direction whose relative change exceeds 1.  Returns
// the iteration count, or -1 after 16 iterations (the reference throws).
template <int N, class R, class Res, class Jac>
CMX_HD inline int cmx_newton(R* g, Res residual, Jac jacobian)
{
    constexpr double eps = 1e-6;
    R r[N], J[N * N], u[N];
    residual(g, r);
    R norm_r = R(0);
CMX_UNROLL
    for (int i = 0; i < N; ++i) norm_r = std::fabs(r[i]) > norm_r ? std::fabs(r[i]) : norm_r;
    R norm_u = R(0);
    int counter = 0;
    do {
        jacobian(g, J);
CMX_UNROLL
        for (int i = 0; i < N; ++i) {
CMX_UNROLL
            for (int ii = i + 1; ii < N; ++ii) {
                const R fact = J[ii * N + i] / J[i * N + i];
CMX_UNROLL
                for (int j = i; j < N; ++j) J[ii * N + j] -= fact * J[i * N + j];
                r[ii] -= fact * r[i];
            }
        }
CMX_UNROLL
        for (int i = N - 1; i >= 0; --i) {
            R v = r[i];
CMX_UNROLL
            for (int j = i + 1; j < N; ++j) v -= J[i * N + j] * u[j];
            u[i] = v / J[i * N + i];
        }
        norm_u = R(0);
CMX_UNROLL
        for (int i = 0; i < N; ++i) norm_u = std::fabs(u[i]) > norm_u ? std::fabs(u[i]) : norm_u;
CMX_UNROLL
        for (int i = 0; i < N; ++i) g[i] -= u[i];
        const R prev = norm_r;
        residual(g, r);
        norm_r = R(0);
CMX_UNROLL
        for (int i = 0; i < N; ++i) norm_r = std::fabs(r[i]) > norm_r ? std::fabs(r[i]) : norm_r;
        if (norm_r > prev && norm_u > R(eps)) {
            R rel_max = R(0);
            int dir = 0;
CMX_UNROLL
            for (int i = 0; i < N; ++i) {
                const R rel = std::fabs(u[i] / g[i]);
                if (rel > rel_max) { rel_max = rel; dir = i; }
            }
            if (rel_max > R(1)) {
CMX_UNROLL
                for (int i = 0; i < N; ++i)
                    if (i == dir) g[i] += R(0.8) * u[i];
                residual(g, r);
                norm_r = R(0);
CMX_UNROLL
                for (int i = 0; i < N; ++i) norm_r = std::fabs(r[i]) > norm_r ? std::fabs(r[i]) : norm_r;
            }
        }
        ++counter;
        if (counter > 15) return -1;
    } while (norm_u > R(eps));
    return counter;
}
