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

// exp(x) - 1 without cancellation for small x (GHK-type currents).  Single
// precision on the device uses Kahan's form (u - 1) x / log(u), u = exp(x),
// exact to a few units in the last place with the hardware exp2/log2
// (W. Kahan, as given in N. J. Higham, Accuracy and Stability of Numerical
// Algorithms, 2nd ed., SIAM 2002, section 1.14.1).
CMX_HD inline double cmx_expm1(double x) { return std::expm1(x); }
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float cmx_expm1(float x)
{
    const float u = __expf(x);
    if (u == 1.0f) return x;
    const float um1 = u - 1.0f;
    if (um1 == -1.0f) return -1.0f;
    return um1 * x * __builtin_amdgcn_rcpf(__logf(u));
}
#else
inline float cmx_expm1(float x) { return std::expm1(x); }
#endif

// exp(E) for an exponent E = k V + c affine in the membrane voltage, with
// k2 = k log2(e) and c2 = c log2(e) folded by the generator.  Double
// precision evaluates the exponent e exactly as the model writes it; single
// precision takes exp2(V k2 + c2), one FMA and the hardware exp2, and the
// unused e is dead code.
CMX_HD inline double cmx_exp_v(double e, double V, double k2, double c2)
{
    (void)V; (void)k2; (void)c2;
    return cmx_exp(e);
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float cmx_exp_v(float e, float V, double k2, double c2)
{
    (void)e;
    return __builtin_amdgcn_exp2f(__builtin_fmaf(V, (float)k2, (float)c2));
}
#else
inline float cmx_exp_v(float e, float V, double k2, double c2)
{
    (void)e;
    return std::exp2(std::fma(V, (float)k2, (float)c2));
}
#endif

// exp(k x) for a constant k, with k2 = k log2(e) folded by the generator:
// double evaluates the exponent e as written, single precision exp2(x k2).
CMX_HD inline double cmx_exp_c(double e, double x, double k2)
{
    (void)x; (void)k2;
    return cmx_exp(e);
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float cmx_exp_c(float e, float x, double k2)
{
    (void)e;
    return __builtin_amdgcn_exp2f(x * (float)k2);
}
#else
inline float cmx_exp_c(float e, float x, double k2)
{
    (void)e;
    return std::exp2(x * (float)k2);
}
#endif

// Rush-Larsen decay factor exp(g_k h) of a gate (g_k = -1/tau).  Single
// precision folds log2(e) into the step length, which the unrolled gate loop
// computes once.
CMX_HD inline double cmx_decay(double gk, double h) { return cmx_exp(gk * h); }
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float cmx_decay(float gk, float h)
{
    return __builtin_amdgcn_exp2f(gk * (h * 1.44269504088896340736f));
}
#else
inline float cmx_decay(float gk, float h) { return std::exp(gk * h); }
#endif

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
