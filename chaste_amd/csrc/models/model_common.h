// Shared helpers for the generated cell-model headers.  The headers compile as
// HIP device/host code under hipcc and as plain C++ under g++ (the CPU oracle).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define CMX_HD __host__ __device__
#else
#define CMX_HD
#endif

#define CMX_MATH_USING \
    using std::exp; using std::log; using std::log10; using std::pow; \
    using std::sqrt; using std::floor; using std::ceil; using std::fabs;

// Integer powers as products: pow(x, 3.0) in the CellML becomes x*x*x.
template <int N, class R>
CMX_HD inline R cmx_ipow(R x)
{
    if constexpr (N == 1) return x;
    else if constexpr (N == 2) return x * x;
    else if constexpr (N == 3) return x * x * x;
    else return (x * x) * (x * x);
}
