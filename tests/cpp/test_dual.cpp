// Host check of the vector forward mode (models/model_dual.h) on the
// generated cell models: the gradient of the total ionic current with
// respect to every state, V and parameter, taken four columns per pass, is
// bitwise the one-column-per-pass result and agrees with central differences
// of the double-precision model to 1e-3 relative (the differences' own error
// at a step of 1e-6 of each input reaches ~1e-4 on the small concentrations;
// a wrong chain rule shows as an O(1) error).  Prints one line per case and
// exits 1 on a mismatch.
#include <cmath>
#include <cstdio>
#include <initializer_list>

#include "models/lr91.h"
#include "models/model_dual.h"
#include "models/tt06_epi.h"

template <int K, int NS, int NP, class F>
void columns(F f, const double* y0, double V0, const double* p0, double* out)
{
    using D = cmx_dualn<K>;
    for (int k0 = 0; k0 < NS + NP; k0 += K) {
        D y[NS], p[NP];
        D V(V0);
        for (int j = 0; j < K; ++j) V.d[j] = k0 + j == 0 ? 1.0 : 0.0;
        for (int s = 1; s < NS; ++s) {
            y[s] = D(y0[s]);
            for (int j = 0; j < K; ++j) y[s].d[j] = s == k0 + j ? 1.0 : 0.0;
        }
        for (int q = 0; q < NP; ++q) {
            p[q] = D(p0[q]);
            for (int j = 0; j < K; ++j) p[q].d[j] = NS + q == k0 + j ? 1.0 : 0.0;
        }
        const D I = f(y, V, p);
        for (int j = 0; j < K && k0 + j < NS + NP; ++j) out[k0 + j] = I.d[j];
    }
}

template <int NS, int NP, class F4, class F1, class FD>
int check(const char* name, F4 f4, F1 f1, FD fd, const double* y0, double V0, const double* p0)
{
    double g4[NS + NP], g1[NS + NP];
    columns<4, NS, NP>(f4, y0, V0, p0, g4);
    columns<1, NS, NP>(f1, y0, V0, p0, g1);
    int bad = 0;
    double worst = 0.0;
    int worst_k = 0;
    for (int k = 0; k < NS + NP; ++k) {
        if (g4[k] != g1[k]) ++bad;
        // central difference of the double model along input k
        double y[NS], p[NP], V = V0;
        for (int s = 0; s < NS; ++s) y[s] = y0[s];
        for (int q = 0; q < NP; ++q) p[q] = p0[q];
        double* x = k == 0 ? &V : (k < NS ? &y[k] : &p[k - NS]);
        const double x0 = *x, h = 1e-6 * (std::fabs(x0) + 1e-3);
        *x = x0 + h;
        const double up = fd(y, V, p);
        *x = x0 - h;
        const double dn = fd(y, V, p);
        const double c = (up - dn) / (2.0 * h);
        // relative to the derivative, above the difference quotient's rounding floor
        const double floor = 1e-13 * (std::fabs(up) + 1.0) / h;
        const double err = std::fabs(c - g4[k]) / (std::fabs(c) + floor);
        if (err > worst) {
            worst = err;
            worst_k = k;
        }
    }
    std::printf("%s at V = %g: %d columns, %d differ between 4 and 1 tangents, worst central-difference "
                "error %.2e (input %d)\n", name, V0, NS + NP, bad, worst, worst_k);
    return bad == 0 && worst < 1e-3 ? 0 : 1;
}

int main()
{
    int rc = 0;
    {
        namespace m = cmx::lr91;
        constexpr int NS = m::kNumStates, NP = m::kNumParams;
        double y[NS], p[NP];
        for (int s = 0; s < NS; ++s) y[s] = m::kInitialState[s];
        for (int q = 0; q < NP; ++q) p[q] = m::kParamDefaults[q];
        for (const double V : {-84.0, -30.0, 20.0}) {
            auto f4 = [](const cmx_dualn<4>* y, cmx_dualn<4> V, const cmx_dualn<4>* p) {
                return m::iionic<cmx_dualn<4>>(y, V, p, cmx_dualn<4>(1.0));
            };
            auto f1 = [](const cmx_dual* y, cmx_dual V, const cmx_dual* p) {
                return m::iionic<cmx_dual>(y, V, p, cmx_dual(1.0));
            };
            auto fd = [](const double* y, double V, const double* p) { return m::iionic<double>(y, V, p, 1.0); };
            rc |= check<NS, NP>("LR91 Iion", f4, f1, fd, y, V, p);
        }
    }
    {
        namespace m = cmx::tt06_epi;
        constexpr int NS = m::kNumStates, NP = m::kNumParams;
        double y[NS], p[NP];
        for (int s = 0; s < NS; ++s) y[s] = m::kInitialState[s];
        for (int q = 0; q < NP; ++q) p[q] = m::kParamDefaults[q];
        for (const double V : {-86.0, -30.0, 20.0}) {
            auto f4 = [](const cmx_dualn<4>* y, cmx_dualn<4> V, const cmx_dualn<4>* p) {
                return m::iionic<cmx_dualn<4>>(y, V, p, cmx_dualn<4>(1.0));
            };
            auto f1 = [](const cmx_dual* y, cmx_dual V, const cmx_dual* p) {
                return m::iionic<cmx_dual>(y, V, p, cmx_dual(1.0));
            };
            auto fd = [](const double* y, double V, const double* p) { return m::iionic<double>(y, V, p, 1.0); };
            rc |= check<NS, NP>("TT06 Iion", f4, f1, fd, y, V, p);
        }
    }
    return rc;
}
