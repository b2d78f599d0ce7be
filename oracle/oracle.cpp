// FP64 CPU oracle -- test infrastructure only (see oracle.h for the reference
// files each part restates).  Built by oracle/Makefile into oracle/liboracle.so.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../chaste_amd/csrc/models/dn85.h"
#include "../chaste_amd/csrc/models/fox2002.h"
#include "../chaste_amd/csrc/models/fr2000.h"
#include "../chaste_amd/csrc/models/hh52.h"
#include "../chaste_amd/csrc/models/lr91.h"
#include "../chaste_amd/csrc/models/mahajan2008.h"
#include "../chaste_amd/csrc/models/maleckar2008.h"
#include "../chaste_amd/csrc/models/n98.h"
#include "../chaste_amd/csrc/models/ord_endo.h"
#include "../chaste_amd/csrc/models/tt06_epi.h"

namespace {

// ---------------------------------------------------------------------------
// LR91, restated from heart/dynamic/DynamicallyLoadableLr91.cpp:45-280.  The
// reference keeps the state as (h, j, m, Cai, V, d, f, X); this restatement
// works on the CellML/codegen order (V, m, h, j, d, f, X, Cai) used everywhere
// else here, so only the indexing differs.
// ---------------------------------------------------------------------------
struct Lr91Hand {
    static constexpr double kC = 1.0, kF = 96484.6, kR = 8314.0, kT = 310.0;
    static constexpr double kEb = -59.87, kGb = 0.03921, kGNa = 23.0;
    static constexpr double kKi = 145.0, kKo = 5.4, kNai = 18.0, kNao = 140.0;
    static constexpr double kGKp = 0.0183, kPRNaK = 0.01833;

    // Ionic currents at state y (V = y[0]); returns the total, fills Cai' and
    // the gate rates when rates != nullptr.
    static double currents(const double* y, double* dy)
    {
        const double V = y[0], m = y[1], h = y[2], j = y[3], d = y[4], f = y[5], X = y[6], Cai = y[7];
        const double RTonF = kR * kT / kF;
        const double ENa = RTonF * std::log(kNao / kNai);
        const double iNa = kGNa * std::pow(m, 3.0) * h * j * (V - ENa);
        const double Esi = 7.7 - 13.0287 * std::log(Cai);
        const double isi = 0.09 * d * f * (V - Esi);
        const double gK = 0.282 * std::sqrt(kKo / 5.4);
        const double Xi = (V > -100.0)
            ? 2.837 * (std::exp(0.04 * (V + 77.0)) - 1.0) / ((V + 77.0) * std::exp(0.04 * (V + 35.0)))
            : 1.0;
        const double EK = RTonF * std::log((kKo + kPRNaK * kNao) / (kKi + kPRNaK * kNai));
        const double iK = gK * X * Xi * (V - EK);
        const double gK1 = 0.6047 * std::sqrt(kKo / 5.4);
        const double EK1 = RTonF * std::log(kKo / kKi);
        const double aK1 = 1.02 / (1.0 + std::exp(0.2385 * (V - EK1 - 59.215)));
        const double bK1 = (0.49124 * std::exp(0.08032 * (V + 5.476 - EK1))
                            + std::exp(0.06175 * (V - (EK1 + 594.31))))
                           / (1.0 + std::exp(-0.5143 * (V - EK1 + 4.753)));
        const double iK1 = gK1 * (aK1 / (aK1 + bK1)) * (V - EK1);
        const double Kp = 1.0 / (1.0 + std::exp((7.488 - V) / 5.98));
        const double iKp = kGKp * Kp * (V - EK1);
        const double ib = kGb * (V - kEb);
        if (dy) {
            const bool low = V < -40.0;
            const double ah = low ? 0.135 * std::exp((80.0 + V) / -6.8) : 0.0;
            const double bh = low ? 3.56 * std::exp(0.079 * V) + 3.1e5 * std::exp(0.35 * V)
                                  : 1.0 / (0.13 * (1.0 + std::exp((V + 10.66) / -11.1)));
            const double aj = low ? (-1.2714e5 * std::exp(0.2444 * V) - 3.474e-5 * std::exp(-0.04391 * V))
                                        * (V + 37.78) / (1.0 + std::exp(0.311 * (V + 79.23)))
                                  : 0.0;
            const double bj = low ? 0.1212 * std::exp(-0.01052 * V) / (1.0 + std::exp(-0.1378 * (V + 40.14)))
                                  : 0.3 * std::exp(-2.535e-7 * V) / (1.0 + std::exp(-0.1 * (V + 32.0)));
            const double am = 0.32 * (V + 47.13) / (1.0 - std::exp(-0.1 * (V + 47.13)));
            const double bm = 0.08 * std::exp(-V / 11.0);
            const double ad = 0.095 * std::exp(-0.01 * (V - 5.0)) / (1.0 + std::exp(-0.072 * (V - 5.0)));
            const double bd = 0.07 * std::exp(-0.017 * (V + 44.0)) / (1.0 + std::exp(0.05 * (V + 44.0)));
            const double af = 0.012 * std::exp(-0.008 * (V + 28.0)) / (1.0 + std::exp(0.15 * (V + 28.0)));
            const double bf = 0.0065 * std::exp(-0.02 * (V + 30.0)) / (1.0 + std::exp(-0.2 * (V + 30.0)));
            const double aX = 0.0005 * std::exp(0.083 * (V + 50.0)) / (1.0 + std::exp(0.057 * (V + 50.0)));
            const double bX = 0.0013 * std::exp(-0.06 * (V + 20.0)) / (1.0 + std::exp(-0.04 * (V + 20.0)));
            dy[1] = am * (1.0 - m) - bm * m;
            dy[2] = ah * (1.0 - h) - bh * h;
            dy[3] = aj * (1.0 - j) - bj * j;
            dy[4] = ad * (1.0 - d) - bd * d;
            dy[5] = af * (1.0 - f) - bf * f;
            dy[6] = aX * (1.0 - X) - bX * X;
            dy[7] = -1e-4 * isi + 0.07 * (1e-4 - Cai);
        }
        return iNa + isi + iK + iK1 + iKp + ib;
    }
};

// Stimulus functions (heart/src/stimulus/SimpleStimulus.cpp:44-76, RegularStimulus.cpp:58-80).
struct Stimulus {
    double mag = 0, dur = 0, start = 0, period = -1, stop = DBL_MAX;
    Stimulus() = default;
    Stimulus(double m, double d, double s, double p) : mag(m), dur(d), start(s), period(p)
    {
        // SimpleStimulus pads the duration by (duration + start) * eps;
        // RegularStimulus by period * eps.
        if (period > 0) dur += period * DBL_EPSILON;
        else dur += (dur + start) * DBL_EPSILON;
    }
    double at(double t) const
    {
        if (mag == 0.0) return 0.0;
        if (period > 0) {
            if (t < start) return 0.0;
            const double beat = std::fmod(t - start, period);
            return (beat >= 0 && beat <= dur && t <= stop) ? mag : 0.0;
        }
        return (start <= t && t <= dur + start) ? mag : 0.0;
    }
};

// A model generated from CellML (or written to the generator's interface):
// its state and parameter tables and each integrator's substep with V
// clamped unless update_v (the reference's ComputeExceptVoltage / Compute).
// Forward Euler and Rush-Larsen: AbstractCardiacCell / AbstractRushLarsenCardiacCell
// (heart/src/odes/AbstractRushLarsenCardiacCell.cpp:57-153); backward Euler
// advances V first and then steps the other states at the new V
// (AbstractBackwardEulerCardiacCell.hpp:300-330).
struct ModelEntry {
    int S, P;
    const double* init;
    const double* pdef;
    bool implicit;           // backward-Euler and generalized Rush-Larsen forms
    double (*substep)(int integ, const double* p, double cm, double* y, double dt, double stim_area, bool update_v);
    double (*iion)(const double* p, double cm, const double* y);
};

// Newton failures of backward-Euler substeps (the reference throws).
std::atomic<int> g_newton_failures{0};

// The backward-Euler and generalized Rush-Larsen substeps of a generated model.
#define ORC_IMPLICIT_STEPS(ns)                                                                                  \
    if (integ == ORC_GRL1 || integ == ORC_GRL2) {                                                               \
        if (integ == ORC_GRL1) ns::grl_step<double, 1>(y, stim_area, p, cm, dt, update_v);                      \
        else ns::grl_step<double, 2>(y, stim_area, p, cm, dt, update_v);                                        \
        return 0.0;                                                                                             \
    }                                                                                                           \
    if (integ == ORC_BE) {                                                                                      \
        if (update_v) y[0] += dt * (dv = ns::dvdt<double>(y, y[0], stim_area, p, cm));                          \
        if (ns::be_step<double>(y, y[0], stim_area, p, cm, dt) < 0) ++g_newton_failures;                       \
        return dv;                                                                                              \
    }

#define ORC_MODEL_FNS(ns, IMPLICIT_STEPS)                                                                       \
    [](int integ, const double* p, double cm, double* y, double dt, double stim_area, bool update_v) -> double { \
        double dy[64] = {}, ginf[64], gk[64], dv = 0.0;                                                         \
        IMPLICIT_STEPS                                                                                          \
        const bool rl = integ == ORC_RL;                                                                        \
        if (update_v) dv = ns::dvdt<double>(y, y[0], stim_area, p, cm);                                        \
        if (rl) ns::evaluate<double, true>(y, y[0], stim_area, p, cm, dy, ginf, gk);                            \
        else ns::evaluate<double, false>(y, y[0], stim_area, p, cm, dy, ginf, gk);                              \
        for (int i = 1; i < ns::kNumStates; ++i) {                                                              \
            if (rl && ns::kIsGate[i]) y[i] = ginf[i] + (y[i] - ginf[i]) * std::exp(gk[i] * dt);                \
            else y[i] += dt * dy[i];                                                                            \
        }                                                                                                       \
        if (update_v) y[0] += dt * dv;                                                                          \
        return dv;                                                                                              \
    },                                                                                                          \
        [](const double* p, double cm, const double* y) { return ns::iionic<double>(y, y[0], p, cm); }

// A generated model (all integrators) and one with forward-Euler and
// Rush-Larsen forms only.
#define ORC_MODEL(ns) \
    ModelEntry { ns::kNumStates, ns::kNumParams, ns::kInitialState, ns::kParamDefaults, true, ORC_MODEL_FNS(ns, ORC_IMPLICIT_STEPS(ns)) }
#define ORC_MODEL_EXPLICIT(ns) \
    ModelEntry { ns::kNumStates, ns::kNumParams, ns::kInitialState, ns::kParamDefaults, false, ORC_MODEL_FNS(ns, ) }

const ModelEntry* model_entry(int model)
{
    static const ModelEntry table[] = {
        ORC_MODEL(cmx::lr91),                // ORC_LR91_HAND: table values; the steps are Lr91Hand's
        ORC_MODEL(cmx::lr91),                // ORC_LR91_CELLML
        ORC_MODEL(cmx::tt06_epi),            // ORC_TT06_EPI
        ORC_MODEL_EXPLICIT(cmx::ord_endo),   // ORC_ORD_ENDO
        ORC_MODEL(cmx::hh52),                // ORC_HH52
        ORC_MODEL(cmx::fox2002),             // ORC_FOX2002
        ORC_MODEL(cmx::mahajan2008),         // ORC_MAHAJAN2008
        ORC_MODEL(cmx::dn85),                // ORC_DN85
        ORC_MODEL(cmx::fr2000),              // ORC_FR2000
        ORC_MODEL(cmx::maleckar2008),        // ORC_MALECKAR2008
        ORC_MODEL(cmx::n98),                 // ORC_N98
    };
    if (model < 0 || model >= (int)(sizeof(table) / sizeof(table[0]))) return nullptr;
    return &table[model];
}

int num_states(int model)
{
    const ModelEntry* m = model_entry(model);
    return m ? m->S : -1;
}

const double* initial_state(int model) { return model_entry(model)->init; }

// The O'Hara-Rudy cell has forward-Euler and Rush-Larsen forms only, the hand
// LR91 restatement forward Euler only.
bool integrator_ok(int model, int integ)
{
    if (integ < ORC_FE || integ > ORC_GRL2 || !model_entry(model)) return false;
    if (model == ORC_LR91_HAND) return integ == ORC_FE;
    return integ == ORC_FE || integ == ORC_RL || model_entry(model)->implicit;
}

// One cell substep over [t, t+dt] with V clamped unless update_v.  stim_area
// is the area stimulus (uA/cm^2) at time t.  Returns dV/dt when update_v.
double cell_substep(int model, int integ, const double* p, double cm, double* y, double t, double dt,
                    double stim_area, bool update_v)
{
    (void)t;
    if (model == ORC_LR91_HAND) {
        double dy[64] = {};
        const double iion = Lr91Hand::currents(y, dy);
        const double dv = -(iion + stim_area) / Lr91Hand::kC;
        // The hand restatement only has the forward-Euler form (the
        // reference class is an AbstractCardiacCell).
        if (update_v) y[0] += dt * dv;
        for (int i = 1; i < num_states(model); ++i) y[i] += dt * dy[i];
        return dv;
    }
    return model_entry(model)->substep(integ, p, cm, y, dt, stim_area, update_v);
}

double cell_iionic(int model, const double* p, double cm, const double* y)
{
    if (model == ORC_LR91_HAND) return Lr91Hand::currents(y, nullptr);
    return model_entry(model)->iion(p, cm, y);
}

const double* default_params(int model) { return model_entry(model)->pdef; }

int num_params(int model) { return model_entry(model)->P; }

// TimeStepper (global/src/TimeStepper.cpp:43-200): t_k = start + k*dt, the last
// step snapped to the end when within eps.
struct Stepper {
    double start, end, dt, eps, time, next;
    unsigned long long k = 0;
    Stepper(double s, double e, double d) : start(s), end(e), dt(d), time(s)
    {
        eps = end > 1.0 ? DBL_EPSILON * end : DBL_EPSILON;
        next = calc();
    }
    double calc() const
    {
        double n = start + (k + 1) * dt;
        if (end - n <= eps) n = end;
        return n;
    }
    bool at_end() const { return time >= end; }
    double step_size() const { return next >= end ? end - time : dt; }
    void advance() { ++k; time = next; next = calc(); }
};

// Second-order Gaussian rules on the reference simplex (GaussianQuadratureRule
// with order 2, pde/src/solver/GaussianQuadratureRule.cpp:81-91, 126-135,
// 184-202): 2 points in 1-D, 3 in 2-D, 4 in 3-D.
void quad_rule(int dim, std::vector<std::vector<double>>& qp, std::vector<double>& qw)
{
    if (dim == 1) {
        const double s = std::sqrt(1.0 / 3.0);
        qw = {0.5, 0.5};
        qp = {{(-s + 1.0) / 2.0}, {(s + 1.0) / 2.0}};
    } else if (dim == 2) {
        qw = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
        qp = {{2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}};
    } else {
        const double s = std::sqrt(0.2);
        const double a = (1.0 + 3.0 * s) / 4.0, b = (1.0 - s) / 4.0;
        qw = {1.0 / 24, 1.0 / 24, 1.0 / 24, 1.0 / 24};
        qp = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
    }
}

// |det J| of a simplex (J[r][c] = x_{c+1,r} - x_{0,r}).
double simplex_det(int dim, const double* coords, const int* nd)
{
    double J[3][3] = {};
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) J[r][c] = coords[nd[c + 1] * dim + r] - coords[nd[0] * dim + r];
    double det;
    if (dim == 1) det = J[0][0];
    else if (dim == 2) det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    else
        det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    return std::fabs(det);
}

// State-variable interpolation correction of the right-hand side
// (AbstractCorrectionTermAssembler.cpp:98-156, MonodomainCorrectionTermAssembler.cpp:
// 48-74): on every element whose nodal ionic currents differ pairwise by more
// than 1 uA/cm^2, each quadrature point adds
//   w |J| phi_a (-Am) (Iion(sum_j phi_j y_j) - sum_j phi_j Iion_j)
// to node a; the interpolated state includes V and takes the cell parameters
// of the element's first node.  Y: state-major planes (plane 0 unused), V.
void svi_correction(int dim, int n, const double* coords, int ne, const int* elems, int model, double am, double cm,
                    const double* Y, const double* V, const double* iion, const double* params, int n_param,
                    double* out)
{
    const int nb = dim + 1;
    const int S = num_states(model), P = num_params(model);
    const double* pdef = default_params(model);
    std::vector<std::vector<double>> qp;
    std::vector<double> qw;
    quad_rule(dim, qp, qw);
    std::fill(out, out + n, 0.0);
    for (int e = 0; e < ne; ++e) {
        const int* nd = elems + (size_t)e * nb;
        double di = 0;
        for (int a = 0; a < nb; ++a)
            for (int b = a + 1; b < nb; ++b) di = std::max(di, std::fabs(iion[nd[a]] - iion[nd[b]]));
        if (!(di > 1.0)) continue;
        const double adet = simplex_det(dim, coords, nd);
        double p[64];
        for (int q = 0; q < P; ++q) p[q] = (params && q < n_param) ? params[(size_t)q * n + nd[0]] : pdef[q];
        double be[4] = {};
        for (size_t q = 0; q < qw.size(); ++q) {
            double phi[4], sum = 0;
            for (int c = 0; c < dim; ++c) { phi[c + 1] = qp[q][c]; sum += qp[q][c]; }
            phi[0] = 1.0 - sum;
            double yq[64] = {}, ii = 0;
            for (int a = 0; a < nb; ++a) {
                ii += phi[a] * iion[nd[a]];
                yq[0] += phi[a] * V[nd[a]];
                for (int s2 = 1; s2 < S; ++s2) yq[s2] += phi[a] * Y[(size_t)s2 * n + nd[a]];
            }
            const double isv = cell_iionic(model, p, cm, yq);
            const double wJ = qw[q] * adet;
            for (int a = 0; a < nb; ++a) be[a] += phi[a] * (-am) * (isv - ii) * wJ;
        }
        for (int a = 0; a < nb; ++a) out[nd[a]] += be[a];
    }
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int orc_model_num_states(int model) { return num_states(model); }

int orc_cell_compute(int model, int integrator, const double* params, double* y,
                     double t0, double t_end, double dt,
                     double stim_mag, double stim_dur, double stim_start, double stim_period,
                     int steps_per_sample, double* samples, int max_samples)
{
    const int S = num_states(model);
    if (S < 0 || dt <= 0 || t_end <= t0 || steps_per_sample <= 0) return -1;
    if (!integrator_ok(model, integrator)) return -1;
    const int fails0 = g_newton_failures.load();
    const double* p = params ? params : default_params(model);
    Stimulus stim(stim_mag, stim_dur, stim_start, stim_period);
    int ns = 0;
    auto record = [&]() {
        if (samples && ns < max_samples) std::memcpy(samples + (size_t)ns * S, y, sizeof(double) * S);
        ++ns;
    };
    record();
    // AbstractCardiacCell::Compute with tSamp = dt: an outer TimeStepper at dt,
    // each sample interval integrated by one inner step.  Rush-Larsen and
    // backward-Euler cells use curr_time = tStart + i*dt
    // (AbstractRushLarsenCardiacCell.cpp:89-103,
    // AbstractBackwardEulerCardiacCell.hpp:146-200).
    const bool fixed_dt = integrator != ORC_FE;
    Stepper st(t0, t_end, dt);
    unsigned long long count = 0;
    while (!st.at_end()) {
        const double t = fixed_dt ? t0 + st.k * dt : st.time;
        const double h = fixed_dt ? dt : st.step_size();
        cell_substep(model, integrator, p, 1.0, y, t, h, stim.at(t), true);
        st.advance();
        ++count;
        if (count % steps_per_sample == 0) record();
    }
    if (g_newton_failures.load() != fails0) return -1;
    return ns;
}

int orc_cell_compute_v(int model, int integrator, const double* params, double* y, double t_end, double dt,
                       double stim_mag, double stim_dur, double stim_start, double stim_period, double* v_out)
{
    // orc_cell_compute from t = 0 in steps of dt, keeping only V (v_out[0] at
    // t = 0, then after every step)
    const int S = num_states(model);
    if (S < 0 || dt <= 0 || t_end <= 0) return -1;
    if (!integrator_ok(model, integrator)) return -1;
    const int fails0 = g_newton_failures.load();
    const double* p = params ? params : default_params(model);
    Stimulus stim(stim_mag, stim_dur, stim_start, stim_period);
    const bool fixed_dt = integrator != ORC_FE;
    Stepper st(0.0, t_end, dt);
    int ns = 0;
    v_out[ns++] = y[0];
    while (!st.at_end()) {
        const double t = fixed_dt ? st.k * dt : st.time;
        cell_substep(model, integrator, p, 1.0, y, t, fixed_dt ? dt : st.step_size(), stim.at(t), true);
        st.advance();
        v_out[ns++] = y[0];
    }
    if (g_newton_failures.load() != fails0) return -1;
    return ns;
}

double orc_cell_compute_except_voltage(int model, int integrator, const double* params, double* y,
                                       double t0, double t_end, double dt)
{
    // AbstractCardiacCell::ComputeExceptVoltage (AbstractCardiacCell.cpp:90-105):
    // the same stepping as Compute with V clamped (so the stimulus, which
    // only enters dV/dt, has no effect), then GetIIonic at the result.
    const int S = num_states(model);
    if (S < 0 || dt <= 0 || t_end <= t0) return NAN;
    if (!integrator_ok(model, integrator)) return NAN;
    const int fails0 = g_newton_failures.load();
    const double* p = params ? params : default_params(model);
    const bool fixed_dt = integrator != ORC_FE;
    Stepper st(t0, t_end, dt);
    while (!st.at_end()) {
        const double t = fixed_dt ? t0 + st.k * dt : st.time;
        cell_substep(model, integrator, p, 1.0, y, t, fixed_dt ? dt : st.step_size(), 0.0, false);
        st.advance();
    }
    if (g_newton_failures.load() != fails0) return NAN;
    return cell_iionic(model, p, 1.0, y);
}

double orc_cell_iionic(int model, const double* params, const double* y)
{
    return cell_iionic(model, params ? params : default_params(model), 1.0, y);
}

int orc_slab_sizes(int dim, int ex, int ey, int ez, int* n_nodes, int* n_elems)
{
    if (dim == 1) { *n_nodes = ex + 1; *n_elems = ex; return 0; }
    if (dim == 2) { *n_nodes = (ex + 1) * (ey + 1); *n_elems = 2 * ex * ey; return 0; }
    if (dim == 3) { *n_nodes = (ex + 1) * (ey + 1) * (ez + 1); *n_elems = 6 * ex * ey * ez; return 0; }
    return -1;
}

// ConstructLinearMesh / ConstructRectangularMesh (stagger) / ConstructCuboid,
// then Scale(h): mesh/src/common/AbstractTetrahedralMesh.cpp:277-614.
int orc_slab_fill(int dim, int ex, int ey, int ez, double h, double* coords, int* elems)
{
    if (dim == 1) {
        for (int i = 0; i <= ex; ++i) coords[i] = i * h;
        for (int e = 0; e < ex; ++e) { elems[2 * e] = e; elems[2 * e + 1] = e + 1; }
        return 0;
    }
    if (dim == 2) {
        const int W = ex, H = ey;
        for (int j = 0; j <= H; ++j)
            for (int i = 0; i <= W; ++i) {
                const int n = j * (W + 1) + i;
                coords[2 * n] = i * h;
                coords[2 * n + 1] = j * h;
            }
        int e = 0;
        for (int j = 0; j < H; ++j)
            for (int i = 0; i < W; ++i) {
                const int parity = (i + (H - j)) % 2;
                const int nw = (j + 1) * (W + 1) + i, sw = j * (W + 1) + i;
                int* up = elems + 3 * e++;
                up[0] = nw; up[1] = nw + 1; up[2] = parity == 1 ? sw + 1 : sw;
                int* lo = elems + 3 * e++;
                lo[0] = sw + 1; lo[1] = sw; lo[2] = parity == 1 ? nw : nw + 1;
            }
        return 0;
    }
    if (dim == 3) {
        const int W = ex, H = ey, D = ez;
        for (int k = 0; k <= D; ++k)
            for (int j = 0; j <= H; ++j)
                for (int i = 0; i <= W; ++i) {
                    const long n = ((long)k * (H + 1) + j) * (W + 1) + i;
                    coords[3 * n] = i * h; coords[3 * n + 1] = j * h; coords[3 * n + 2] = k * h;
                }
        static const int tets[6][4] = {{0, 1, 5, 7}, {0, 1, 3, 7}, {0, 2, 3, 7},
                                       {0, 2, 6, 7}, {0, 4, 6, 7}, {0, 4, 5, 7}};
        long e = 0;
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H; ++j)
                for (int i = 0; i < W; ++i) {
                    int corner[8], c = 0;
                    for (int z = 0; z < 2; ++z)
                        for (int y = 0; y < 2; ++y)
                            for (int x = 0; x < 2; ++x)
                                corner[c++] = i + x + (W + 1) * (j + y + (H + 1) * (k + z));
                    for (int m = 0; m < 6; ++m, ++e)
                        for (int q = 0; q < 4; ++q) elems[4 * e + q] = corner[tets[m][q]];
                }
        return 0;
    }
    return -1;
}

// Element integrals with the reference's 2nd-order Gauss rules
// (pde/src/solver/GaussianQuadratureRule.cpp): 1D 2 points, triangles 3
// points, tetrahedra 4 points; P1 basis phi_0 = 1 - sum(xi), phi_k = xi_k.
int orc_model_initial_state(int model, double* y)
{
    const int S = num_states(model);
    if (S < 0) return -1;
    const double* y0 = initial_state(model);
    for (int s = 0; s < S; ++s) y[s] = y0[s];
    return S;
}

int orc_assemble(int dim, int n, const double* coords, int ne, const int* elems,
                 const double* sigma_all, int sigma_per_elem, int lumped, int* nnz_out,
                 int* row_ptr, int* col, double* mass, double* stiff)
{
    if (dim < 1 || dim > 3) return -1;
    const int nb = dim + 1;
    // Pattern: node adjacency through elements.
    std::vector<std::vector<int>> adj(n);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) adj[elems[e * nb + a]].push_back(elems[e * nb + b]);
    long nnz = 0;
    for (auto& r : adj) {
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        nnz += r.size();
    }
    *nnz_out = (int)nnz;
    if (!col) return 0;
    row_ptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        row_ptr[i + 1] = row_ptr[i] + (int)adj[i].size();
        std::copy(adj[i].begin(), adj[i].end(), col + row_ptr[i]);
    }
    std::fill(mass, mass + nnz, 0.0);
    std::fill(stiff, stiff + nnz, 0.0);

    std::vector<std::vector<double>> qp;
    std::vector<double> qw;
    quad_rule(dim, qp, qw);

    auto find = [&](int r, int c) {
        const int* b = col + row_ptr[r];
        const int* e = col + row_ptr[r + 1];
        return (int)(std::lower_bound(b, e, c) - col);
    };

    for (int e = 0; e < ne; ++e) {
        const int* nd = elems + e * nb;
        const double* sigma = sigma_per_elem ? sigma_all + (size_t)e * dim * dim : sigma_all;
        // Jacobian J[r][c] = d x_r / d xi_c = x_{c+1,r} - x_{0,r}
        double J[3][3] = {}, Ji[3][3] = {};
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < dim; ++c) J[r][c] = coords[nd[c + 1] * dim + r] - coords[nd[0] * dim + r];
        double det;
        if (dim == 1) { det = J[0][0]; Ji[0][0] = 1.0 / det; }
        else if (dim == 2) {
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            Ji[0][0] = J[1][1] / det; Ji[0][1] = -J[0][1] / det;
            Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
        } else {
            det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        }
        const double adet = std::fabs(det);
        // Physical gradients: grad phi_k = Ji^T grad_ref phi_k (constant per element).
        double gref[4][3] = {};
        for (int c = 0; c < dim; ++c) { gref[0][c] = -1.0; gref[c + 1][c] = 1.0; }
        double g[4][3] = {};
        for (int k = 0; k < nb; ++k)
            for (int r = 0; r < dim; ++r) {
                double s = 0;
                for (int c = 0; c < dim; ++c) s += Ji[c][r] * gref[k][c];
                g[k][r] = s;
            }
        double Me[4][4] = {}, Ke[4][4] = {};
        for (size_t q = 0; q < qw.size(); ++q) {
            double phi[4];
            double sum = 0;
            for (int c = 0; c < dim; ++c) { phi[c + 1] = qp[q][c]; sum += qp[q][c]; }
            phi[0] = 1.0 - sum;
            const double wJ = qw[q] * adet;
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b) {
                    Me[a][b] += wJ * phi[a] * phi[b];
                    double s = 0;
                    for (int r = 0; r < dim; ++r)
                        for (int c = 0; c < dim; ++c) s += g[a][r] * sigma[r * dim + c] * g[b][c];
                    Ke[a][b] += wJ * s;
                }
        }
        if (lumped) {
            // MassMatrixAssembler.hpp:88-101: row sums onto the diagonal.
            for (int a = 0; a < nb; ++a) {
                double s = 0;
                for (int b = 0; b < nb; ++b) { s += Me[a][b]; Me[a][b] = 0; }
                Me[a][a] = s;
            }
        }
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) {
                const int p = find(nd[a], nd[b]);
                mass[p] += Me[a][b];
                stiff[p] += Ke[a][b];
            }
    }
    return 0;
}

int orc_pcg(int n, const int* row_ptr, const int* col, const double* val,
            const double* b, double* x, double rtol, double atol, int max_its)
{
    std::vector<double> r(n), z(n), p(n), q(n), dinv(n);
    for (int i = 0; i < n; ++i) {
        double d = 0;
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
            if (col[k] == i) d = val[k];
        dinv[i] = 1.0 / d;
    }
    auto spmv = [&](const double* v, double* out) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) s += val[k] * v[col[k]];
            out[i] = s;
        }
    };
    double bnorm2 = 0;
#pragma omp parallel for reduction(+ : bnorm2) schedule(static)
    for (int i = 0; i < n; ++i) bnorm2 += (dinv[i] * b[i]) * (dinv[i] * b[i]);
    const double tol = std::max(rtol * std::sqrt(bnorm2), atol);
    spmv(x, q.data());
    double rz = 0, zz = 0;
#pragma omp parallel for reduction(+ : rz, zz) schedule(static)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
        rz += r[i] * z[i];
        zz += z[i] * z[i];
    }
    if (std::sqrt(zz) <= tol) return 0;
    for (int it = 1; it <= max_its; ++it) {
        spmv(p.data(), q.data());
        double pq = 0;
#pragma omp parallel for reduction(+ : pq) schedule(static)
        for (int i = 0; i < n; ++i) pq += p[i] * q[i];
        const double alpha = rz / pq;
        double rz_new = 0;
        zz = 0;
#pragma omp parallel for reduction(+ : rz_new, zz) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = dinv[i] * r[i];
            rz_new += r[i] * z[i];
            zz += z[i] * z[i];
        }
        if (std::sqrt(zz) <= tol) return it;
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return -1;
}

int orc_mono_run(const orc_mono_cfg* cfg, double* V_out, double* state_out,
                 double* act_out, double act_threshold,
                 int n_trace, const int* trace_nodes, double* v_trace,
                 orc_mono_stats* stats)
{
    const auto t_start = std::chrono::steady_clock::now();
#ifdef _OPENMP
    if (cfg->threads > 0) omp_set_num_threads(cfg->threads);
#endif
    const int dim = cfg->dim, model = cfg->model;
    const int integ = cfg->integrator;
    const int fails0 = g_newton_failures.load();
    int n, ne;
    std::vector<double> coords;
    std::vector<int> elems;
    if (cfg->mesh_coords) {
        n = cfg->mesh_n_nodes;
        ne = cfg->mesh_n_elems;
        coords.assign(cfg->mesh_coords, cfg->mesh_coords + (size_t)n * dim);
        elems.assign(cfg->mesh_elems, cfg->mesh_elems + (size_t)ne * (dim + 1));
    } else {
        if (orc_slab_sizes(dim, cfg->ex, cfg->ey, cfg->ez, &n, &ne)) return -1;
        coords.resize((size_t)n * dim);
        elems.resize((size_t)ne * (dim + 1));
        orc_slab_fill(dim, cfg->ex, cfg->ey, cfg->ez, cfg->h, coords.data(), elems.data());
    }
    // Cell type of each node; S = the largest state count.  Rush-Larsen and
    // backward-Euler cells step by dt_ode always
    // (AbstractRushLarsenCardiacCell.cpp:106-153,
    // AbstractBackwardEulerCardiacCell.hpp:204-245), forward Euler by the
    // stepper's next step (AbstractOneStepIvpOdeSolver.cpp:115-153).
    std::vector<int> nmodel(n, model), ninteg(n, integ);
    if (cfg->node_model) nmodel.assign(cfg->node_model, cfg->node_model + n);
    if (cfg->node_integrator) ninteg.assign(cfg->node_integrator, cfg->node_integrator + n);
    int S = 0;
    for (int i = 0; i < n; ++i) {
        const int Si = num_states(nmodel[i]);
        if (Si < 0 || !integrator_ok(nmodel[i], ninteg[i]))
            return -1;
        S = std::max(S, Si);
    }
    if (S <= 0) return -1;
    const bool mixed_cells = cfg->node_model || cfg->node_integrator;
    if (mixed_cells && (cfg->svi || cfg->params)) return -1;
    double sigma[9] = {};
    for (int r = 0; r < dim; ++r) sigma[r * dim + r] = cfg->sigma[r];
    const double* sig = cfg->sigma_elem ? cfg->sigma_elem : sigma;
    const int per_elem = cfg->sigma_elem ? 1 : 0;
    int nnz;
    orc_assemble(dim, n, coords.data(), ne, elems.data(), sig, per_elem, cfg->lumped, &nnz, nullptr, nullptr,
                 nullptr, nullptr);
    std::vector<int> rp(n + 1), col(nnz);
    std::vector<double> M(nnz), K(nnz), A(nnz);
    orc_assemble(dim, n, coords.data(), ne, elems.data(), sig, per_elem, cfg->lumped, &nnz, rp.data(), col.data(),
                 M.data(), K.data());
    // A = (Am Cm / dt) M + K  (MonodomainAssembler.cpp:42-66)
    const double am = cfg->am, cm = cfg->cm, dt = cfg->dt_pde;
    for (int k = 0; k < nnz; ++k) A[k] = (1.0 / dt) * (am * cm * M[k]) + K[k];

    // Cells: state-major planes y[s*n + i]; V lives in its own vector.
    std::vector<double> Y((size_t)S * n);
    std::vector<double> V(n), Vprev(n), iion(n), istim(n), z(n), b(n), svi(cfg->svi ? n : 0);
    for (int i = 0; i < n; ++i) {
        const double* y0 = initial_state(nmodel[i]);
        for (int s = 0; s < num_states(nmodel[i]); ++s) Y[(size_t)s * n + i] = y0[s];
        V[i] = y0[0];
    }
    std::vector<char> stimulated(n);
    for (int i = 0; i < n; ++i)
        stimulated[i] = cfg->stim_mask ? cfg->stim_mask[i] != 0 : coords[(size_t)i * dim] < cfg->stim_xlim;
    Stimulus stim(cfg->stim_mag, cfg->stim_dur, cfg->stim_start, cfg->stim_period > 0 ? cfg->stim_period : -1.0);
    // ActivationOutputModifier::ProcessSolutionAtTimeStep
    // (heart/src/problem/ActivationOutputModifier.cpp:94-125), called with the
    // initial condition and after every printing step
    // (heart/src/problem/AbstractCardiacProblem.cpp:459-464, 552-555).
    auto modifier = [&](double time) {
        double* m = cfg->modifier_out;
        const double thr = cfg->modifier_threshold;
        for (int i = 0; i < n; ++i) {
            const double v = V[i];
            if (v > thr && m[i] < 0.0) m[i] = time;
            else if (m[i] >= 0.0 && m[n + i] < 0.0 && v < thr) m[n + i] = time;
            else if (m[n + i] >= 0.0 && m[2 * n + i] < 0.0 && v > thr) m[2 * n + i] = time;
            else if (m[2 * n + i] >= 0.0 && m[3 * n + i] < 0.0 && v < thr) m[3 * n + i] = time;
        }
    };
    if (cfg->modifier_out) {
        std::fill(cfg->modifier_out, cfg->modifier_out + 4 * (size_t)n, -1.0);
        modifier(0.0);
    }
    if (act_out) std::fill(act_out, act_out + n, -1.0);

    orc_mono_stats st{};
    st.n_nodes = n;
    st.nnz = nnz;
    // Operator splitting (heart/src/solver/electrics/monodomain/
    // OperatorSplittingMonodomainSolver.cpp:100-116): SolveCellSystems with
    // updateVoltage over [t0, t1], every node's full state (V included) stepped
    // by SolveAndUpdateState (heart/src/tissue/AbstractCardiacTissue.cpp:531-536).
    std::vector<double> Vhalf(n);
    auto cells_with_voltage = [&](double t0, double t1) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const int mi = nmodel[i], ii = ninteg[i], Si = num_states(mi), P = num_params(mi);
            const double* pdef = default_params(mi);
            const bool fixed_dt = ii != ORC_FE;
            double y[64], p[64];
            for (int s = 0; s < Si; ++s) y[s] = Y[(size_t)s * n + i];
            for (int q = 0; q < P; ++q)
                p[q] = (cfg->params && q < cfg->n_param) ? cfg->params[(size_t)q * n + i] : pdef[q];
            y[0] = V[i];
            Stepper ode(t0, t1, cfg->dt_ode);
            while (!ode.at_end()) {
                const double ts = ode.time;
                const double sa = stimulated[i] ? stim.at(ts) / am : 0.0;
                cell_substep(mi, ii, p, cm, y, ts, fixed_dt ? cfg->dt_ode : ode.step_size(), sa, true);
                ode.advance();
            }
            V[i] = y[0];
            for (int s = 1; s < Si; ++s) Y[(size_t)s * n + i] = y[s];
        }
    };
    Stepper pde(0.0, cfg->t_end, dt);
    int step = 0;
    while (!pde.at_end()) {
        const double t = pde.time, t_next = pde.next;
        if (cfg->split) {
            const double th = t + 0.5 * (t_next - t);
            Vprev = V;
            auto c0 = std::chrono::steady_clock::now();
            cells_with_voltage(t, th);
            Vhalf = V;
            st.t_ode += seconds_since(c0);
            // diffusion only: b = M (Am Cm V / dt), initial guess V
            c0 = std::chrono::steady_clock::now();
            const double c = am * cm / (t_next - t);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                double s = 0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) s += M[k] * (c * V[col[k]]);
                b[i] = s;
            }
            st.t_rhs += seconds_since(c0);
            c0 = std::chrono::steady_clock::now();
            const int its = orc_pcg(n, rp.data(), col.data(), A.data(), b.data(), V.data(),
                                    cfg->ksp_rtol, cfg->ksp_atol, cfg->max_its);
            st.t_ksp += seconds_since(c0);
            if (its < 0) return -2;
            st.cg_iterations += its;
            c0 = std::chrono::steady_clock::now();
            cells_with_voltage(th, t_next);
            st.t_ode += seconds_since(c0);
            if (g_newton_failures.load() != fails0) return -3;
            pde.advance();
            if (cfg->modifier_out && (step + 1) % std::max(cfg->modifier_every, 1) == 0) modifier(t_next);
            if (act_out) {
                for (int i = 0; i < n; ++i) {
                    if (act_out[i] >= 0) continue;
                    if (Vprev[i] < act_threshold && Vhalf[i] >= act_threshold)
                        act_out[i] = t + (th - t) * (act_threshold - Vprev[i]) / (Vhalf[i] - Vprev[i]);
                    else if (Vhalf[i] < act_threshold && V[i] >= act_threshold)
                        act_out[i] = th + (t_next - th) * (act_threshold - Vhalf[i]) / (V[i] - Vhalf[i]);
                }
            }
            if (v_trace)
                for (int q = 0; q < n_trace; ++q) v_trace[(size_t)step * n_trace + q] = V[trace_nodes[q]];
            ++step;
            continue;
        }
        // SolveCellSystems(V, t, t_next): every node integrated with V clamped
        // over [t, t_next] in ODE substeps, then Iion and Istim(t_next) cached.
        auto c0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const int mi = nmodel[i], ii = ninteg[i], Si = num_states(mi), P = num_params(mi);
            const double* pdef = default_params(mi);
            const bool fixed_dt = ii != ORC_FE;
            double y[64], p[64];
            for (int s = 0; s < Si; ++s) y[s] = Y[(size_t)s * n + i];
            for (int q = 0; q < P; ++q)
                p[q] = (cfg->params && q < cfg->n_param) ? cfg->params[(size_t)q * n + i] : pdef[q];
            y[0] = V[i];
            Stepper ode(t, t_next, cfg->dt_ode);
            while (!ode.at_end()) {
                const double ts = ode.time;
                const double sa = stimulated[i] ? stim.at(ts) / am : 0.0;
                cell_substep(mi, ii, p, cm, y, ts, fixed_dt ? cfg->dt_ode : ode.step_size(), sa, false);
                ode.advance();
            }
            y[0] = V[i];
            for (int s = 1; s < Si; ++s) Y[(size_t)s * n + i] = y[s];
            iion[i] = cell_iionic(mi, p, cm, y);
            istim[i] = stimulated[i] ? stim.at(t_next) : 0.0;
        }
        st.t_ode += seconds_since(c0);
        // z = Am Cm V / dt - Am Iion - Istim ; b = M z
        c0 = std::chrono::steady_clock::now();
        const double inv_dt = 1.0 / (t_next - t);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) z[i] = am * cm * V[i] * inv_dt + (-am * iion[i] - istim[i]);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = rp[i]; k < rp[i + 1]; ++k) s += M[k] * z[col[k]];
            b[i] = s;
        }
        if (cfg->svi) {
            // MonodomainSolver.cpp:122-127: the correction is added to b = M z
            svi_correction(dim, n, coords.data(), ne, elems.data(), model, am, cm, Y.data(), V.data(), iion.data(),
                           cfg->params, cfg->n_param, svi.data());
            for (int i = 0; i < n; ++i) b[i] += svi[i];
        }
        st.t_rhs += seconds_since(c0);
        c0 = std::chrono::steady_clock::now();
        Vprev = V;
        const int its = orc_pcg(n, rp.data(), col.data(), A.data(), b.data(), V.data(),
                                cfg->ksp_rtol, cfg->ksp_atol, cfg->max_its);
        st.t_ksp += seconds_since(c0);
        if (its < 0) return -2;
        st.cg_iterations += its;
        if (g_newton_failures.load() != fails0) return -3;
        pde.advance();
        if (cfg->modifier_out && (step + 1) % std::max(cfg->modifier_every, 1) == 0) modifier(t_next);
        if (act_out) {
            for (int i = 0; i < n; ++i)
                if (act_out[i] < 0 && Vprev[i] < act_threshold && V[i] >= act_threshold)
                    act_out[i] = t + (t_next - t) * (act_threshold - Vprev[i]) / (V[i] - Vprev[i]);
        }
        if (v_trace)
            for (int q = 0; q < n_trace; ++q) v_trace[(size_t)step * n_trace + q] = V[trace_nodes[q]];
        ++step;
    }
    st.steps = step;
    st.t_total = seconds_since(t_start);
    if (V_out) std::copy(V.begin(), V.end(), V_out);
    if (state_out) {
        std::copy(Y.begin(), Y.end(), state_out);
        std::copy(V.begin(), V.end(), state_out);
    }
    if (stats) *stats = st;
    return 0;
}

}  // extern "C"
