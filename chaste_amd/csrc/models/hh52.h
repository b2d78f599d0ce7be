// Generated by chaste_amd/codegen/cellml2cpp.py from HodgkinHuxley1952.cellml -- do not edit.
// Model: hodgkin_huxley_squid_axon_model_1952_modified
#pragma once
#include "model_common.h"

namespace cmx { namespace hh52 {

constexpr int kNumStates = 4;
constexpr int kVIndex = 0;
constexpr int kNumParams = 3;
constexpr bool kPerCapacitanceUnits = false;

// State variables (CellML component__variable), V first then document order.
static const char* const kStateNames[kNumStates] = {
    "membrane__V",
    "sodium_channel_m_gate__m",
    "sodium_channel_h_gate__h",
    "potassium_channel_n_gate__n",
};
static constexpr double kInitialState[kNumStates] = {
    -75.0,  // membrane__V
    0.05,  // sodium_channel_m_gate__m
    0.6,  // sodium_channel_h_gate__h
    0.325,  // potassium_channel_n_gate__n
};
static const char* const kParamNames[3] = {
    "sodium_channel__g_Na",
    "potassium_channel__g_K",
    "leakage_current__g_L",
};
// The names the reference gives the same parameters (the CellML oxmeta term,
// else the variable's cmeta:id; empty when the file annotates neither).
static const char* const kParamAliases[3] = {
    "membrane_fast_sodium_current_conductance",
    "",
    "",
};
static constexpr double kParamDefaults[3] = {
    120.0,
    36.0,
    0.3,
};
// 1 = integrated by Rush-Larsen (affine gate, V-dependent coefficients only)
static constexpr int kIsGate[kNumStates] = {0, 1, 1, 1};

// Right-hand side with V clamped.  Non-gates (and, with RL=false, every
// state) get dy[i]; with RL=true gate i gets g_inf[i], g_k[i] such that
// y_i(t+dt) = g_inf[i] + (y_i - g_inf[i]) * exp(g_k[i] * dt).
// i_stim_area is in uA/cm^2; cm is the tissue capacitance in uF/cm^2.
template <class R, bool RL>
CMX_HD inline void evaluate(const R* y, R V, R i_stim_area, const R* p, R cm,
                            R* dy, R* g_inf, R* g_k)
{
    CMX_MATH_USING
    (void)p; (void)cm; (void)g_inf; (void)g_k; (void)i_stim_area;
    const R i_stim = i_stim_area;
    (void)i_stim;
    const R potassium_channel_n_gate__alpha_n = cmx_remove_singularity(V, R(-65.00000000000023), R(1.0000000000000036e-06), [&](const R V) { return ((((V < R(-64.9999)) && (V > R(-65.0001)))) ? (R(0.1)) : (cmx_div(((-R(0.01)) * (V + R(65.0))), cmx_expm1(((-(V + R(65.0))) * (R(1) / R(10.0))))))); });
    const R potassium_channel_n_gate__beta_n = (R(0.125) * cmx_exp_v(((V + R(75.0)) * (R(1) / R(80.0))), V, 0.018033688011112044, 1.3525266008334031));
    const R sodium_channel_h_gate__alpha_h = (R(0.07) * cmx_exp_v(((-(V + R(75.0))) * (R(1) / R(20.0))), V, -0.07213475204444818, -5.410106403333613));
    const R sodium_channel_h_gate__beta_h = cmx_div(R(1.0), (cmx_exp_v(((-(V + R(45.0))) * (R(1) / R(10.0))), V, -0.14426950408889636, -6.4921276840003355) + R(1.0)));
    const R sodium_channel_m_gate__alpha_m = cmx_remove_singularity(V, R(-50.00000000000018), R(1.0000000000000036e-06), [&](const R V) { return ((((V < R(-49.99999)) && (V > R(-50.00001)))) ? (R(1.0)) : (cmx_div(((-R(0.1)) * (V + R(50.0))), cmx_expm1(((-(V + R(50.0))) * (R(1) / R(10.0))))))); });
    const R sodium_channel_m_gate__beta_m = (R(4.0) * cmx_exp_v(((-(V + R(75.0))) * (R(1) / R(18.0))), V, -0.0801497244938313, -6.0112293370373475));
    if (RL) {
        g_k[1] = ((sodium_channel_m_gate__alpha_m * (-R(1))) - (sodium_channel_m_gate__beta_m * R(1)));
        g_inf[1] = cmx_div(-((sodium_channel_m_gate__alpha_m * R(1.0))), ((sodium_channel_m_gate__alpha_m * (-R(1))) - (sodium_channel_m_gate__beta_m * R(1))));
    } else {
        dy[1] = ((sodium_channel_m_gate__alpha_m * (R(1.0) - y[1])) - (sodium_channel_m_gate__beta_m * y[1]));
    }
    if (RL) {
        g_k[2] = ((sodium_channel_h_gate__alpha_h * (-R(1))) - (sodium_channel_h_gate__beta_h * R(1)));
        g_inf[2] = cmx_div(-((sodium_channel_h_gate__alpha_h * R(1.0))), ((sodium_channel_h_gate__alpha_h * (-R(1))) - (sodium_channel_h_gate__beta_h * R(1))));
    } else {
        dy[2] = ((sodium_channel_h_gate__alpha_h * (R(1.0) - y[2])) - (sodium_channel_h_gate__beta_h * y[2]));
    }
    if (RL) {
        g_k[3] = ((potassium_channel_n_gate__alpha_n * (-R(1))) - (potassium_channel_n_gate__beta_n * R(1)));
        g_inf[3] = cmx_div(-((potassium_channel_n_gate__alpha_n * R(1.0))), ((potassium_channel_n_gate__alpha_n * (-R(1))) - (potassium_channel_n_gate__beta_n * R(1))));
    } else {
        dy[3] = ((potassium_channel_n_gate__alpha_n * (R(1.0) - y[3])) - (potassium_channel_n_gate__beta_n * y[3]));
    }
}

// dV/dt including the stimulus (single-cell integration).
template <class R>
CMX_HD inline R dvdt(const R* y, R V, R i_stim_area, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__E_R = R(-75.0);
    const R membrane__Cm = R(1.0);
    const R i_stim = i_stim_area;
    (void)i_stim;
    const R leakage_current__E_L = (membrane__E_R + R(10.613));
    const R leakage_current__i_L = (p[2] * (V - leakage_current__E_L));
    const R potassium_channel__E_K = (membrane__E_R - R(12.0));
    const R potassium_channel__i_K = (p[1] * cmx_ipow<4>(y[3]) * (V - potassium_channel__E_K));
    const R sodium_channel__E_Na = (membrane__E_R + R(115.0));
    const R sodium_channel__i_Na = (p[0] * cmx_ipow<3>(y[1]) * y[2] * (V - sodium_channel__E_Na));
    return ((-(i_stim + sodium_channel__i_Na + potassium_channel__i_K + leakage_current__i_L)) * (R(1) / membrane__Cm));
}

// Total ionic current in uA/cm^2 (GetIIonic): minus dV/dt at zero stimulus,
// times the membrane capacitance that converts it to a current density.
template <class R>
CMX_HD inline R iionic(const R* y, R V, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__E_R = R(-75.0);
    const R membrane__Cm = R(1.0);
    const R i_stim = R(0);
    (void)i_stim;
    const R leakage_current__E_L = (membrane__E_R + R(10.613));
    const R leakage_current__i_L = (p[2] * (V - leakage_current__E_L));
    const R potassium_channel__E_K = (membrane__E_R - R(12.0));
    const R potassium_channel__i_K = (p[1] * cmx_ipow<4>(y[3]) * (V - potassium_channel__E_K));
    const R sodium_channel__E_Na = (membrane__E_R + R(115.0));
    const R sodium_channel__i_Na = (p[0] * cmx_ipow<3>(y[1]) * y[2] * (V - sodium_channel__E_Na));
    const R dv = ((-(i_stim + sodium_channel__i_Na + potassium_channel__i_K + leakage_current__i_L)) * (R(1) / membrane__Cm));
    return -dv * membrane__Cm;
}

// Backward Euler: the states solved by Newton (all but the linear gates).
constexpr int kNumNonlinear = 0;
static constexpr int kNonlinear[1] = {0};

// Linear states: y <- (y + a dt) / (1 - b dt) for dy/dt = a + b y, a and b
// functions of V (ComputeOneStepExceptVoltage, closed-form part).
template <class R>
CMX_HD inline void be_linear(R* y, R V, const R* p, R cm, R dt)
{
    CMX_MATH_USING
    (void)p; (void)cm; (void)V;
    const R potassium_channel_n_gate__alpha_n = cmx_remove_singularity(V, R(-65.00000000000023), R(1.0000000000000036e-06), [&](const R V) { return ((((V < R(-64.9999)) && (V > R(-65.0001)))) ? (R(0.1)) : (cmx_div(((-R(0.01)) * (V + R(65.0))), cmx_expm1(((-(V + R(65.0))) * (R(1) / R(10.0))))))); });
    const R potassium_channel_n_gate__beta_n = (R(0.125) * cmx_exp_v(((V + R(75.0)) * (R(1) / R(80.0))), V, 0.018033688011112044, 1.3525266008334031));
    const R sodium_channel_h_gate__alpha_h = (R(0.07) * cmx_exp_v(((-(V + R(75.0))) * (R(1) / R(20.0))), V, -0.07213475204444818, -5.410106403333613));
    const R sodium_channel_h_gate__beta_h = cmx_div(R(1.0), (cmx_exp_v(((-(V + R(45.0))) * (R(1) / R(10.0))), V, -0.14426950408889636, -6.4921276840003355) + R(1.0)));
    const R sodium_channel_m_gate__alpha_m = cmx_remove_singularity(V, R(-50.00000000000018), R(1.0000000000000036e-06), [&](const R V) { return ((((V < R(-49.99999)) && (V > R(-50.00001)))) ? (R(1.0)) : (cmx_div(((-R(0.1)) * (V + R(50.0))), cmx_expm1(((-(V + R(50.0))) * (R(1) / R(10.0))))))); });
    const R sodium_channel_m_gate__beta_m = (R(4.0) * cmx_exp_v(((-(V + R(75.0))) * (R(1) / R(18.0))), V, -0.0801497244938313, -6.0112293370373475));
    y[1] = (y[1] + ((sodium_channel_m_gate__alpha_m * R(1.0))) * dt) / (R(1) - (((sodium_channel_m_gate__alpha_m * (-R(1))) - (sodium_channel_m_gate__beta_m * R(1)))) * dt);
    y[2] = (y[2] + ((sodium_channel_h_gate__alpha_h * R(1.0))) * dt) / (R(1) - (((sodium_channel_h_gate__alpha_h * (-R(1))) - (sodium_channel_h_gate__beta_h * R(1)))) * dt);
    y[3] = (y[3] + ((potassium_channel_n_gate__alpha_n * R(1.0))) * dt) / (R(1) - (((potassium_channel_n_gate__alpha_n * (-R(1))) - (potassium_channel_n_gate__beta_n * R(1)))) * dt);
}

// Newton residual of the nonlinear states at guess g (y holds the other
// states, linear ones already advanced): g - y_old - dt f(g).
template <class R>
CMX_HD inline void be_residual(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* res)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R i_stim = i_stim_area;
    (void)i_stim;
}

// Jacobian of be_residual in g: J[i*N + j] = delta_ij - dt d f_i / d g_j.
template <class R>
CMX_HD inline void be_jacobian(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* J)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R i_stim = i_stim_area;
    (void)i_stim;
}

// One backward-Euler step of every state but V at the given V
// (ComputeOneStepExceptVoltage): linear states in closed form, then the
// nonlinear ones by Newton from their current values.  Returns the Newton
// iteration count, -1 if it did not converge.
template <class R>
CMX_HD inline int be_step(R* y, R V, R i_stim_area, const R* p, R cm, R dt)
{
    be_linear<R>(y, V, p, cm, dt);
    (void)i_stim_area;
    return 0;
}

// Rates of the full system: dV/dt (0 when V is clamped) and the states.
template <class R>
CMX_HD inline void grl_rates(const R* y, R i_stim_area, const R* p, R cm, bool update_v, R* f)
{
    R gi[kNumStates], gk[kNumStates];
    evaluate<R, false>(y, y[0], i_stim_area, p, cm, f, gi, gk);
    f[0] = update_v ? dvdt<R>(y, y[0], i_stim_area, p, cm) : R(0);
}

template <class R>
CMX_HD inline R grl_advance(R y0, R f, R a, R dt)
{
    using std::fabs; using std::exp;
    return fabs(a) < R(1e-8) ? y0 + f * dt : y0 + (f / a) * (exp(a * dt) - R(1));
}

// One generalized Rush-Larsen step of length dt (ORDER 1 or 2); V is
// advanced too when update_v, else held.
template <class R, int ORDER>
CMX_HD inline void grl_step(R* y, R i_stim_area, const R* p, R cm, R dt, bool update_v)
{
    constexpr int S = kNumStates;
    const R delta = R(1e-8);
    R f[S], a[S], y0[S], t[S];
    const int first = update_v ? 0 : 1;
CMX_UNROLL
    for (int i = 0; i < S; ++i) y0[i] = y[i];
    grl_rates<R>(y, i_stim_area, p, cm, update_v, f);
CMX_UNROLL
    for (int i = 0; i < S; ++i) {
        if (i < first) continue;
        const R ys = y[i];
        y[i] = ys + delta;
        grl_rates<R>(y, i_stim_area, p, cm, update_v, t);
        a[i] = (t[i] - f[i]) / delta;
        y[i] = ys;
    }
    if constexpr (ORDER == 1) {
CMX_UNROLL
        for (int i = 0; i < S; ++i)
            if (i >= first) y[i] = grl_advance<R>(y0[i], f[i], a[i], dt);
        return;
    }
    R half = dt * R(0.5);
CMX_UNROLL
    for (int i = 0; i < S; ++i)
        if (i >= first) y[i] = grl_advance<R>(y0[i], f[i], a[i], half);
    R g[S];
CMX_UNROLL
    for (int i = 0; i < S; ++i) {
        if (i < first) continue;
        const R ym = y[i];
        y[i] = y0[i];
        grl_rates<R>(y, i_stim_area, p, cm, update_v, t);
        g[i] = t[i];
        y[i] = y0[i] + delta;
        grl_rates<R>(y, i_stim_area, p, cm, update_v, t);
        a[i] = (t[i] - g[i]) / delta;
        y[i] = ym;
    }
CMX_UNROLL
    for (int i = 0; i < S; ++i)
        if (i >= first) y[i] = grl_advance<R>(y0[i], g[i], a[i], dt);
}

}}  // namespace cmx::hh52
