// Generated by chaste_amd/codegen/cellml2cpp.py from FoxModel2002.cellml -- do not edit.
// Model: fox_model_2002
#pragma once
#include "model_common.h"

namespace cmx { namespace fox2002 {

constexpr int kNumStates = 13;
constexpr int kVIndex = 0;
constexpr int kNumParams = 5;
constexpr bool kPerCapacitanceUnits = true;

// State variables (CellML component__variable), V first then document order.
static const char* const kStateNames[kNumStates] = {
    "membrane__V",
    "fast_sodium_current_m_gate__m",
    "fast_sodium_current_h_gate__h",
    "fast_sodium_current_j_gate__j",
    "rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__X_kr",
    "slow_activating_delayed_rectifiyer_K_current_X_ks_gate__X_ks",
    "transient_outward_potassium_current_X_to_gate__X_to",
    "transient_outward_potassium_current_Y_to_gate__Y_to",
    "L_type_Ca_current_f_gate__f",
    "L_type_Ca_current_d_gate__d",
    "L_type_Ca_current_f_Ca_gate__f_Ca",
    "calcium_dynamics__Ca_i",
    "calcium_dynamics__Ca_SR",
};
static constexpr double kInitialState[kNumStates] = {
    -94.7,  // membrane__V
    0.00024676,  // fast_sodium_current_m_gate__m
    0.99869,  // fast_sodium_current_h_gate__h
    0.99887,  // fast_sodium_current_j_gate__j
    0.229,  // rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__X_kr
    0.0001,  // slow_activating_delayed_rectifiyer_K_current_X_ks_gate__X_ks
    3.742e-05,  // transient_outward_potassium_current_X_to_gate__X_to
    1.0,  // transient_outward_potassium_current_Y_to_gate__Y_to
    0.983,  // L_type_Ca_current_f_gate__f
    0.0001,  // L_type_Ca_current_d_gate__d
    0.942,  // L_type_Ca_current_f_Ca_gate__f_Ca
    0.0472,  // calcium_dynamics__Ca_i
    320.0,  // calcium_dynamics__Ca_SR
};
static const char* const kParamNames[5] = {
    "fast_sodium_current__g_Na",
    "rapid_activating_delayed_rectifiyer_K_current__g_Kr",
    "slow_activating_delayed_rectifiyer_K_current__g_Ks",
    "time_independent_potassium_current__g_K1",
    "L_type_Ca_current__P_Ca",
};
// The names the reference gives the same parameters (the CellML oxmeta term,
// else the variable's cmeta:id; empty when the file annotates neither).
static const char* const kParamAliases[5] = {
    "",
    "",
    "",
    "",
    "",
};
static constexpr double kParamDefaults[5] = {
    12.8,
    0.0136,
    0.0245,
    2.8,
    1.26e-05,
};
// 1 = integrated by Rush-Larsen (affine gate, V-dependent coefficients only)
static constexpr int kIsGate[kNumStates] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0};

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
    const R membrane__R = R(8.314);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96.5);
    const R Na_Ca_exchanger__K_mCa = R(1380.0);
    const R Na_Ca_exchanger__K_mNa = R(87.5);
    const R Na_Ca_exchanger__K_NaCa = R(1500.0);
    const R Na_Ca_exchanger__K_sat = R(0.2);
    const R Na_Ca_exchanger__eta = R(0.35);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.05);
    const R sarcolemmal_calcium_pump__i_pCa_max = R(0.05);
    const R calcium_background_current__g_Cab = R(0.0003842);
    const R L_type_Ca_current__C_sc = R(1.0);
    const R L_type_Ca_current_f_Ca_gate__K_mfCa = R(0.18);
    const R calcium_dynamics__K_mCMDN = R(2.0);
    const R calcium_dynamics__CMDN_tot = R(10.0);
    const R calcium_dynamics__V_myo = R(2.584e-05);
    const R calcium_dynamics__A_Cap = R(0.0001534);
    const R calcium_dynamics__P_rel = R(6.0);
    const R calcium_dynamics__P_leak = R(1e-06);
    const R calcium_dynamics__K_mCSQN = R(600.0);
    const R calcium_dynamics__CSQN_tot = R(10000.0);
    const R calcium_dynamics__V_SR = R(2e-06);
    const R calcium_dynamics__V_up = R(0.1);
    const R calcium_dynamics__K_mup = R(0.32);
    const R standard_ionic_concentrations__Na_i = R(10.0);
    const R standard_ionic_concentrations__Na_o = R(138.0);
    const R standard_ionic_concentrations__Ca_o = R(2000.0);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_Ca_max = cmx_remove_singularity(V, R(-0.0), R(1.3354093264248706e-06), [&](const R V) { return cmx_div(((((p[4] * (R(1) / L_type_Ca_current__C_sc)) * R(4.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((y[11] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10803391981328421, 0.0)) - (R(0.341) * standard_ionic_concentrations__Ca_o))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current__i_Ca = (L_type_Ca_current__i_Ca_max * y[8] * y[9] * y[10]);
    const R L_type_Ca_current_d_gate__d_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(10.0)) * (R(1) / (-R(6.24)))), V, -0.23120112834759027, -2.312011283475903)));
    const R L_type_Ca_current_d_gate__E0_m = (V + R(40.0));
    const R L_type_Ca_current_d_gate__tau_d = cmx_div(R(1.0), (cmx_div((R(0.25) * cmx_exp_v(((-R(0.01)) * V), V, -0.014426950408889635, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.07)) * V), V, -0.10098865286222744, -0.0))) + cmx_div((R(0.07) * cmx_exp_c(((-R(0.05)) * L_type_Ca_current_d_gate__E0_m), (L_type_Ca_current_d_gate__E0_m), -0.07213475204444818)), (R(1.0) + cmx_exp_c((R(0.05) * L_type_Ca_current_d_gate__E0_m), (L_type_Ca_current_d_gate__E0_m), 0.07213475204444818)))));
    const R L_type_Ca_current_f_Ca_gate__f_Ca_infinity = cmx_div(R(1.0), (R(1.0) + cmx_ipow<3>((y[11] * (R(1) / L_type_Ca_current_f_Ca_gate__K_mfCa)))));
    const R L_type_Ca_current_f_Ca_gate__tau_f_Ca = R(30.0);
    const R L_type_Ca_current_f_gate__f_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(12.5)) * (R(1) / R(5.0))), V, 0.2885390081777927, 3.6067376022224087)));
    const R L_type_Ca_current_f_gate__tau_f = (R(30.0) + cmx_div(R(200.0), (R(1.0) + cmx_exp_v(((V + R(20.0)) * (R(1) / R(9.5))), V, 0.15186263588304877, 3.0372527176609756))));
    const R Na_Ca_exchanger__i_NaCa = (cmx_div(Na_Ca_exchanger__K_NaCa, ((cmx_ipow<3>(Na_Ca_exchanger__K_mNa) + cmx_ipow<3>(standard_ionic_concentrations__Na_o)) * (Na_Ca_exchanger__K_mCa + standard_ionic_concentrations__Ca_o) * (R(1.0) + (Na_Ca_exchanger__K_sat * cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0))))) * ((cmx_exp_v(((Na_Ca_exchanger__eta * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.018905935967324736, 0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_i) * standard_ionic_concentrations__Ca_o) - (cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_o) * y[11])));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(standard_ionic_concentrations__Ca_o, y[11])));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R calcium_dynamics__J_leak = (calcium_dynamics__P_leak * (y[12] - y[11]));
    const R calcium_dynamics__gamma = cmx_div(R(1.0), (R(1.0) + cmx_ipow<3>(cmx_div(R(2000.0), y[12]))));
    const R calcium_dynamics__J_rel = cmx_div((calcium_dynamics__P_rel * y[8] * y[9] * y[10] * ((calcium_dynamics__gamma * y[12]) - y[11])), (R(1.0) + (R(1.65) * cmx_exp_v((V * (R(1) / R(20.0))), V, 0.07213475204444818, 0.0))));
    const R calcium_dynamics__J_up = cmx_div(calcium_dynamics__V_up, (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__K_mup, y[11]))));
    const R calcium_dynamics__beta_SR = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_tot * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + y[12])))));
    const R calcium_dynamics__beta_i = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_tot * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + y[11])))));
    const R fast_sodium_current_h_gate__alpha_h = (R(0.135) * cmx_exp_v(((V + R(80.0)) * (R(1) / (-R(6.8)))), V, -0.2121610354248476, -16.972882833987807));
    const R fast_sodium_current_h_gate__beta_h = cmx_div(R(7.5), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(11.0))), V, -0.14426950408889636, -1.5869645449778598)));
    const R fast_sodium_current_j_gate__alpha_j = cmx_div((R(0.175) * cmx_exp_v(((V + R(100.0)) * (R(1) / (-R(23.0)))), V, -0.06272587134299841, -6.27258713429984)), (R(1.0) + cmx_exp_v((R(0.15) * (V + R(79.0))), V, 0.21640425613334446, 17.095936234534214)));
    const R fast_sodium_current_j_gate__beta_j = cmx_div(R(0.3), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(32.0))), V, -0.14426950408889636, -4.616624130844683)));
    const R fast_sodium_current_m_gate__E0_m = (V + R(47.13));
    const R fast_sodium_current_m_gate__alpha_m = cmx_div((R(0.32) * fast_sodium_current_m_gate__E0_m), (-cmx_expm1(((-R(0.1)) * fast_sodium_current_m_gate__E0_m))));
    const R fast_sodium_current_m_gate__beta_m = (R(0.08) * cmx_exp_v(((-V) * (R(1) / R(11.0))), V, -0.1311540946262694, -0.0));
    const R rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__X_kr_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-R(2.182)) - (R(0.1819) * V)), V, -0.2624262279377025, -3.147960579219718)));
    const R rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__tau_X_kr = (R(43.0) + cmx_div(R(1.0), (cmx_exp_v(((-R(5.495)) + (R(0.1691) * V)), V, 0.2439597314143237, -7.927609249684854) + cmx_exp_v(((-R(7.677)) - (R(0.0128) * V)), V, -0.01846649652337874, -11.075569828904571))));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__i_pCa_max * y[11]), (sarcolemmal_calcium_pump__K_mpCa + y[11]));
    const R slow_activating_delayed_rectifiyer_K_current_X_ks_gate__X_ks_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(16.0)) * (R(1) / (-R(13.6)))), V, -0.10608051771242379, 1.6972882833987806)));
    const R slow_activating_delayed_rectifiyer_K_current_X_ks_gate__tau_X_ks = cmx_remove_singularity(V, R(9.999999999999991), R(6.75675675675675e-07), [&](const R V) { return cmx_div(R(1.0), (cmx_div((R(7.19e-05) * (V - R(10.0))), (-cmx_expm1(((-R(0.148)) * (V - R(10.0)))))) + cmx_div((R(0.000131) * (V - R(10.0))), cmx_expm1((R(0.0687) * (V - R(10.0))))))); });
    const R transient_outward_potassium_current_X_to_gate__alpha_X_to = (R(0.04516) * cmx_exp_v((R(0.03577) * V), V, 0.051605201612598225, 0.0));
    const R transient_outward_potassium_current_X_to_gate__beta_X_to = (R(0.0989) * cmx_exp_v(((-R(0.06237)) * V), V, -0.08998088970024463, -0.0));
    const R transient_outward_potassium_current_Y_to_gate__alpha_Y_to = cmx_div((R(0.005415) * cmx_exp_v(((V + R(33.5)) * (R(1) / (-R(5.0)))), V, -0.2885390081777927, -9.666056773956054)), (R(1.0) + (R(0.051335) * cmx_exp_v(((V + R(33.5)) * (R(1) / (-R(5.0)))), V, -0.2885390081777927, -9.666056773956054))));
    const R transient_outward_potassium_current_Y_to_gate__beta_Y_to = cmx_div((R(0.005415) * cmx_exp_v(((V + R(33.5)) * (R(1) / R(5.0))), V, 0.2885390081777927, 9.666056773956054)), (R(1.0) + (R(0.051335) * cmx_exp_v(((V + R(33.5)) * (R(1) / R(5.0))), V, 0.2885390081777927, 9.666056773956054))));
    if (RL) {
        g_k[1] = ((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1)));
        g_inf[1] = cmx_div(-((fast_sodium_current_m_gate__alpha_m * R(1.0))), ((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1))));
    } else {
        dy[1] = ((fast_sodium_current_m_gate__alpha_m * (R(1.0) - y[1])) - (fast_sodium_current_m_gate__beta_m * y[1]));
    }
    if (RL) {
        g_k[2] = ((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1)));
        g_inf[2] = cmx_div(-((fast_sodium_current_h_gate__alpha_h * R(1.0))), ((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1))));
    } else {
        dy[2] = ((fast_sodium_current_h_gate__alpha_h * (R(1.0) - y[2])) - (fast_sodium_current_h_gate__beta_h * y[2]));
    }
    if (RL) {
        g_k[3] = ((fast_sodium_current_j_gate__alpha_j * (-R(1))) - (fast_sodium_current_j_gate__beta_j * R(1)));
        g_inf[3] = cmx_div(-((fast_sodium_current_j_gate__alpha_j * R(1.0))), ((fast_sodium_current_j_gate__alpha_j * (-R(1))) - (fast_sodium_current_j_gate__beta_j * R(1))));
    } else {
        dy[3] = ((fast_sodium_current_j_gate__alpha_j * (R(1.0) - y[3])) - (fast_sodium_current_j_gate__beta_j * y[3]));
    }
    if (RL) {
        g_k[4] = cmx_div(R(-1), rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__tau_X_kr);
        g_inf[4] = rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__X_kr_inf;
    } else {
        dy[4] = cmx_div((rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__X_kr_inf - y[4]), rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__tau_X_kr);
    }
    if (RL) {
        g_k[5] = cmx_div(R(-1), slow_activating_delayed_rectifiyer_K_current_X_ks_gate__tau_X_ks);
        g_inf[5] = slow_activating_delayed_rectifiyer_K_current_X_ks_gate__X_ks_infinity;
    } else {
        dy[5] = cmx_div((slow_activating_delayed_rectifiyer_K_current_X_ks_gate__X_ks_infinity - y[5]), slow_activating_delayed_rectifiyer_K_current_X_ks_gate__tau_X_ks);
    }
    if (RL) {
        g_k[6] = ((transient_outward_potassium_current_X_to_gate__alpha_X_to * (-R(1))) - (transient_outward_potassium_current_X_to_gate__beta_X_to * R(1)));
        g_inf[6] = cmx_div(-((transient_outward_potassium_current_X_to_gate__alpha_X_to * R(1.0))), ((transient_outward_potassium_current_X_to_gate__alpha_X_to * (-R(1))) - (transient_outward_potassium_current_X_to_gate__beta_X_to * R(1))));
    } else {
        dy[6] = ((transient_outward_potassium_current_X_to_gate__alpha_X_to * (R(1.0) - y[6])) - (transient_outward_potassium_current_X_to_gate__beta_X_to * y[6]));
    }
    if (RL) {
        g_k[7] = ((transient_outward_potassium_current_Y_to_gate__alpha_Y_to * (-R(1))) - (transient_outward_potassium_current_Y_to_gate__beta_Y_to * R(1)));
        g_inf[7] = cmx_div(-((transient_outward_potassium_current_Y_to_gate__alpha_Y_to * R(1.0))), ((transient_outward_potassium_current_Y_to_gate__alpha_Y_to * (-R(1))) - (transient_outward_potassium_current_Y_to_gate__beta_Y_to * R(1))));
    } else {
        dy[7] = ((transient_outward_potassium_current_Y_to_gate__alpha_Y_to * (R(1.0) - y[7])) - (transient_outward_potassium_current_Y_to_gate__beta_Y_to * y[7]));
    }
    if (RL) {
        g_k[8] = cmx_div(R(-1), L_type_Ca_current_f_gate__tau_f);
        g_inf[8] = L_type_Ca_current_f_gate__f_infinity;
    } else {
        dy[8] = cmx_div((L_type_Ca_current_f_gate__f_infinity - y[8]), L_type_Ca_current_f_gate__tau_f);
    }
    if (RL) {
        g_k[9] = cmx_div(R(-1), L_type_Ca_current_d_gate__tau_d);
        g_inf[9] = L_type_Ca_current_d_gate__d_infinity;
    } else {
        dy[9] = cmx_div((L_type_Ca_current_d_gate__d_infinity - y[9]), L_type_Ca_current_d_gate__tau_d);
    }
    dy[10] = ((L_type_Ca_current_f_Ca_gate__f_Ca_infinity - y[10]) * (R(1) / L_type_Ca_current_f_Ca_gate__tau_f_Ca));
    dy[11] = (calcium_dynamics__beta_i * (((calcium_dynamics__J_rel + calcium_dynamics__J_leak) - calcium_dynamics__J_up) - (((calcium_dynamics__A_Cap * L_type_Ca_current__C_sc) * (R(1) / (R(2.0) * membrane__F * calcium_dynamics__V_myo))) * ((L_type_Ca_current__i_Ca + calcium_background_current__i_Ca_b + sarcolemmal_calcium_pump__i_p_Ca) - (R(2.0) * Na_Ca_exchanger__i_NaCa)))));
    dy[12] = ((calcium_dynamics__beta_SR * ((calcium_dynamics__J_up - calcium_dynamics__J_leak) - calcium_dynamics__J_rel) * calcium_dynamics__V_myo) * (R(1) / calcium_dynamics__V_SR));
}

// dV/dt including the stimulus (single-cell integration).
template <class R>
CMX_HD inline R dvdt(const R* y, R V, R i_stim_area, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8.314);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96.5);
    const R time_independent_potassium_current__K_mK1 = R(13.0);
    const R transient_outward_potassium_current__g_to = R(0.23815);
    const R plateau_potassium_current__g_Kp = R(0.002216);
    const R sodium_potassium_pump__i_NaK_max = R(0.693);
    const R sodium_potassium_pump__K_mNai = R(10.0);
    const R sodium_potassium_pump__K_mKo = R(1.5);
    const R Na_Ca_exchanger__K_mCa = R(1380.0);
    const R Na_Ca_exchanger__K_mNa = R(87.5);
    const R Na_Ca_exchanger__K_NaCa = R(1500.0);
    const R Na_Ca_exchanger__K_sat = R(0.2);
    const R Na_Ca_exchanger__eta = R(0.35);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.05);
    const R sarcolemmal_calcium_pump__i_pCa_max = R(0.05);
    const R calcium_background_current__g_Cab = R(0.0003842);
    const R sodium_background_current__g_Nab = R(0.0031);
    const R L_type_Ca_current__P_CaK = R(5.79e-07);
    const R L_type_Ca_current__i_Ca_half = R(-0.265);
    const R L_type_Ca_current__C_sc = R(1.0);
    const R standard_ionic_concentrations__Na_i = R(10.0);
    const R standard_ionic_concentrations__Na_o = R(138.0);
    const R standard_ionic_concentrations__Ca_o = R(2000.0);
    const R standard_ionic_concentrations__K_i = R(149.4);
    const R standard_ionic_concentrations__K_o = R(4.0);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_Ca_max = cmx_remove_singularity(V, R(-0.0), R(1.3354093264248706e-06), [&](const R V) { return cmx_div(((((p[4] * (R(1) / L_type_Ca_current__C_sc)) * R(4.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((y[11] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10803391981328421, 0.0)) - (R(0.341) * standard_ionic_concentrations__Ca_o))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current__i_Ca = (L_type_Ca_current__i_Ca_max * y[8] * y[9] * y[10]);
    const R L_type_Ca_current__i_CaK = cmx_div((((cmx_div(((L_type_Ca_current__P_CaK * (R(1) / L_type_Ca_current__C_sc)) * y[8] * y[9] * y[10]), (R(1.0) + (L_type_Ca_current__i_Ca_max * (R(1) / L_type_Ca_current__i_Ca_half)))) * R(1000.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((standard_ionic_concentrations__K_i * cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.054016959906642105, 0.0)) - standard_ionic_concentrations__K_o)), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T)))));
    const R Na_Ca_exchanger__i_NaCa = (cmx_div(Na_Ca_exchanger__K_NaCa, ((cmx_ipow<3>(Na_Ca_exchanger__K_mNa) + cmx_ipow<3>(standard_ionic_concentrations__Na_o)) * (Na_Ca_exchanger__K_mCa + standard_ionic_concentrations__Ca_o) * (R(1.0) + (Na_Ca_exchanger__K_sat * cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0))))) * ((cmx_exp_v(((Na_Ca_exchanger__eta * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.018905935967324736, 0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_i) * standard_ionic_concentrations__Ca_o) - (cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_o) * y[11])));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(standard_ionic_concentrations__Ca_o, y[11])));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R fast_sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log((standard_ionic_concentrations__Na_o * (R(1) / standard_ionic_concentrations__Na_i))));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[1]) * y[2] * y[3] * (V - fast_sodium_current__E_Na));
    const R plateau_potassium_current_Kp_gate__Kp_V = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(7.488) - V) * (R(1) / R(5.98))), V, -0.2412533513192246, 1.806505094678354)));
    const R rapid_activating_delayed_rectifiyer_K_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log((standard_ionic_concentrations__K_o * (R(1) / standard_ionic_concentrations__K_i))));
    const R plateau_potassium_current__i_Kp = (plateau_potassium_current__g_Kp * plateau_potassium_current_Kp_gate__Kp_V * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    const R rapid_activating_delayed_rectifiyer_K_current__R_V = cmx_div(R(1.0), (R(1.0) + (R(2.5) * cmx_exp_v((R(0.1) * (V + R(28.0))), V, 0.14426950408889636, 4.039546114489098))));
    const R rapid_activating_delayed_rectifiyer_K_current__i_Kr = (p[1] * rapid_activating_delayed_rectifiyer_K_current__R_V * y[4] * sqrt((standard_ionic_concentrations__K_o * (R(1) / R(4.0)))) * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__i_pCa_max * y[11]), (sarcolemmal_calcium_pump__K_mpCa + y[11]));
    const R slow_activating_delayed_rectifiyer_K_current__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(((standard_ionic_concentrations__K_o + (R(0.01833) * standard_ionic_concentrations__Na_o)) * (R(1) / (standard_ionic_concentrations__K_i + (R(0.01833) * standard_ionic_concentrations__Na_i))))));
    const R slow_activating_delayed_rectifiyer_K_current__i_Ks = (p[2] * cmx_ipow<2>(y[5]) * (V - slow_activating_delayed_rectifiyer_K_current__E_Ks));
    const R sodium_background_current__i_Na_b = (sodium_background_current__g_Nab * (V - fast_sodium_current__E_Na));
    const R sodium_potassium_pump__sigma = ((R(1.0) * (R(1) / R(7.0))) * cmx_expm1((standard_ionic_concentrations__Na_o * (R(1) / R(67.3)))));
    const R sodium_potassium_pump__f_NaK = cmx_div(R(1.0), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005401695990664211, -0.0)) + (R(0.0365) * sodium_potassium_pump__sigma * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.054016959906642105, -0.0))));
    const R sodium_potassium_pump__i_NaK = ((((sodium_potassium_pump__i_NaK_max * sodium_potassium_pump__f_NaK) * (R(1) / (R(1.0) + pow((sodium_potassium_pump__K_mNai * (R(1) / standard_ionic_concentrations__Na_i)), R(1.5))))) * standard_ionic_concentrations__K_o) * (R(1) / (standard_ionic_concentrations__K_o + sodium_potassium_pump__K_mKo)));
    const R time_independent_potassium_current_K1_gate__K1_infinity = cmx_div(R(1.0), (R(2.0) + cmx_exp_v((((R(1.62) * membrane__F) * (R(1) / (membrane__R * membrane__T))) * (V - rapid_activating_delayed_rectifiyer_K_current__E_K)), V, 0.08750747504876022, 8.461318867401037)));
    const R time_independent_potassium_current__i_K1 = (((p[3] * time_independent_potassium_current_K1_gate__K1_infinity * standard_ionic_concentrations__K_o) * (R(1) / (standard_ionic_concentrations__K_o + time_independent_potassium_current__K_mK1))) * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    const R transient_outward_potassium_current__i_to = (transient_outward_potassium_current__g_to * y[6] * y[7] * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    return (-(fast_sodium_current__i_Na + L_type_Ca_current__i_Ca + L_type_Ca_current__i_CaK + rapid_activating_delayed_rectifiyer_K_current__i_Kr + slow_activating_delayed_rectifiyer_K_current__i_Ks + transient_outward_potassium_current__i_to + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + Na_Ca_exchanger__i_NaCa + sodium_potassium_pump__i_NaK + sarcolemmal_calcium_pump__i_p_Ca + sodium_background_current__i_Na_b + calcium_background_current__i_Ca_b + i_stim));
}

// Total ionic current in uA/cm^2 (GetIIonic): minus dV/dt at zero stimulus,
// times the membrane capacitance that converts it to a current density.
template <class R>
CMX_HD inline R iionic(const R* y, R V, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8.314);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96.5);
    const R time_independent_potassium_current__K_mK1 = R(13.0);
    const R transient_outward_potassium_current__g_to = R(0.23815);
    const R plateau_potassium_current__g_Kp = R(0.002216);
    const R sodium_potassium_pump__i_NaK_max = R(0.693);
    const R sodium_potassium_pump__K_mNai = R(10.0);
    const R sodium_potassium_pump__K_mKo = R(1.5);
    const R Na_Ca_exchanger__K_mCa = R(1380.0);
    const R Na_Ca_exchanger__K_mNa = R(87.5);
    const R Na_Ca_exchanger__K_NaCa = R(1500.0);
    const R Na_Ca_exchanger__K_sat = R(0.2);
    const R Na_Ca_exchanger__eta = R(0.35);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.05);
    const R sarcolemmal_calcium_pump__i_pCa_max = R(0.05);
    const R calcium_background_current__g_Cab = R(0.0003842);
    const R sodium_background_current__g_Nab = R(0.0031);
    const R L_type_Ca_current__P_CaK = R(5.79e-07);
    const R L_type_Ca_current__i_Ca_half = R(-0.265);
    const R L_type_Ca_current__C_sc = R(1.0);
    const R standard_ionic_concentrations__Na_i = R(10.0);
    const R standard_ionic_concentrations__Na_o = R(138.0);
    const R standard_ionic_concentrations__Ca_o = R(2000.0);
    const R standard_ionic_concentrations__K_i = R(149.4);
    const R standard_ionic_concentrations__K_o = R(4.0);
    const R i_stim = R(0);
    (void)i_stim;
    const R L_type_Ca_current__i_Ca_max = cmx_remove_singularity(V, R(-0.0), R(1.3354093264248706e-06), [&](const R V) { return cmx_div(((((p[4] * (R(1) / L_type_Ca_current__C_sc)) * R(4.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((y[11] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10803391981328421, 0.0)) - (R(0.341) * standard_ionic_concentrations__Ca_o))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current__i_Ca = (L_type_Ca_current__i_Ca_max * y[8] * y[9] * y[10]);
    const R L_type_Ca_current__i_CaK = cmx_div((((cmx_div(((L_type_Ca_current__P_CaK * (R(1) / L_type_Ca_current__C_sc)) * y[8] * y[9] * y[10]), (R(1.0) + (L_type_Ca_current__i_Ca_max * (R(1) / L_type_Ca_current__i_Ca_half)))) * R(1000.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((standard_ionic_concentrations__K_i * cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.054016959906642105, 0.0)) - standard_ionic_concentrations__K_o)), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T)))));
    const R Na_Ca_exchanger__i_NaCa = (cmx_div(Na_Ca_exchanger__K_NaCa, ((cmx_ipow<3>(Na_Ca_exchanger__K_mNa) + cmx_ipow<3>(standard_ionic_concentrations__Na_o)) * (Na_Ca_exchanger__K_mCa + standard_ionic_concentrations__Ca_o) * (R(1.0) + (Na_Ca_exchanger__K_sat * cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0))))) * ((cmx_exp_v(((Na_Ca_exchanger__eta * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.018905935967324736, 0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_i) * standard_ionic_concentrations__Ca_o) - (cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_o) * y[11])));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(standard_ionic_concentrations__Ca_o, y[11])));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R fast_sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log((standard_ionic_concentrations__Na_o * (R(1) / standard_ionic_concentrations__Na_i))));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[1]) * y[2] * y[3] * (V - fast_sodium_current__E_Na));
    const R plateau_potassium_current_Kp_gate__Kp_V = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(7.488) - V) * (R(1) / R(5.98))), V, -0.2412533513192246, 1.806505094678354)));
    const R rapid_activating_delayed_rectifiyer_K_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log((standard_ionic_concentrations__K_o * (R(1) / standard_ionic_concentrations__K_i))));
    const R plateau_potassium_current__i_Kp = (plateau_potassium_current__g_Kp * plateau_potassium_current_Kp_gate__Kp_V * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    const R rapid_activating_delayed_rectifiyer_K_current__R_V = cmx_div(R(1.0), (R(1.0) + (R(2.5) * cmx_exp_v((R(0.1) * (V + R(28.0))), V, 0.14426950408889636, 4.039546114489098))));
    const R rapid_activating_delayed_rectifiyer_K_current__i_Kr = (p[1] * rapid_activating_delayed_rectifiyer_K_current__R_V * y[4] * sqrt((standard_ionic_concentrations__K_o * (R(1) / R(4.0)))) * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__i_pCa_max * y[11]), (sarcolemmal_calcium_pump__K_mpCa + y[11]));
    const R slow_activating_delayed_rectifiyer_K_current__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(((standard_ionic_concentrations__K_o + (R(0.01833) * standard_ionic_concentrations__Na_o)) * (R(1) / (standard_ionic_concentrations__K_i + (R(0.01833) * standard_ionic_concentrations__Na_i))))));
    const R slow_activating_delayed_rectifiyer_K_current__i_Ks = (p[2] * cmx_ipow<2>(y[5]) * (V - slow_activating_delayed_rectifiyer_K_current__E_Ks));
    const R sodium_background_current__i_Na_b = (sodium_background_current__g_Nab * (V - fast_sodium_current__E_Na));
    const R sodium_potassium_pump__sigma = ((R(1.0) * (R(1) / R(7.0))) * cmx_expm1((standard_ionic_concentrations__Na_o * (R(1) / R(67.3)))));
    const R sodium_potassium_pump__f_NaK = cmx_div(R(1.0), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005401695990664211, -0.0)) + (R(0.0365) * sodium_potassium_pump__sigma * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.054016959906642105, -0.0))));
    const R sodium_potassium_pump__i_NaK = ((((sodium_potassium_pump__i_NaK_max * sodium_potassium_pump__f_NaK) * (R(1) / (R(1.0) + pow((sodium_potassium_pump__K_mNai * (R(1) / standard_ionic_concentrations__Na_i)), R(1.5))))) * standard_ionic_concentrations__K_o) * (R(1) / (standard_ionic_concentrations__K_o + sodium_potassium_pump__K_mKo)));
    const R time_independent_potassium_current_K1_gate__K1_infinity = cmx_div(R(1.0), (R(2.0) + cmx_exp_v((((R(1.62) * membrane__F) * (R(1) / (membrane__R * membrane__T))) * (V - rapid_activating_delayed_rectifiyer_K_current__E_K)), V, 0.08750747504876022, 8.461318867401037)));
    const R time_independent_potassium_current__i_K1 = (((p[3] * time_independent_potassium_current_K1_gate__K1_infinity * standard_ionic_concentrations__K_o) * (R(1) / (standard_ionic_concentrations__K_o + time_independent_potassium_current__K_mK1))) * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    const R transient_outward_potassium_current__i_to = (transient_outward_potassium_current__g_to * y[6] * y[7] * (V - rapid_activating_delayed_rectifiyer_K_current__E_K));
    const R dv = (-(fast_sodium_current__i_Na + L_type_Ca_current__i_Ca + L_type_Ca_current__i_CaK + rapid_activating_delayed_rectifiyer_K_current__i_Kr + slow_activating_delayed_rectifiyer_K_current__i_Ks + transient_outward_potassium_current__i_to + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + Na_Ca_exchanger__i_NaCa + sodium_potassium_pump__i_NaK + sarcolemmal_calcium_pump__i_p_Ca + sodium_background_current__i_Na_b + calcium_background_current__i_Ca_b + i_stim));
    return -dv * cm;
}

// Backward Euler: the states solved by Newton (all but the linear gates).
constexpr int kNumNonlinear = 3;
static constexpr int kNonlinear[3] = {10, 11, 12};

// Linear states: y <- (y + a dt) / (1 - b dt) for dy/dt = a + b y, a and b
// functions of V (ComputeOneStepExceptVoltage, closed-form part).
template <class R>
CMX_HD inline void be_linear(R* y, R V, const R* p, R cm, R dt)
{
    CMX_MATH_USING
    (void)p; (void)cm; (void)V;
    const R L_type_Ca_current_d_gate__d_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(10.0)) * (R(1) / (-R(6.24)))), V, -0.23120112834759027, -2.312011283475903)));
    const R L_type_Ca_current_d_gate__E0_m = (V + R(40.0));
    const R L_type_Ca_current_d_gate__tau_d = cmx_div(R(1.0), (cmx_div((R(0.25) * cmx_exp_v(((-R(0.01)) * V), V, -0.014426950408889635, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.07)) * V), V, -0.10098865286222744, -0.0))) + cmx_div((R(0.07) * cmx_exp_c(((-R(0.05)) * L_type_Ca_current_d_gate__E0_m), (L_type_Ca_current_d_gate__E0_m), -0.07213475204444818)), (R(1.0) + cmx_exp_c((R(0.05) * L_type_Ca_current_d_gate__E0_m), (L_type_Ca_current_d_gate__E0_m), 0.07213475204444818)))));
    const R L_type_Ca_current_f_gate__f_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(12.5)) * (R(1) / R(5.0))), V, 0.2885390081777927, 3.6067376022224087)));
    const R L_type_Ca_current_f_gate__tau_f = (R(30.0) + cmx_div(R(200.0), (R(1.0) + cmx_exp_v(((V + R(20.0)) * (R(1) / R(9.5))), V, 0.15186263588304877, 3.0372527176609756))));
    const R fast_sodium_current_h_gate__alpha_h = (R(0.135) * cmx_exp_v(((V + R(80.0)) * (R(1) / (-R(6.8)))), V, -0.2121610354248476, -16.972882833987807));
    const R fast_sodium_current_h_gate__beta_h = cmx_div(R(7.5), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(11.0))), V, -0.14426950408889636, -1.5869645449778598)));
    const R fast_sodium_current_j_gate__alpha_j = cmx_div((R(0.175) * cmx_exp_v(((V + R(100.0)) * (R(1) / (-R(23.0)))), V, -0.06272587134299841, -6.27258713429984)), (R(1.0) + cmx_exp_v((R(0.15) * (V + R(79.0))), V, 0.21640425613334446, 17.095936234534214)));
    const R fast_sodium_current_j_gate__beta_j = cmx_div(R(0.3), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(32.0))), V, -0.14426950408889636, -4.616624130844683)));
    const R fast_sodium_current_m_gate__E0_m = (V + R(47.13));
    const R fast_sodium_current_m_gate__alpha_m = cmx_div((R(0.32) * fast_sodium_current_m_gate__E0_m), (-cmx_expm1(((-R(0.1)) * fast_sodium_current_m_gate__E0_m))));
    const R fast_sodium_current_m_gate__beta_m = (R(0.08) * cmx_exp_v(((-V) * (R(1) / R(11.0))), V, -0.1311540946262694, -0.0));
    const R rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__X_kr_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-R(2.182)) - (R(0.1819) * V)), V, -0.2624262279377025, -3.147960579219718)));
    const R rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__tau_X_kr = (R(43.0) + cmx_div(R(1.0), (cmx_exp_v(((-R(5.495)) + (R(0.1691) * V)), V, 0.2439597314143237, -7.927609249684854) + cmx_exp_v(((-R(7.677)) - (R(0.0128) * V)), V, -0.01846649652337874, -11.075569828904571))));
    const R slow_activating_delayed_rectifiyer_K_current_X_ks_gate__X_ks_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(16.0)) * (R(1) / (-R(13.6)))), V, -0.10608051771242379, 1.6972882833987806)));
    const R slow_activating_delayed_rectifiyer_K_current_X_ks_gate__tau_X_ks = cmx_remove_singularity(V, R(9.999999999999991), R(6.75675675675675e-07), [&](const R V) { return cmx_div(R(1.0), (cmx_div((R(7.19e-05) * (V - R(10.0))), (-cmx_expm1(((-R(0.148)) * (V - R(10.0)))))) + cmx_div((R(0.000131) * (V - R(10.0))), cmx_expm1((R(0.0687) * (V - R(10.0))))))); });
    const R transient_outward_potassium_current_X_to_gate__alpha_X_to = (R(0.04516) * cmx_exp_v((R(0.03577) * V), V, 0.051605201612598225, 0.0));
    const R transient_outward_potassium_current_X_to_gate__beta_X_to = (R(0.0989) * cmx_exp_v(((-R(0.06237)) * V), V, -0.08998088970024463, -0.0));
    const R transient_outward_potassium_current_Y_to_gate__alpha_Y_to = cmx_div((R(0.005415) * cmx_exp_v(((V + R(33.5)) * (R(1) / (-R(5.0)))), V, -0.2885390081777927, -9.666056773956054)), (R(1.0) + (R(0.051335) * cmx_exp_v(((V + R(33.5)) * (R(1) / (-R(5.0)))), V, -0.2885390081777927, -9.666056773956054))));
    const R transient_outward_potassium_current_Y_to_gate__beta_Y_to = cmx_div((R(0.005415) * cmx_exp_v(((V + R(33.5)) * (R(1) / R(5.0))), V, 0.2885390081777927, 9.666056773956054)), (R(1.0) + (R(0.051335) * cmx_exp_v(((V + R(33.5)) * (R(1) / R(5.0))), V, 0.2885390081777927, 9.666056773956054))));
    y[1] = (y[1] + ((fast_sodium_current_m_gate__alpha_m * R(1.0))) * dt) / (R(1) - (((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1)))) * dt);
    y[2] = (y[2] + ((fast_sodium_current_h_gate__alpha_h * R(1.0))) * dt) / (R(1) - (((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1)))) * dt);
    y[3] = (y[3] + ((fast_sodium_current_j_gate__alpha_j * R(1.0))) * dt) / (R(1) - (((fast_sodium_current_j_gate__alpha_j * (-R(1))) - (fast_sodium_current_j_gate__beta_j * R(1)))) * dt);
    y[4] = (y[4] + ((rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__X_kr_inf / rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__tau_X_kr)) * dt) / (R(1) - (((-R(1)) / rapid_activating_delayed_rectifiyer_K_current_X_kr_gate__tau_X_kr)) * dt);
    y[5] = (y[5] + ((slow_activating_delayed_rectifiyer_K_current_X_ks_gate__X_ks_infinity / slow_activating_delayed_rectifiyer_K_current_X_ks_gate__tau_X_ks)) * dt) / (R(1) - (((-R(1)) / slow_activating_delayed_rectifiyer_K_current_X_ks_gate__tau_X_ks)) * dt);
    y[6] = (y[6] + ((transient_outward_potassium_current_X_to_gate__alpha_X_to * R(1.0))) * dt) / (R(1) - (((transient_outward_potassium_current_X_to_gate__alpha_X_to * (-R(1))) - (transient_outward_potassium_current_X_to_gate__beta_X_to * R(1)))) * dt);
    y[7] = (y[7] + ((transient_outward_potassium_current_Y_to_gate__alpha_Y_to * R(1.0))) * dt) / (R(1) - (((transient_outward_potassium_current_Y_to_gate__alpha_Y_to * (-R(1))) - (transient_outward_potassium_current_Y_to_gate__beta_Y_to * R(1)))) * dt);
    y[8] = (y[8] + ((L_type_Ca_current_f_gate__f_infinity / L_type_Ca_current_f_gate__tau_f)) * dt) / (R(1) - (((-R(1)) / L_type_Ca_current_f_gate__tau_f)) * dt);
    y[9] = (y[9] + ((L_type_Ca_current_d_gate__d_infinity / L_type_Ca_current_d_gate__tau_d)) * dt) / (R(1) - (((-R(1)) / L_type_Ca_current_d_gate__tau_d)) * dt);
}

// Newton residual of the nonlinear states at guess g (y holds the other
// states, linear ones already advanced): g - y_old - dt f(g).
template <class R>
CMX_HD inline void be_residual(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* res)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8.314);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96.5);
    const R Na_Ca_exchanger__K_mCa = R(1380.0);
    const R Na_Ca_exchanger__K_mNa = R(87.5);
    const R Na_Ca_exchanger__K_NaCa = R(1500.0);
    const R Na_Ca_exchanger__K_sat = R(0.2);
    const R Na_Ca_exchanger__eta = R(0.35);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.05);
    const R sarcolemmal_calcium_pump__i_pCa_max = R(0.05);
    const R calcium_background_current__g_Cab = R(0.0003842);
    const R L_type_Ca_current__C_sc = R(1.0);
    const R L_type_Ca_current_f_Ca_gate__K_mfCa = R(0.18);
    const R calcium_dynamics__K_mCMDN = R(2.0);
    const R calcium_dynamics__CMDN_tot = R(10.0);
    const R calcium_dynamics__V_myo = R(2.584e-05);
    const R calcium_dynamics__A_Cap = R(0.0001534);
    const R calcium_dynamics__P_rel = R(6.0);
    const R calcium_dynamics__P_leak = R(1e-06);
    const R calcium_dynamics__K_mCSQN = R(600.0);
    const R calcium_dynamics__CSQN_tot = R(10000.0);
    const R calcium_dynamics__V_SR = R(2e-06);
    const R calcium_dynamics__V_up = R(0.1);
    const R calcium_dynamics__K_mup = R(0.32);
    const R standard_ionic_concentrations__Na_i = R(10.0);
    const R standard_ionic_concentrations__Na_o = R(138.0);
    const R standard_ionic_concentrations__Ca_o = R(2000.0);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_Ca_max = cmx_remove_singularity(V, R(-0.0), R(1.3354093264248706e-06), [&](const R V) { return cmx_div(((((p[4] * (R(1) / L_type_Ca_current__C_sc)) * R(4.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((g[1] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10803391981328421, 0.0)) - (R(0.341) * standard_ionic_concentrations__Ca_o))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current__i_Ca = (L_type_Ca_current__i_Ca_max * y[8] * y[9] * g[0]);
    const R L_type_Ca_current_f_Ca_gate__f_Ca_infinity = cmx_div(R(1.0), (R(1.0) + cmx_ipow<3>((g[1] * (R(1) / L_type_Ca_current_f_Ca_gate__K_mfCa)))));
    const R L_type_Ca_current_f_Ca_gate__tau_f_Ca = R(30.0);
    const R Na_Ca_exchanger__i_NaCa = (cmx_div(Na_Ca_exchanger__K_NaCa, ((cmx_ipow<3>(Na_Ca_exchanger__K_mNa) + cmx_ipow<3>(standard_ionic_concentrations__Na_o)) * (Na_Ca_exchanger__K_mCa + standard_ionic_concentrations__Ca_o) * (R(1.0) + (Na_Ca_exchanger__K_sat * cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0))))) * ((cmx_exp_v(((Na_Ca_exchanger__eta * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.018905935967324736, 0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_i) * standard_ionic_concentrations__Ca_o) - (cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_o) * g[1])));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(standard_ionic_concentrations__Ca_o, g[1])));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R calcium_dynamics__J_leak = (calcium_dynamics__P_leak * (g[2] - g[1]));
    const R calcium_dynamics__gamma = cmx_div(R(1.0), (R(1.0) + cmx_ipow<3>(cmx_div(R(2000.0), g[2]))));
    const R calcium_dynamics__J_rel = cmx_div((calcium_dynamics__P_rel * y[8] * y[9] * g[0] * ((calcium_dynamics__gamma * g[2]) - g[1])), (R(1.0) + (R(1.65) * cmx_exp_v((V * (R(1) / R(20.0))), V, 0.07213475204444818, 0.0))));
    const R calcium_dynamics__J_up = cmx_div(calcium_dynamics__V_up, (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__K_mup, g[1]))));
    const R calcium_dynamics__beta_SR = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_tot * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[2])))));
    const R calcium_dynamics__beta_i = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_tot * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[1])))));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__i_pCa_max * g[1]), (sarcolemmal_calcium_pump__K_mpCa + g[1]));
    res[0] = g[0] - y[10] - dt * ((L_type_Ca_current_f_Ca_gate__f_Ca_infinity - g[0]) * (R(1) / L_type_Ca_current_f_Ca_gate__tau_f_Ca));
    res[1] = g[1] - y[11] - dt * (calcium_dynamics__beta_i * (((calcium_dynamics__J_rel + calcium_dynamics__J_leak) - calcium_dynamics__J_up) - (((calcium_dynamics__A_Cap * L_type_Ca_current__C_sc) * (R(1) / (R(2.0) * membrane__F * calcium_dynamics__V_myo))) * ((L_type_Ca_current__i_Ca + calcium_background_current__i_Ca_b + sarcolemmal_calcium_pump__i_p_Ca) - (R(2.0) * Na_Ca_exchanger__i_NaCa)))));
    res[2] = g[2] - y[12] - dt * ((calcium_dynamics__beta_SR * ((calcium_dynamics__J_up - calcium_dynamics__J_leak) - calcium_dynamics__J_rel) * calcium_dynamics__V_myo) * (R(1) / calcium_dynamics__V_SR));
}

// Jacobian of be_residual in g: J[i*N + j] = delta_ij - dt d f_i / d g_j.
template <class R>
CMX_HD inline void be_jacobian(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* J)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8.314);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96.5);
    const R Na_Ca_exchanger__K_mCa = R(1380.0);
    const R Na_Ca_exchanger__K_mNa = R(87.5);
    const R Na_Ca_exchanger__K_NaCa = R(1500.0);
    const R Na_Ca_exchanger__K_sat = R(0.2);
    const R Na_Ca_exchanger__eta = R(0.35);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.05);
    const R sarcolemmal_calcium_pump__i_pCa_max = R(0.05);
    const R calcium_background_current__g_Cab = R(0.0003842);
    const R L_type_Ca_current__C_sc = R(1.0);
    const R L_type_Ca_current_f_Ca_gate__K_mfCa = R(0.18);
    const R calcium_dynamics__K_mCMDN = R(2.0);
    const R calcium_dynamics__CMDN_tot = R(10.0);
    const R calcium_dynamics__V_myo = R(2.584e-05);
    const R calcium_dynamics__A_Cap = R(0.0001534);
    const R calcium_dynamics__P_rel = R(6.0);
    const R calcium_dynamics__P_leak = R(1e-06);
    const R calcium_dynamics__K_mCSQN = R(600.0);
    const R calcium_dynamics__CSQN_tot = R(10000.0);
    const R calcium_dynamics__V_SR = R(2e-06);
    const R calcium_dynamics__V_up = R(0.1);
    const R calcium_dynamics__K_mup = R(0.32);
    const R standard_ionic_concentrations__Na_i = R(10.0);
    const R standard_ionic_concentrations__Na_o = R(138.0);
    const R standard_ionic_concentrations__Ca_o = R(2000.0);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_Ca_max = cmx_remove_singularity(V, R(-0.0), R(1.3354093264248706e-06), [&](const R V) { return cmx_div(((((p[4] * (R(1) / L_type_Ca_current__C_sc)) * R(4.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((g[1] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10803391981328421, 0.0)) - (R(0.341) * standard_ionic_concentrations__Ca_o))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current__i_Ca = (L_type_Ca_current__i_Ca_max * y[8] * y[9] * g[0]);
    const R L_type_Ca_current_f_Ca_gate__f_Ca_infinity = cmx_div(R(1.0), (R(1.0) + cmx_ipow<3>((g[1] * (R(1) / L_type_Ca_current_f_Ca_gate__K_mfCa)))));
    const R L_type_Ca_current_f_Ca_gate__tau_f_Ca = R(30.0);
    const R Na_Ca_exchanger__i_NaCa = (cmx_div(Na_Ca_exchanger__K_NaCa, ((cmx_ipow<3>(Na_Ca_exchanger__K_mNa) + cmx_ipow<3>(standard_ionic_concentrations__Na_o)) * (Na_Ca_exchanger__K_mCa + standard_ionic_concentrations__Ca_o) * (R(1.0) + (Na_Ca_exchanger__K_sat * cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0))))) * ((cmx_exp_v(((Na_Ca_exchanger__eta * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.018905935967324736, 0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_i) * standard_ionic_concentrations__Ca_o) - (cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_o) * g[1])));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(standard_ionic_concentrations__Ca_o, g[1])));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R calcium_dynamics__J_leak = (calcium_dynamics__P_leak * (g[2] - g[1]));
    const R calcium_dynamics__gamma = cmx_div(R(1.0), (R(1.0) + cmx_ipow<3>(cmx_div(R(2000.0), g[2]))));
    const R calcium_dynamics__J_rel = cmx_div((calcium_dynamics__P_rel * y[8] * y[9] * g[0] * ((calcium_dynamics__gamma * g[2]) - g[1])), (R(1.0) + (R(1.65) * cmx_exp_v((V * (R(1) / R(20.0))), V, 0.07213475204444818, 0.0))));
    const R calcium_dynamics__J_up = cmx_div(calcium_dynamics__V_up, (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__K_mup, g[1]))));
    const R calcium_dynamics__beta_SR = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_tot * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[2])))));
    const R calcium_dynamics__beta_i = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_tot * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[1])))));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__i_pCa_max * g[1]), (sarcolemmal_calcium_pump__K_mpCa + g[1]));
    const R d_L_type_Ca_current__i_Ca_max__d1 = cmx_remove_singularity(V, R(-0.0), R(1.3354093264248706e-06), [&](const R V) { return cmx_div(((((R(1) * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10803391981328421, 0.0))) * (((p[4] * (R(1) / L_type_Ca_current__C_sc)) * R(4.0) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_L_type_Ca_current__i_Ca__d0 = ((R(1) * L_type_Ca_current__i_Ca_max * y[8] * y[9]));
    const R d_L_type_Ca_current__i_Ca__d1 = ((d_L_type_Ca_current__i_Ca_max__d1 * y[8] * y[9] * g[0]));
    const R d_L_type_Ca_current_f_Ca_gate__f_Ca_infinity__d1 = cmx_div((-(R(1.0) * ((R(3.0) * cmx_ipow<2>((g[1] * (R(1) / L_type_Ca_current_f_Ca_gate__K_mfCa))) * cmx_div(R(1), L_type_Ca_current_f_Ca_gate__K_mfCa))))), ((R(1.0) + cmx_ipow<3>((g[1] * (R(1) / L_type_Ca_current_f_Ca_gate__K_mfCa)))) * (R(1.0) + cmx_ipow<3>((g[1] * (R(1) / L_type_Ca_current_f_Ca_gate__K_mfCa))))));
    const R d_Na_Ca_exchanger__i_NaCa__d1 = (((-((R(1) * cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0) * cmx_ipow<3>(standard_ionic_concentrations__Na_o)))) * cmx_div(Na_Ca_exchanger__K_NaCa, ((cmx_ipow<3>(Na_Ca_exchanger__K_mNa) + cmx_ipow<3>(standard_ionic_concentrations__Na_o)) * (Na_Ca_exchanger__K_mCa + standard_ionic_concentrations__Ca_o) * (R(1.0) + (Na_Ca_exchanger__K_sat * cmx_exp_v((((Na_Ca_exchanger__eta - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.035111023939317365, -0.0)))))));
    const R d_calcium_background_current__E_Ca__d1 = ((cmx_div(cmx_div((-(standard_ionic_concentrations__Ca_o * R(1))), (g[1] * g[1])), cmx_div(standard_ionic_concentrations__Ca_o, g[1])) * ((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F)))));
    const R d_calcium_background_current__i_Ca_b__d1 = (((-d_calcium_background_current__E_Ca__d1) * calcium_background_current__g_Cab));
    const R d_calcium_dynamics__J_leak__d1 = (((-R(1)) * calcium_dynamics__P_leak));
    const R d_calcium_dynamics__J_leak__d2 = ((R(1) * calcium_dynamics__P_leak));
    const R d_calcium_dynamics__gamma__d2 = cmx_div((-(R(1.0) * ((R(3.0) * cmx_ipow<2>(cmx_div(R(2000.0), g[2])) * cmx_div((-(R(2000.0) * R(1))), (g[2] * g[2])))))), ((R(1.0) + cmx_ipow<3>(cmx_div(R(2000.0), g[2]))) * (R(1.0) + cmx_ipow<3>(cmx_div(R(2000.0), g[2])))));
    const R d_calcium_dynamics__J_rel__d0 = cmx_div(((R(1) * calcium_dynamics__P_rel * y[8] * y[9] * ((calcium_dynamics__gamma * g[2]) - g[1]))), (R(1.0) + (R(1.65) * cmx_exp_v((V * (R(1) / R(20.0))), V, 0.07213475204444818, 0.0))));
    const R d_calcium_dynamics__J_rel__d1 = cmx_div((((-R(1)) * calcium_dynamics__P_rel * y[8] * y[9] * g[0])), (R(1.0) + (R(1.65) * cmx_exp_v((V * (R(1) / R(20.0))), V, 0.07213475204444818, 0.0))));
    const R d_calcium_dynamics__J_rel__d2 = cmx_div(((((d_calcium_dynamics__gamma__d2 * g[2]) + (R(1) * calcium_dynamics__gamma)) * calcium_dynamics__P_rel * y[8] * y[9] * g[0])), (R(1.0) + (R(1.65) * cmx_exp_v((V * (R(1) / R(20.0))), V, 0.07213475204444818, 0.0))));
    const R d_calcium_dynamics__J_up__d1 = cmx_div((-(calcium_dynamics__V_up * ((R(2.0) * cmx_ipow<1>(cmx_div(calcium_dynamics__K_mup, g[1])) * cmx_div((-(calcium_dynamics__K_mup * R(1))), (g[1] * g[1])))))), ((R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__K_mup, g[1]))) * (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__K_mup, g[1])))));
    const R d_calcium_dynamics__beta_SR__d2 = cmx_div((-(R(1.0) * (cmx_div((-((calcium_dynamics__CSQN_tot * calcium_dynamics__K_mCSQN) * (R(2.0) * cmx_ipow<1>((calcium_dynamics__K_mCSQN + g[2])) * (R(1))))), (cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[2])) * cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[2]))))))), ((R(1.0) + cmx_div((calcium_dynamics__CSQN_tot * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[2])))) * (R(1.0) + cmx_div((calcium_dynamics__CSQN_tot * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[2]))))));
    const R d_calcium_dynamics__beta_i__d1 = cmx_div((-(R(1.0) * (cmx_div((-((calcium_dynamics__CMDN_tot * calcium_dynamics__K_mCMDN) * (R(2.0) * cmx_ipow<1>((calcium_dynamics__K_mCMDN + g[1])) * (R(1))))), (cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[1])) * cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[1]))))))), ((R(1.0) + cmx_div((calcium_dynamics__CMDN_tot * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[1])))) * (R(1.0) + cmx_div((calcium_dynamics__CMDN_tot * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[1]))))));
    const R d_sarcolemmal_calcium_pump__i_p_Ca__d1 = cmx_div((((R(1) * sarcolemmal_calcium_pump__i_pCa_max)) * (sarcolemmal_calcium_pump__K_mpCa + g[1]) - (sarcolemmal_calcium_pump__i_pCa_max * g[1]) * (R(1))), ((sarcolemmal_calcium_pump__K_mpCa + g[1]) * (sarcolemmal_calcium_pump__K_mpCa + g[1])));
    J[0] = R(1) - dt * cmx_div((-R(1)), L_type_Ca_current_f_Ca_gate__tau_f_Ca);
    J[1] = R(0) - dt * cmx_div(d_L_type_Ca_current_f_Ca_gate__f_Ca_infinity__d1, L_type_Ca_current_f_Ca_gate__tau_f_Ca);
    J[2] = R(0);
    J[3] = R(0) - dt * ((((d_calcium_dynamics__J_rel__d0) - (((d_L_type_Ca_current__i_Ca__d0) * ((calcium_dynamics__A_Cap * L_type_Ca_current__C_sc) * (R(1) / (R(2.0) * membrane__F * calcium_dynamics__V_myo)))))) * calcium_dynamics__beta_i));
    J[4] = R(1) - dt * ((d_calcium_dynamics__beta_i__d1 * (((calcium_dynamics__J_rel + calcium_dynamics__J_leak) - calcium_dynamics__J_up) - (((calcium_dynamics__A_Cap * L_type_Ca_current__C_sc) * (R(1) / (R(2.0) * membrane__F * calcium_dynamics__V_myo))) * ((L_type_Ca_current__i_Ca + calcium_background_current__i_Ca_b + sarcolemmal_calcium_pump__i_p_Ca) - (R(2.0) * Na_Ca_exchanger__i_NaCa))))) + ((((d_calcium_dynamics__J_rel__d1 + d_calcium_dynamics__J_leak__d1) - d_calcium_dynamics__J_up__d1) - ((((d_L_type_Ca_current__i_Ca__d1 + d_calcium_background_current__i_Ca_b__d1 + d_sarcolemmal_calcium_pump__i_p_Ca__d1) - ((d_Na_Ca_exchanger__i_NaCa__d1 * R(2.0)))) * ((calcium_dynamics__A_Cap * L_type_Ca_current__C_sc) * (R(1) / (R(2.0) * membrane__F * calcium_dynamics__V_myo)))))) * calcium_dynamics__beta_i));
    J[5] = R(0) - dt * (((d_calcium_dynamics__J_rel__d2 + d_calcium_dynamics__J_leak__d2) * calcium_dynamics__beta_i));
    J[6] = R(0) - dt * cmx_div((((-d_calcium_dynamics__J_rel__d0) * calcium_dynamics__beta_SR * calcium_dynamics__V_myo)), calcium_dynamics__V_SR);
    J[7] = R(0) - dt * cmx_div(((((d_calcium_dynamics__J_up__d1 - d_calcium_dynamics__J_leak__d1) - d_calcium_dynamics__J_rel__d1) * calcium_dynamics__beta_SR * calcium_dynamics__V_myo)), calcium_dynamics__V_SR);
    J[8] = R(1) - dt * cmx_div(((d_calcium_dynamics__beta_SR__d2 * ((calcium_dynamics__J_up - calcium_dynamics__J_leak) - calcium_dynamics__J_rel) * calcium_dynamics__V_myo) + (((-d_calcium_dynamics__J_leak__d2) - d_calcium_dynamics__J_rel__d2) * calcium_dynamics__beta_SR * calcium_dynamics__V_myo)), calcium_dynamics__V_SR);
}

// One backward-Euler step of every state but V at the given V
// (ComputeOneStepExceptVoltage): linear states in closed form, then the
// nonlinear ones by Newton from their current values.  Returns the Newton
// iteration count, -1 if it did not converge.
template <class R>
CMX_HD inline int be_step(R* y, R V, R i_stim_area, const R* p, R cm, R dt)
{
    be_linear<R>(y, V, p, cm, dt);
    R g[kNumNonlinear];
CMX_UNROLL
    for (int i = 0; i < kNumNonlinear; ++i) g[i] = y[kNonlinear[i]];
    const int its = cmx_newton<kNumNonlinear, R>(
        g, [&](const R* x, R* r) { be_residual<R>(y, x, V, i_stim_area, p, cm, dt, r); },
        [&](const R* x, R* J) { be_jacobian<R>(y, x, V, i_stim_area, p, cm, dt, J); });
CMX_UNROLL
    for (int i = 0; i < kNumNonlinear; ++i) y[kNonlinear[i]] = g[i];
    return its;
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

}}  // namespace cmx::fox2002
