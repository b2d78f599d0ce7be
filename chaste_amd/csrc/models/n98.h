// Generated by chaste_amd/codegen/cellml2cpp.py from NobleVargheseKohlNoble1998a.cellml -- do not edit.
// Model: noble_varghese_kohl_noble_1998_basic
#pragma once
#include "model_common.h"

namespace cmx { namespace n98 {

constexpr int kNumStates = 22;
constexpr int kVIndex = 0;
constexpr int kNumParams = 5;
constexpr bool kPerCapacitanceUnits = true;

// State variables (CellML component__variable), V first then document order.
static const char* const kStateNames[kNumStates] = {
    "membrane__V",
    "rapid_delayed_rectifier_potassium_current_xr1_gate__xr1",
    "rapid_delayed_rectifier_potassium_current_xr2_gate__xr2",
    "slow_delayed_rectifier_potassium_current_xs_gate__xs",
    "fast_sodium_current_m_gate__m",
    "fast_sodium_current_h_gate__h",
    "L_type_Ca_channel_d_gate__d",
    "L_type_Ca_channel_f_gate__f",
    "L_type_Ca_channel_f2_gate__f2",
    "L_type_Ca_channel_f2ds_gate__f2ds",
    "transient_outward_current_s_gate__s",
    "transient_outward_current_r_gate__r",
    "calcium_release__ActFrac",
    "calcium_release__ProdFrac",
    "intracellular_sodium_concentration__Na_i",
    "intracellular_potassium_concentration__K_i",
    "intracellular_calcium_concentration__Ca_i",
    "intracellular_calcium_concentration__Ca_ds",
    "intracellular_calcium_concentration__Ca_up",
    "intracellular_calcium_concentration__Ca_rel",
    "intracellular_calcium_concentration__Ca_Calmod",
    "intracellular_calcium_concentration__Ca_Trop",
};
static constexpr double kInitialState[kNumStates] = {
    -92.849333,  // membrane__V
    1.03e-05,  // rapid_delayed_rectifier_potassium_current_xr1_gate__xr1
    2e-07,  // rapid_delayed_rectifier_potassium_current_xr2_gate__xr2
    0.001302,  // slow_delayed_rectifier_potassium_current_xs_gate__xs
    0.0016203,  // fast_sodium_current_m_gate__m
    0.9944036,  // fast_sodium_current_h_gate__h
    0.0,  // L_type_Ca_channel_d_gate__d
    1.0,  // L_type_Ca_channel_f_gate__f
    0.9349197,  // L_type_Ca_channel_f2_gate__f2
    0.9651958,  // L_type_Ca_channel_f2ds_gate__f2ds
    0.9948645,  // transient_outward_current_s_gate__s
    0.0,  // transient_outward_current_r_gate__r
    0.0042614,  // calcium_release__ActFrac
    0.4068154,  // calcium_release__ProdFrac
    7.3321223,  // intracellular_sodium_concentration__Na_i
    136.5644281,  // intracellular_potassium_concentration__K_i
    1.4e-05,  // intracellular_calcium_concentration__Ca_i
    1.88e-05,  // intracellular_calcium_concentration__Ca_ds
    0.4531889,  // intracellular_calcium_concentration__Ca_up
    0.4481927,  // intracellular_calcium_concentration__Ca_rel
    0.0005555,  // intracellular_calcium_concentration__Ca_Calmod
    0.0003542,  // intracellular_calcium_concentration__Ca_Trop
};
static const char* const kParamNames[5] = {
    "fast_sodium_current__g_Na",
    "time_independent_potassium_current__g_K1",
    "L_type_Ca_channel__P_Ca_L",
    "slow_delayed_rectifier_potassium_current__g_Ks",
    "rapid_delayed_rectifier_potassium_current__g_Kr1",
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
    2.5,
    0.5,
    0.1,
    0.0026,
    0.0021,
};
// 1 = integrated by Rush-Larsen (affine gate, V-dependent coefficients only)
static constexpr int kIsGate[kNumStates] = {0, 1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

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
    const R membrane__R = R(8314.472);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.3415);
    const R reversal_potentials__P_kna = R(0.03);
    const R time_independent_potassium_current__K_mk1 = R(10.0);
    const R rapid_delayed_rectifier_potassium_current__g_Kr2 = R(0.0013);
    const R fast_sodium_current_m_gate__delta_m = R(1e-05);
    const R fast_sodium_current_h_gate__shift_h = R(0.0);
    const R persistent_sodium_current__g_pna = R(0.004);
    const R sodium_background_current__g_bna = R(0.0006);
    const R L_type_Ca_channel__P_CaK = R(0.002);
    const R L_type_Ca_channel__P_CaNa = R(0.01);
    const R L_type_Ca_channel__Km_f2 = R(100000.0);
    const R L_type_Ca_channel__Km_f2ds = R(0.001);
    const R L_type_Ca_channel__R_decay = R(20.0);
    const R L_type_Ca_channel__FrICa = R(1.0);
    const R L_type_Ca_channel_d_gate__speed_d = R(3.0);
    const R L_type_Ca_channel_f_gate__speed_f = R(0.3);
    const R L_type_Ca_channel_f_gate__delta_f = R(0.0001);
    const R calcium_background_current__g_bca = R(0.00025);
    const R transient_outward_current__g_to = R(0.005);
    const R transient_outward_current__g_tos = R(0.0);
    const R sodium_potassium_pump__i_NaK_max = R(0.7);
    const R sodium_potassium_pump__K_mK = R(1.0);
    const R sodium_potassium_pump__K_mNa = R(40.0);
    const R sodium_calcium_exchanger__k_NaCa = R(0.0005);
    const R sodium_calcium_exchanger__n_NaCa = R(3.0);
    const R sodium_calcium_exchanger__d_NaCa = R(0.0);
    const R sodium_calcium_exchanger__gamma = R(0.5);
    const R sodium_calcium_exchanger__FRiNaCa = R(0.001);
    const R sarcoplasmic_reticulum_calcium_pump__K_cyca = R(0.0003);
    const R sarcoplasmic_reticulum_calcium_pump__K_xcs = R(0.4);
    const R sarcoplasmic_reticulum_calcium_pump__K_srca = R(0.5);
    const R sarcoplasmic_reticulum_calcium_pump__alpha_up = R(0.4);
    const R sarcoplasmic_reticulum_calcium_pump__beta_up = R(0.03);
    const R calcium_release__K_leak_rate = R(0.05);
    const R calcium_release__K_m_rel = R(250.0);
    const R calcium_release__K_m_Ca_cyt = R(0.0005);
    const R calcium_release__K_m_Ca_ds = R(0.01);
    const R extracellular_sodium_concentration__Na_o = R(140.0);
    const R extracellular_calcium_concentration__Ca_o = R(2.0);
    const R extracellular_potassium_concentration__K_o = R(4.0);
    const R intracellular_calcium_concentration__Calmod = R(0.02);
    const R intracellular_calcium_concentration__Trop = R(0.05);
    const R intracellular_calcium_concentration__alpha_Calmod = R(100000.0);
    const R intracellular_calcium_concentration__beta_Calmod = R(50.0);
    const R intracellular_calcium_concentration__alpha_Trop = R(100000.0);
    const R intracellular_calcium_concentration__beta_Trop = R(200.0);
    const R intracellular_calcium_concentration__radius = R(0.012);
    const R intracellular_calcium_concentration__length = R(0.074);
    const R intracellular_calcium_concentration__V_ds_ratio = R(0.1);
    const R intracellular_calcium_concentration__V_rel_ratio = R(0.1);
    const R intracellular_calcium_concentration__V_e_ratio = R(0.4);
    const R intracellular_calcium_concentration__V_up_ratio = R(0.01);
    const R intracellular_calcium_concentration__Kdecay = R(10.0);
    const R i_stim = (i_stim_area / cm) * R(0.09499999999999999);
    (void)i_stim;
    const R L_type_Ca_channel__i_Ca_L_Ca_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * R(4.0) * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((y[16] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_Ca_ds = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * R(4.0) * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((y[16] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_K_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[15] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_K_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[15] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[14] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[14] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel_d_gate__E0_d = ((V + R(24.0)) - R(5.0));
    const R L_type_Ca_channel_d_gate__alpha_d = (((fabs(L_type_Ca_channel_d_gate__E0_d) < R(0.0001))) ? (R(120.0)) : (cmx_div((R(30.0) * L_type_Ca_channel_d_gate__E0_d), (-cmx_expm1(((-L_type_Ca_channel_d_gate__E0_d) * (R(1) / R(4.0))))))));
    const R L_type_Ca_channel_d_gate__beta_d = (((fabs(L_type_Ca_channel_d_gate__E0_d) < R(0.0001))) ? (R(120.0)) : (cmx_div((R(12.0) * L_type_Ca_channel_d_gate__E0_d), cmx_expm1((L_type_Ca_channel_d_gate__E0_d * (R(1) / R(10.0)))))));
    const R L_type_Ca_channel_f_gate__E0_f = (V + R(34.0));
    const R L_type_Ca_channel_f_gate__alpha_f = (((fabs(L_type_Ca_channel_f_gate__E0_f) < L_type_Ca_channel_f_gate__delta_f)) ? (R(25.0)) : (cmx_div((R(6.25) * L_type_Ca_channel_f_gate__E0_f), cmx_expm1((L_type_Ca_channel_f_gate__E0_f * (R(1) / R(4.0)))))));
    const R L_type_Ca_channel_f_gate__beta_f = cmx_div(R(12.0), (R(1.0) + cmx_exp_v((((-R(1.0)) * (V + R(34.0))) * (R(1) / R(4.0))), V, -0.36067376022224085, -12.262907847556189)));
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_calcium_concentration__Ca_o, y[16])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_release__CadsReg = cmx_div(y[17], (y[17] + calcium_release__K_m_Ca_ds));
    const R calcium_release__CaiReg = cmx_div(y[16], (y[16] + calcium_release__K_m_Ca_cyt));
    const R calcium_release__RegBindSite = (calcium_release__CaiReg + ((R(1.0) - calcium_release__CaiReg) * calcium_release__CadsReg));
    const R calcium_release__VoltDep = cmx_exp_v((R(0.08) * (V - R(40.0))), V, 0.11541560327111708, -4.616624130844683);
    const R calcium_release__ActRate = ((R(0.0) * calcium_release__VoltDep) + (R(500.0) * cmx_ipow<2>(calcium_release__RegBindSite)));
    const R calcium_release__InactRate = (R(60.0) + (R(500.0) * cmx_ipow<2>(calcium_release__RegBindSite)));
    const R calcium_release__PrecFrac = ((R(1.0) - y[12]) - y[13]);
    const R calcium_release__SpeedRel = (((V < (-R(50.0)))) ? (R(5.0)) : (R(1.0)));
    const R calcium_release__i_rel = (((cmx_ipow<2>(cmx_div(y[12], (y[12] + R(0.25)))) * calcium_release__K_m_rel) + calcium_release__K_leak_rate) * y[19]);
    const R calcium_translocation__i_trans = (R(50.0) * (y[18] - y[19]));
    const R reversal_potentials__E_mh = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)), (y[14] + (R(0.12) * y[15])))));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * (V - reversal_potentials__E_mh));
    const R fast_sodium_current_h_gate__alpha_h = (R(20.0) * cmx_exp_v(((-R(0.125)) * ((V + R(75.0)) - fast_sodium_current_h_gate__shift_h)), V, -0.18033688011112042, -13.525266008334032));
    const R fast_sodium_current_h_gate__beta_h = cmx_div(R(2000.0), (R(1.0) + (R(320.0) * cmx_exp_v(((-R(0.1)) * ((V + R(75.0)) - fast_sodium_current_h_gate__shift_h)), V, -0.14426950408889636, -10.820212806667225))));
    const R fast_sodium_current_m_gate__E0_m = (V + R(41.0));
    const R fast_sodium_current_m_gate__alpha_m = (((fabs(fast_sodium_current_m_gate__E0_m) < fast_sodium_current_m_gate__delta_m)) ? (R(2000.0)) : (cmx_div((R(200.0) * fast_sodium_current_m_gate__E0_m), (-cmx_expm1(((-R(0.1)) * fast_sodium_current_m_gate__E0_m))))));
    const R fast_sodium_current_m_gate__beta_m = (R(8000.0) * cmx_exp_v(((-R(0.056)) * (V + R(66.0))), V, -0.08079092228978195, -5.332200871125609));
    const R intracellular_calcium_concentration__V_Cell = (R(3.141592654) * cmx_ipow<2>(intracellular_calcium_concentration__radius) * intracellular_calcium_concentration__length);
    const R intracellular_calcium_concentration__V_i_ratio = (((R(1.0) - intracellular_calcium_concentration__V_e_ratio) - intracellular_calcium_concentration__V_up_ratio) - intracellular_calcium_concentration__V_rel_ratio);
    const R intracellular_calcium_concentration__V_i = (intracellular_calcium_concentration__V_Cell * intracellular_calcium_concentration__V_i_ratio);
    const R intracellular_calcium_concentration__d_Ca_Calmod_dt = ((intracellular_calcium_concentration__alpha_Calmod * y[16] * (intracellular_calcium_concentration__Calmod - y[20])) - (intracellular_calcium_concentration__beta_Calmod * y[20]));
    const R intracellular_calcium_concentration__d_Ca_Trop_dt = ((intracellular_calcium_concentration__alpha_Trop * y[16] * (intracellular_calcium_concentration__Trop - y[21])) - (intracellular_calcium_concentration__beta_Trop * y[21]));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_sodium_concentration__Na_o, y[14])));
    const R persistent_sodium_current__i_p_Na = (cmx_div((persistent_sodium_current__g_pna * R(1.0)), (R(1.0) + cmx_exp_v(((-(V + R(52.0))) * (R(1) / R(8.0))), V, -0.18033688011112042, -9.377517765778261))) * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_potassium_concentration__K_o, y[15])));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (cmx_div((((p[4] * y[1]) + (rapid_delayed_rectifier_potassium_current__g_Kr2 * y[2])) * R(1.0)), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003))) * (V - reversal_potentials__E_K));
    const R rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 = cmx_div(R(50.0), (R(1.0) + cmx_exp_v(((-(V - R(5.0))) * (R(1) / R(9.0))), V, -0.16029944898766257, 0.801497244938313)));
    const R rapid_delayed_rectifier_potassium_current_xr1_gate__beta_xr1 = (R(0.05) * cmx_exp_v(((-(V - R(20.0))) * (R(1) / R(15.0))), V, -0.09617966939259756, 1.923593387851951));
    const R rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 = cmx_div(R(50.0), (R(1.0) + cmx_exp_v(((-(V - R(5.0))) * (R(1) / R(9.0))), V, -0.16029944898766257, 0.801497244938313)));
    const R rapid_delayed_rectifier_potassium_current_xr2_gate__beta_xr2 = (R(0.4) * cmx_exp((-cmx_ipow<3>(((V + R(30.0)) * (R(1) / R(30.0)))))));
    const R sarcoplasmic_reticulum_calcium_pump__K_1 = ((sarcoplasmic_reticulum_calcium_pump__K_cyca * sarcoplasmic_reticulum_calcium_pump__K_xcs) * (R(1) / sarcoplasmic_reticulum_calcium_pump__K_srca));
    const R sarcoplasmic_reticulum_calcium_pump__K_2 = (y[16] + (y[18] * sarcoplasmic_reticulum_calcium_pump__K_1) + (sarcoplasmic_reticulum_calcium_pump__K_cyca * sarcoplasmic_reticulum_calcium_pump__K_xcs) + sarcoplasmic_reticulum_calcium_pump__K_cyca);
    const R sarcoplasmic_reticulum_calcium_pump__i_up = ((cmx_div(y[16], sarcoplasmic_reticulum_calcium_pump__K_2) * sarcoplasmic_reticulum_calcium_pump__alpha_up) - (cmx_div((y[18] * sarcoplasmic_reticulum_calcium_pump__K_1), sarcoplasmic_reticulum_calcium_pump__K_2) * sarcoplasmic_reticulum_calcium_pump__beta_up));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)), (y[15] + (reversal_potentials__P_kna * y[14])))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs = cmx_div(R(14.0), (R(1.0) + cmx_exp_v(((-(V - R(40.0))) * (R(1) / R(9.0))), V, -0.1602994489876626, 6.411977959506504)));
    const R slow_delayed_rectifier_potassium_current_xs_gate__beta_xs = (R(1.0) * cmx_exp_v(((-V) * (R(1) / R(45.0))), V, -0.03205988979753251, -0.0));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger__i_NaCa_cyt = cmx_div(((R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(y[14], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * y[16]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((y[16] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(y[14], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (y[16] * (R(1) / R(0.0069))))));
    const R sodium_calcium_exchanger__i_NaCa_ds = cmx_div((sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(y[14], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * y[17]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((y[17] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(y[14], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (y[17] * (R(1) / R(0.0069))))));
    const R sodium_potassium_pump__i_NaK = cmx_div((((sodium_potassium_pump__i_NaK_max * extracellular_potassium_concentration__K_o) * (R(1) / (sodium_potassium_pump__K_mK + extracellular_potassium_concentration__K_o))) * y[14]), (sodium_potassium_pump__K_mNa + y[14]));
    const R time_independent_potassium_current__i_K1 = cmx_div((((p[1] * extracellular_potassium_concentration__K_o) * (R(1) / (extracellular_potassium_concentration__K_o + time_independent_potassium_current__K_mk1))) * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07)));
    const R transient_outward_current__i_to = (transient_outward_current__g_to * (transient_outward_current__g_tos + (y[10] * (R(1.0) - transient_outward_current__g_tos))) * y[11] * (V - reversal_potentials__E_K));
    const R transient_outward_current_s_gate__alpha_s = (R(0.033) * cmx_exp_v(((-V) * (R(1) / R(17.0))), V, -0.08486441416993903, -0.0));
    const R transient_outward_current_s_gate__beta_s = cmx_div(R(33.0), (R(1.0) + cmx_exp_v(((-R(0.125)) * (V + R(10.0))), V, -0.18033688011112042, -1.8033688011112043)));
    if (RL) {
        g_k[1] = cmx_div(((rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 * (-R(1))) - (rapid_delayed_rectifier_potassium_current_xr1_gate__beta_xr1 * R(1))), R(1000.0));
        g_inf[1] = cmx_div(-((rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 * R(1.0))), ((rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 * (-R(1))) - (rapid_delayed_rectifier_potassium_current_xr1_gate__beta_xr1 * R(1))));
    } else {
        dy[1] = (((rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 * (R(1.0) - y[1])) - (rapid_delayed_rectifier_potassium_current_xr1_gate__beta_xr1 * y[1])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[2] = cmx_div(((rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 * (-R(1))) - (rapid_delayed_rectifier_potassium_current_xr2_gate__beta_xr2 * R(1))), R(1000.0));
        g_inf[2] = cmx_div(-((rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 * R(1.0))), ((rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 * (-R(1))) - (rapid_delayed_rectifier_potassium_current_xr2_gate__beta_xr2 * R(1))));
    } else {
        dy[2] = (((rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 * (R(1.0) - y[2])) - (rapid_delayed_rectifier_potassium_current_xr2_gate__beta_xr2 * y[2])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[3] = cmx_div(((slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs * (-R(1))) - (slow_delayed_rectifier_potassium_current_xs_gate__beta_xs * R(1))), R(1000.0));
        g_inf[3] = cmx_div(-((slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs * R(1.0))), ((slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs * (-R(1))) - (slow_delayed_rectifier_potassium_current_xs_gate__beta_xs * R(1))));
    } else {
        dy[3] = (((slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs * (R(1.0) - y[3])) - (slow_delayed_rectifier_potassium_current_xs_gate__beta_xs * y[3])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[4] = cmx_div(((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1))), R(1000.0));
        g_inf[4] = cmx_div(-((fast_sodium_current_m_gate__alpha_m * R(1.0))), ((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1))));
    } else {
        dy[4] = (((fast_sodium_current_m_gate__alpha_m * (R(1.0) - y[4])) - (fast_sodium_current_m_gate__beta_m * y[4])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[5] = cmx_div(((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1))), R(1000.0));
        g_inf[5] = cmx_div(-((fast_sodium_current_h_gate__alpha_h * R(1.0))), ((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1))));
    } else {
        dy[5] = (((fast_sodium_current_h_gate__alpha_h * (R(1.0) - y[5])) - (fast_sodium_current_h_gate__beta_h * y[5])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[6] = cmx_div((L_type_Ca_channel_d_gate__speed_d * ((L_type_Ca_channel_d_gate__alpha_d * (-R(1))) - (L_type_Ca_channel_d_gate__beta_d * R(1)))), R(1000.0));
        g_inf[6] = cmx_div(-((L_type_Ca_channel_d_gate__speed_d * (L_type_Ca_channel_d_gate__alpha_d * R(1.0)))), (L_type_Ca_channel_d_gate__speed_d * ((L_type_Ca_channel_d_gate__alpha_d * (-R(1))) - (L_type_Ca_channel_d_gate__beta_d * R(1)))));
    } else {
        dy[6] = ((L_type_Ca_channel_d_gate__speed_d * ((L_type_Ca_channel_d_gate__alpha_d * (R(1.0) - y[6])) - (L_type_Ca_channel_d_gate__beta_d * y[6]))) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[7] = cmx_div((L_type_Ca_channel_f_gate__speed_f * ((L_type_Ca_channel_f_gate__alpha_f * (-R(1))) - (L_type_Ca_channel_f_gate__beta_f * R(1)))), R(1000.0));
        g_inf[7] = cmx_div(-((L_type_Ca_channel_f_gate__speed_f * (L_type_Ca_channel_f_gate__alpha_f * R(1.0)))), (L_type_Ca_channel_f_gate__speed_f * ((L_type_Ca_channel_f_gate__alpha_f * (-R(1))) - (L_type_Ca_channel_f_gate__beta_f * R(1)))));
    } else {
        dy[7] = ((L_type_Ca_channel_f_gate__speed_f * ((L_type_Ca_channel_f_gate__alpha_f * (R(1.0) - y[7])) - (L_type_Ca_channel_f_gate__beta_f * y[7]))) * (R(1) / R(1000.0)));
    }
    dy[8] = ((R(1.0) - (R(1.0) * (cmx_div(y[16], (L_type_Ca_channel__Km_f2 + y[16])) + y[8]))) * (R(1) / R(1000.0)));
    dy[9] = ((L_type_Ca_channel__R_decay * (R(1.0) - (cmx_div(y[17], (L_type_Ca_channel__Km_f2ds + y[17])) + y[9]))) * (R(1) / R(1000.0)));
    if (RL) {
        g_k[10] = cmx_div(((transient_outward_current_s_gate__alpha_s * (-R(1))) - (transient_outward_current_s_gate__beta_s * R(1))), R(1000.0));
        g_inf[10] = cmx_div(-((transient_outward_current_s_gate__alpha_s * R(1.0))), ((transient_outward_current_s_gate__alpha_s * (-R(1))) - (transient_outward_current_s_gate__beta_s * R(1))));
    } else {
        dy[10] = (((transient_outward_current_s_gate__alpha_s * (R(1.0) - y[10])) - (transient_outward_current_s_gate__beta_s * y[10])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[11] = cmx_div((R(333.0) * (-R(1))), R(1000.0));
        g_inf[11] = cmx_div(-((R(333.0) * cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(4.0))) * (R(1) / R(5.0))), V, -0.2885390081777927, -1.1541560327111708))))), (R(333.0) * (-R(1))));
    } else {
        dy[11] = ((R(333.0) * (cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(4.0))) * (R(1) / R(5.0))), V, -0.2885390081777927, -1.1541560327111708))) - y[11])) * (R(1) / R(1000.0)));
    }
    dy[12] = (((calcium_release__PrecFrac * calcium_release__SpeedRel * calcium_release__ActRate) - (y[12] * calcium_release__SpeedRel * calcium_release__InactRate)) * (R(1) / R(1000.0)));
    dy[13] = (((y[12] * calcium_release__SpeedRel * calcium_release__InactRate) - (calcium_release__SpeedRel * R(1.0) * y[13])) * (R(1) / R(1000.0)));
    dy[14] = ((((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))) * (fast_sodium_current__i_Na + persistent_sodium_current__i_p_Na + sodium_background_current__i_b_Na + (R(3.0) * sodium_potassium_pump__i_NaK) + (R(3.0) * sodium_calcium_exchanger__i_NaCa_cyt) + L_type_Ca_channel__i_Ca_L_Na_cyt + L_type_Ca_channel__i_Ca_L_Na_ds)) * (R(1) / R(1000.0)));
    dy[15] = ((((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))) * ((time_independent_potassium_current__i_K1 + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + L_type_Ca_channel__i_Ca_L_K_cyt + L_type_Ca_channel__i_Ca_L_K_ds + transient_outward_current__i_to) - (R(2.0) * sodium_potassium_pump__i_NaK))) * (R(1) / R(1000.0)));
    dy[16] = ((((((((-R(1.0)) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))) * (((L_type_Ca_channel__i_Ca_L_Ca_cyt + calcium_background_current__i_b_Ca) - (R(2.0) * sodium_calcium_exchanger__i_NaCa_cyt)) - (R(2.0) * sodium_calcium_exchanger__i_NaCa_ds))) + (y[17] * intracellular_calcium_concentration__V_ds_ratio * intracellular_calcium_concentration__Kdecay) + ((calcium_release__i_rel * intracellular_calcium_concentration__V_rel_ratio) * (R(1) / intracellular_calcium_concentration__V_i_ratio))) - intracellular_calcium_concentration__d_Ca_Calmod_dt) - intracellular_calcium_concentration__d_Ca_Trop_dt) - sarcoplasmic_reticulum_calcium_pump__i_up) * (R(1) / R(1000.0)));
    dy[17] = (((((-R(1.0)) * L_type_Ca_channel__i_Ca_L_Ca_ds) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_ds_ratio * intracellular_calcium_concentration__V_i * membrane__F))) - (y[17] * intracellular_calcium_concentration__Kdecay)) * (R(1) / R(1000.0)));
    dy[18] = ((((intracellular_calcium_concentration__V_i_ratio * (R(1) / intracellular_calcium_concentration__V_up_ratio)) * sarcoplasmic_reticulum_calcium_pump__i_up) - calcium_translocation__i_trans) * (R(1) / R(1000.0)));
    dy[19] = ((((intracellular_calcium_concentration__V_up_ratio * (R(1) / intracellular_calcium_concentration__V_rel_ratio)) * calcium_translocation__i_trans) - calcium_release__i_rel) * (R(1) / R(1000.0)));
    dy[20] = (((intracellular_calcium_concentration__alpha_Calmod * y[16] * (intracellular_calcium_concentration__Calmod - y[20])) - (intracellular_calcium_concentration__beta_Calmod * y[20])) * (R(1) / R(1000.0)));
    dy[21] = (((intracellular_calcium_concentration__alpha_Trop * y[16] * (intracellular_calcium_concentration__Trop - y[21])) - (intracellular_calcium_concentration__beta_Trop * y[21])) * (R(1) / R(1000.0)));
}

// dV/dt including the stimulus (single-cell integration).
template <class R>
CMX_HD inline R dvdt(const R* y, R V, R i_stim_area, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8314.472);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.3415);
    const R membrane__Cm = R(9.5e-05);
    const R reversal_potentials__P_kna = R(0.03);
    const R time_independent_potassium_current__K_mk1 = R(10.0);
    const R rapid_delayed_rectifier_potassium_current__g_Kr2 = R(0.0013);
    const R persistent_sodium_current__g_pna = R(0.004);
    const R sodium_background_current__g_bna = R(0.0006);
    const R L_type_Ca_channel__P_CaK = R(0.002);
    const R L_type_Ca_channel__P_CaNa = R(0.01);
    const R L_type_Ca_channel__FrICa = R(1.0);
    const R calcium_background_current__g_bca = R(0.00025);
    const R transient_outward_current__g_to = R(0.005);
    const R transient_outward_current__g_tos = R(0.0);
    const R sodium_potassium_pump__i_NaK_max = R(0.7);
    const R sodium_potassium_pump__K_mK = R(1.0);
    const R sodium_potassium_pump__K_mNa = R(40.0);
    const R sodium_calcium_exchanger__k_NaCa = R(0.0005);
    const R sodium_calcium_exchanger__n_NaCa = R(3.0);
    const R sodium_calcium_exchanger__d_NaCa = R(0.0);
    const R sodium_calcium_exchanger__gamma = R(0.5);
    const R sodium_calcium_exchanger__FRiNaCa = R(0.001);
    const R extracellular_sodium_concentration__Na_o = R(140.0);
    const R extracellular_calcium_concentration__Ca_o = R(2.0);
    const R extracellular_potassium_concentration__K_o = R(4.0);
    const R i_stim = (i_stim_area / cm) * R(0.09499999999999999);
    (void)i_stim;
    const R L_type_Ca_channel__i_Ca_L_Ca_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * R(4.0) * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((y[16] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_Ca_ds = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * R(4.0) * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((y[16] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_K_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[15] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_K_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[15] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[14] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[14] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_calcium_concentration__Ca_o, y[16])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R reversal_potentials__E_mh = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)), (y[14] + (R(0.12) * y[15])))));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * (V - reversal_potentials__E_mh));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_sodium_concentration__Na_o, y[14])));
    const R persistent_sodium_current__i_p_Na = (cmx_div((persistent_sodium_current__g_pna * R(1.0)), (R(1.0) + cmx_exp_v(((-(V + R(52.0))) * (R(1) / R(8.0))), V, -0.18033688011112042, -9.377517765778261))) * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_potassium_concentration__K_o, y[15])));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (cmx_div((((p[4] * y[1]) + (rapid_delayed_rectifier_potassium_current__g_Kr2 * y[2])) * R(1.0)), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003))) * (V - reversal_potentials__E_K));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)), (y[15] + (reversal_potentials__P_kna * y[14])))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger__i_NaCa_cyt = cmx_div(((R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(y[14], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * y[16]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((y[16] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(y[14], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (y[16] * (R(1) / R(0.0069))))));
    const R sodium_calcium_exchanger__i_NaCa_ds = cmx_div((sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(y[14], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * y[17]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((y[17] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(y[14], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (y[17] * (R(1) / R(0.0069))))));
    const R sodium_potassium_pump__i_NaK = cmx_div((((sodium_potassium_pump__i_NaK_max * extracellular_potassium_concentration__K_o) * (R(1) / (sodium_potassium_pump__K_mK + extracellular_potassium_concentration__K_o))) * y[14]), (sodium_potassium_pump__K_mNa + y[14]));
    const R time_independent_potassium_current__i_K1 = cmx_div((((p[1] * extracellular_potassium_concentration__K_o) * (R(1) / (extracellular_potassium_concentration__K_o + time_independent_potassium_current__K_mk1))) * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07)));
    const R transient_outward_current__i_to = (transient_outward_current__g_to * (transient_outward_current__g_tos + (y[10] * (R(1.0) - transient_outward_current__g_tos))) * y[11] * (V - reversal_potentials__E_K));
    return ((((-R(1.0)) * (R(1) / membrane__Cm)) * (i_stim + time_independent_potassium_current__i_K1 + transient_outward_current__i_to + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + sodium_potassium_pump__i_NaK + fast_sodium_current__i_Na + sodium_background_current__i_b_Na + persistent_sodium_current__i_p_Na + L_type_Ca_channel__i_Ca_L_Na_cyt + L_type_Ca_channel__i_Ca_L_Na_ds + sodium_calcium_exchanger__i_NaCa_cyt + sodium_calcium_exchanger__i_NaCa_ds + L_type_Ca_channel__i_Ca_L_Ca_cyt + L_type_Ca_channel__i_Ca_L_Ca_ds + L_type_Ca_channel__i_Ca_L_K_cyt + L_type_Ca_channel__i_Ca_L_K_ds + calcium_background_current__i_b_Ca)) * (R(1) / R(1000.0)));
}

// Total ionic current in uA/cm^2 (GetIIonic): minus dV/dt at zero stimulus,
// times the membrane capacitance that converts it to a current density.
template <class R>
CMX_HD inline R iionic(const R* y, R V, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8314.472);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.3415);
    const R membrane__Cm = R(9.5e-05);
    const R reversal_potentials__P_kna = R(0.03);
    const R time_independent_potassium_current__K_mk1 = R(10.0);
    const R rapid_delayed_rectifier_potassium_current__g_Kr2 = R(0.0013);
    const R persistent_sodium_current__g_pna = R(0.004);
    const R sodium_background_current__g_bna = R(0.0006);
    const R L_type_Ca_channel__P_CaK = R(0.002);
    const R L_type_Ca_channel__P_CaNa = R(0.01);
    const R L_type_Ca_channel__FrICa = R(1.0);
    const R calcium_background_current__g_bca = R(0.00025);
    const R transient_outward_current__g_to = R(0.005);
    const R transient_outward_current__g_tos = R(0.0);
    const R sodium_potassium_pump__i_NaK_max = R(0.7);
    const R sodium_potassium_pump__K_mK = R(1.0);
    const R sodium_potassium_pump__K_mNa = R(40.0);
    const R sodium_calcium_exchanger__k_NaCa = R(0.0005);
    const R sodium_calcium_exchanger__n_NaCa = R(3.0);
    const R sodium_calcium_exchanger__d_NaCa = R(0.0);
    const R sodium_calcium_exchanger__gamma = R(0.5);
    const R sodium_calcium_exchanger__FRiNaCa = R(0.001);
    const R extracellular_sodium_concentration__Na_o = R(140.0);
    const R extracellular_calcium_concentration__Ca_o = R(2.0);
    const R extracellular_potassium_concentration__K_o = R(4.0);
    const R i_stim = R(0);
    (void)i_stim;
    const R L_type_Ca_channel__i_Ca_L_Ca_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * R(4.0) * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((y[16] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_Ca_ds = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * R(4.0) * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((y[16] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_K_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[15] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_K_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[15] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * y[8] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[14] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * y[9] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((y[14] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_calcium_concentration__Ca_o, y[16])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R reversal_potentials__E_mh = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)), (y[14] + (R(0.12) * y[15])))));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * (V - reversal_potentials__E_mh));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_sodium_concentration__Na_o, y[14])));
    const R persistent_sodium_current__i_p_Na = (cmx_div((persistent_sodium_current__g_pna * R(1.0)), (R(1.0) + cmx_exp_v(((-(V + R(52.0))) * (R(1) / R(8.0))), V, -0.18033688011112042, -9.377517765778261))) * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_potassium_concentration__K_o, y[15])));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (cmx_div((((p[4] * y[1]) + (rapid_delayed_rectifier_potassium_current__g_Kr2 * y[2])) * R(1.0)), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003))) * (V - reversal_potentials__E_K));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)), (y[15] + (reversal_potentials__P_kna * y[14])))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger__i_NaCa_cyt = cmx_div(((R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(y[14], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * y[16]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((y[16] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(y[14], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (y[16] * (R(1) / R(0.0069))))));
    const R sodium_calcium_exchanger__i_NaCa_ds = cmx_div((sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(y[14], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * y[17]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((y[17] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(y[14], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (y[17] * (R(1) / R(0.0069))))));
    const R sodium_potassium_pump__i_NaK = cmx_div((((sodium_potassium_pump__i_NaK_max * extracellular_potassium_concentration__K_o) * (R(1) / (sodium_potassium_pump__K_mK + extracellular_potassium_concentration__K_o))) * y[14]), (sodium_potassium_pump__K_mNa + y[14]));
    const R time_independent_potassium_current__i_K1 = cmx_div((((p[1] * extracellular_potassium_concentration__K_o) * (R(1) / (extracellular_potassium_concentration__K_o + time_independent_potassium_current__K_mk1))) * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07)));
    const R transient_outward_current__i_to = (transient_outward_current__g_to * (transient_outward_current__g_tos + (y[10] * (R(1.0) - transient_outward_current__g_tos))) * y[11] * (V - reversal_potentials__E_K));
    const R dv = ((((-R(1.0)) * (R(1) / membrane__Cm)) * (i_stim + time_independent_potassium_current__i_K1 + transient_outward_current__i_to + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + sodium_potassium_pump__i_NaK + fast_sodium_current__i_Na + sodium_background_current__i_b_Na + persistent_sodium_current__i_p_Na + L_type_Ca_channel__i_Ca_L_Na_cyt + L_type_Ca_channel__i_Ca_L_Na_ds + sodium_calcium_exchanger__i_NaCa_cyt + sodium_calcium_exchanger__i_NaCa_ds + L_type_Ca_channel__i_Ca_L_Ca_cyt + L_type_Ca_channel__i_Ca_L_Ca_ds + L_type_Ca_channel__i_Ca_L_K_cyt + L_type_Ca_channel__i_Ca_L_K_ds + calcium_background_current__i_b_Ca)) * (R(1) / R(1000.0)));
    return -dv * cm;
}

// Backward Euler: the states solved by Newton (all but the linear gates).
constexpr int kNumNonlinear = 12;
static constexpr int kNonlinear[12] = {8, 9, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21};

// Linear states: y <- (y + a dt) / (1 - b dt) for dy/dt = a + b y, a and b
// functions of V (ComputeOneStepExceptVoltage, closed-form part).
template <class R>
CMX_HD inline void be_linear(R* y, R V, const R* p, R cm, R dt)
{
    CMX_MATH_USING
    (void)p; (void)cm; (void)V;
    const R fast_sodium_current_m_gate__delta_m = R(1e-05);
    const R fast_sodium_current_h_gate__shift_h = R(0.0);
    const R L_type_Ca_channel_d_gate__speed_d = R(3.0);
    const R L_type_Ca_channel_f_gate__speed_f = R(0.3);
    const R L_type_Ca_channel_f_gate__delta_f = R(0.0001);
    const R L_type_Ca_channel_d_gate__E0_d = ((V + R(24.0)) - R(5.0));
    const R L_type_Ca_channel_d_gate__alpha_d = (((fabs(L_type_Ca_channel_d_gate__E0_d) < R(0.0001))) ? (R(120.0)) : (cmx_div((R(30.0) * L_type_Ca_channel_d_gate__E0_d), (-cmx_expm1(((-L_type_Ca_channel_d_gate__E0_d) * (R(1) / R(4.0))))))));
    const R L_type_Ca_channel_d_gate__beta_d = (((fabs(L_type_Ca_channel_d_gate__E0_d) < R(0.0001))) ? (R(120.0)) : (cmx_div((R(12.0) * L_type_Ca_channel_d_gate__E0_d), cmx_expm1((L_type_Ca_channel_d_gate__E0_d * (R(1) / R(10.0)))))));
    const R L_type_Ca_channel_f_gate__E0_f = (V + R(34.0));
    const R L_type_Ca_channel_f_gate__alpha_f = (((fabs(L_type_Ca_channel_f_gate__E0_f) < L_type_Ca_channel_f_gate__delta_f)) ? (R(25.0)) : (cmx_div((R(6.25) * L_type_Ca_channel_f_gate__E0_f), cmx_expm1((L_type_Ca_channel_f_gate__E0_f * (R(1) / R(4.0)))))));
    const R L_type_Ca_channel_f_gate__beta_f = cmx_div(R(12.0), (R(1.0) + cmx_exp_v((((-R(1.0)) * (V + R(34.0))) * (R(1) / R(4.0))), V, -0.36067376022224085, -12.262907847556189)));
    const R fast_sodium_current_h_gate__alpha_h = (R(20.0) * cmx_exp_v(((-R(0.125)) * ((V + R(75.0)) - fast_sodium_current_h_gate__shift_h)), V, -0.18033688011112042, -13.525266008334032));
    const R fast_sodium_current_h_gate__beta_h = cmx_div(R(2000.0), (R(1.0) + (R(320.0) * cmx_exp_v(((-R(0.1)) * ((V + R(75.0)) - fast_sodium_current_h_gate__shift_h)), V, -0.14426950408889636, -10.820212806667225))));
    const R fast_sodium_current_m_gate__E0_m = (V + R(41.0));
    const R fast_sodium_current_m_gate__alpha_m = (((fabs(fast_sodium_current_m_gate__E0_m) < fast_sodium_current_m_gate__delta_m)) ? (R(2000.0)) : (cmx_div((R(200.0) * fast_sodium_current_m_gate__E0_m), (-cmx_expm1(((-R(0.1)) * fast_sodium_current_m_gate__E0_m))))));
    const R fast_sodium_current_m_gate__beta_m = (R(8000.0) * cmx_exp_v(((-R(0.056)) * (V + R(66.0))), V, -0.08079092228978195, -5.332200871125609));
    const R rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 = cmx_div(R(50.0), (R(1.0) + cmx_exp_v(((-(V - R(5.0))) * (R(1) / R(9.0))), V, -0.16029944898766257, 0.801497244938313)));
    const R rapid_delayed_rectifier_potassium_current_xr1_gate__beta_xr1 = (R(0.05) * cmx_exp_v(((-(V - R(20.0))) * (R(1) / R(15.0))), V, -0.09617966939259756, 1.923593387851951));
    const R rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 = cmx_div(R(50.0), (R(1.0) + cmx_exp_v(((-(V - R(5.0))) * (R(1) / R(9.0))), V, -0.16029944898766257, 0.801497244938313)));
    const R rapid_delayed_rectifier_potassium_current_xr2_gate__beta_xr2 = (R(0.4) * cmx_exp((-cmx_ipow<3>(((V + R(30.0)) * (R(1) / R(30.0)))))));
    const R slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs = cmx_div(R(14.0), (R(1.0) + cmx_exp_v(((-(V - R(40.0))) * (R(1) / R(9.0))), V, -0.1602994489876626, 6.411977959506504)));
    const R slow_delayed_rectifier_potassium_current_xs_gate__beta_xs = (R(1.0) * cmx_exp_v(((-V) * (R(1) / R(45.0))), V, -0.03205988979753251, -0.0));
    const R transient_outward_current_s_gate__alpha_s = (R(0.033) * cmx_exp_v(((-V) * (R(1) / R(17.0))), V, -0.08486441416993903, -0.0));
    const R transient_outward_current_s_gate__beta_s = cmx_div(R(33.0), (R(1.0) + cmx_exp_v(((-R(0.125)) * (V + R(10.0))), V, -0.18033688011112042, -1.8033688011112043)));
    y[1] = (y[1] + (((rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((rapid_delayed_rectifier_potassium_current_xr1_gate__alpha_xr1 * (-R(1))) - (rapid_delayed_rectifier_potassium_current_xr1_gate__beta_xr1 * R(1))) / R(1000.0))) * dt);
    y[2] = (y[2] + (((rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((rapid_delayed_rectifier_potassium_current_xr2_gate__alpha_xr2 * (-R(1))) - (rapid_delayed_rectifier_potassium_current_xr2_gate__beta_xr2 * R(1))) / R(1000.0))) * dt);
    y[3] = (y[3] + (((slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((slow_delayed_rectifier_potassium_current_xs_gate__alpha_xs * (-R(1))) - (slow_delayed_rectifier_potassium_current_xs_gate__beta_xs * R(1))) / R(1000.0))) * dt);
    y[4] = (y[4] + (((fast_sodium_current_m_gate__alpha_m * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1))) / R(1000.0))) * dt);
    y[5] = (y[5] + (((fast_sodium_current_h_gate__alpha_h * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1))) / R(1000.0))) * dt);
    y[6] = (y[6] + (((L_type_Ca_channel_d_gate__speed_d * (L_type_Ca_channel_d_gate__alpha_d * R(1.0))) / R(1000.0))) * dt) / (R(1) - (((L_type_Ca_channel_d_gate__speed_d * ((L_type_Ca_channel_d_gate__alpha_d * (-R(1))) - (L_type_Ca_channel_d_gate__beta_d * R(1)))) / R(1000.0))) * dt);
    y[7] = (y[7] + (((L_type_Ca_channel_f_gate__speed_f * (L_type_Ca_channel_f_gate__alpha_f * R(1.0))) / R(1000.0))) * dt) / (R(1) - (((L_type_Ca_channel_f_gate__speed_f * ((L_type_Ca_channel_f_gate__alpha_f * (-R(1))) - (L_type_Ca_channel_f_gate__beta_f * R(1)))) / R(1000.0))) * dt);
    y[10] = (y[10] + (((transient_outward_current_s_gate__alpha_s * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((transient_outward_current_s_gate__alpha_s * (-R(1))) - (transient_outward_current_s_gate__beta_s * R(1))) / R(1000.0))) * dt);
    y[11] = (y[11] + (((R(333.0) * cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(4.0))) * (R(1) / R(5.0))), V, -0.2885390081777927, -1.1541560327111708)))) / R(1000.0))) * dt) / (R(1) - (((R(333.0) * (-R(1))) / R(1000.0))) * dt);
}

// Newton residual of the nonlinear states at guess g (y holds the other
// states, linear ones already advanced): g - y_old - dt f(g).
template <class R>
CMX_HD inline void be_residual(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* res)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8314.472);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.3415);
    const R reversal_potentials__P_kna = R(0.03);
    const R time_independent_potassium_current__K_mk1 = R(10.0);
    const R rapid_delayed_rectifier_potassium_current__g_Kr2 = R(0.0013);
    const R persistent_sodium_current__g_pna = R(0.004);
    const R sodium_background_current__g_bna = R(0.0006);
    const R L_type_Ca_channel__P_CaK = R(0.002);
    const R L_type_Ca_channel__P_CaNa = R(0.01);
    const R L_type_Ca_channel__Km_f2 = R(100000.0);
    const R L_type_Ca_channel__Km_f2ds = R(0.001);
    const R L_type_Ca_channel__R_decay = R(20.0);
    const R L_type_Ca_channel__FrICa = R(1.0);
    const R calcium_background_current__g_bca = R(0.00025);
    const R transient_outward_current__g_to = R(0.005);
    const R transient_outward_current__g_tos = R(0.0);
    const R sodium_potassium_pump__i_NaK_max = R(0.7);
    const R sodium_potassium_pump__K_mK = R(1.0);
    const R sodium_potassium_pump__K_mNa = R(40.0);
    const R sodium_calcium_exchanger__k_NaCa = R(0.0005);
    const R sodium_calcium_exchanger__n_NaCa = R(3.0);
    const R sodium_calcium_exchanger__d_NaCa = R(0.0);
    const R sodium_calcium_exchanger__gamma = R(0.5);
    const R sodium_calcium_exchanger__FRiNaCa = R(0.001);
    const R sarcoplasmic_reticulum_calcium_pump__K_cyca = R(0.0003);
    const R sarcoplasmic_reticulum_calcium_pump__K_xcs = R(0.4);
    const R sarcoplasmic_reticulum_calcium_pump__K_srca = R(0.5);
    const R sarcoplasmic_reticulum_calcium_pump__alpha_up = R(0.4);
    const R sarcoplasmic_reticulum_calcium_pump__beta_up = R(0.03);
    const R calcium_release__K_leak_rate = R(0.05);
    const R calcium_release__K_m_rel = R(250.0);
    const R calcium_release__K_m_Ca_cyt = R(0.0005);
    const R calcium_release__K_m_Ca_ds = R(0.01);
    const R extracellular_sodium_concentration__Na_o = R(140.0);
    const R extracellular_calcium_concentration__Ca_o = R(2.0);
    const R extracellular_potassium_concentration__K_o = R(4.0);
    const R intracellular_calcium_concentration__Calmod = R(0.02);
    const R intracellular_calcium_concentration__Trop = R(0.05);
    const R intracellular_calcium_concentration__alpha_Calmod = R(100000.0);
    const R intracellular_calcium_concentration__beta_Calmod = R(50.0);
    const R intracellular_calcium_concentration__alpha_Trop = R(100000.0);
    const R intracellular_calcium_concentration__beta_Trop = R(200.0);
    const R intracellular_calcium_concentration__radius = R(0.012);
    const R intracellular_calcium_concentration__length = R(0.074);
    const R intracellular_calcium_concentration__V_ds_ratio = R(0.1);
    const R intracellular_calcium_concentration__V_rel_ratio = R(0.1);
    const R intracellular_calcium_concentration__V_e_ratio = R(0.4);
    const R intracellular_calcium_concentration__V_up_ratio = R(0.01);
    const R intracellular_calcium_concentration__Kdecay = R(10.0);
    const R i_stim = (i_stim_area / cm) * R(0.09499999999999999);
    (void)i_stim;
    const R L_type_Ca_channel__i_Ca_L_Ca_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * R(4.0) * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((g[6] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_Ca_ds = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * R(4.0) * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((g[6] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_K_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[5] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_K_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[5] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[4] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[4] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_calcium_concentration__Ca_o, g[6])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_release__CadsReg = cmx_div(g[7], (g[7] + calcium_release__K_m_Ca_ds));
    const R calcium_release__CaiReg = cmx_div(g[6], (g[6] + calcium_release__K_m_Ca_cyt));
    const R calcium_release__RegBindSite = (calcium_release__CaiReg + ((R(1.0) - calcium_release__CaiReg) * calcium_release__CadsReg));
    const R calcium_release__VoltDep = cmx_exp_v((R(0.08) * (V - R(40.0))), V, 0.11541560327111708, -4.616624130844683);
    const R calcium_release__ActRate = ((R(0.0) * calcium_release__VoltDep) + (R(500.0) * cmx_ipow<2>(calcium_release__RegBindSite)));
    const R calcium_release__InactRate = (R(60.0) + (R(500.0) * cmx_ipow<2>(calcium_release__RegBindSite)));
    const R calcium_release__PrecFrac = ((R(1.0) - g[2]) - g[3]);
    const R calcium_release__SpeedRel = (((V < (-R(50.0)))) ? (R(5.0)) : (R(1.0)));
    const R calcium_release__i_rel = (((cmx_ipow<2>(cmx_div(g[2], (g[2] + R(0.25)))) * calcium_release__K_m_rel) + calcium_release__K_leak_rate) * g[9]);
    const R calcium_translocation__i_trans = (R(50.0) * (g[8] - g[9]));
    const R reversal_potentials__E_mh = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)), (g[4] + (R(0.12) * g[5])))));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * (V - reversal_potentials__E_mh));
    const R intracellular_calcium_concentration__V_Cell = (R(3.141592654) * cmx_ipow<2>(intracellular_calcium_concentration__radius) * intracellular_calcium_concentration__length);
    const R intracellular_calcium_concentration__V_i_ratio = (((R(1.0) - intracellular_calcium_concentration__V_e_ratio) - intracellular_calcium_concentration__V_up_ratio) - intracellular_calcium_concentration__V_rel_ratio);
    const R intracellular_calcium_concentration__V_i = (intracellular_calcium_concentration__V_Cell * intracellular_calcium_concentration__V_i_ratio);
    const R intracellular_calcium_concentration__d_Ca_Calmod_dt = ((intracellular_calcium_concentration__alpha_Calmod * g[6] * (intracellular_calcium_concentration__Calmod - g[10])) - (intracellular_calcium_concentration__beta_Calmod * g[10]));
    const R intracellular_calcium_concentration__d_Ca_Trop_dt = ((intracellular_calcium_concentration__alpha_Trop * g[6] * (intracellular_calcium_concentration__Trop - g[11])) - (intracellular_calcium_concentration__beta_Trop * g[11]));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_sodium_concentration__Na_o, g[4])));
    const R persistent_sodium_current__i_p_Na = (cmx_div((persistent_sodium_current__g_pna * R(1.0)), (R(1.0) + cmx_exp_v(((-(V + R(52.0))) * (R(1) / R(8.0))), V, -0.18033688011112042, -9.377517765778261))) * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_potassium_concentration__K_o, g[5])));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (cmx_div((((p[4] * y[1]) + (rapid_delayed_rectifier_potassium_current__g_Kr2 * y[2])) * R(1.0)), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003))) * (V - reversal_potentials__E_K));
    const R sarcoplasmic_reticulum_calcium_pump__K_1 = ((sarcoplasmic_reticulum_calcium_pump__K_cyca * sarcoplasmic_reticulum_calcium_pump__K_xcs) * (R(1) / sarcoplasmic_reticulum_calcium_pump__K_srca));
    const R sarcoplasmic_reticulum_calcium_pump__K_2 = (g[6] + (g[8] * sarcoplasmic_reticulum_calcium_pump__K_1) + (sarcoplasmic_reticulum_calcium_pump__K_cyca * sarcoplasmic_reticulum_calcium_pump__K_xcs) + sarcoplasmic_reticulum_calcium_pump__K_cyca);
    const R sarcoplasmic_reticulum_calcium_pump__i_up = ((cmx_div(g[6], sarcoplasmic_reticulum_calcium_pump__K_2) * sarcoplasmic_reticulum_calcium_pump__alpha_up) - (cmx_div((g[8] * sarcoplasmic_reticulum_calcium_pump__K_1), sarcoplasmic_reticulum_calcium_pump__K_2) * sarcoplasmic_reticulum_calcium_pump__beta_up));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)), (g[5] + (reversal_potentials__P_kna * g[4])))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger__i_NaCa_cyt = cmx_div(((R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[6]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))));
    const R sodium_calcium_exchanger__i_NaCa_ds = cmx_div((sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[7]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))));
    const R sodium_potassium_pump__i_NaK = cmx_div((((sodium_potassium_pump__i_NaK_max * extracellular_potassium_concentration__K_o) * (R(1) / (sodium_potassium_pump__K_mK + extracellular_potassium_concentration__K_o))) * g[4]), (sodium_potassium_pump__K_mNa + g[4]));
    const R time_independent_potassium_current__i_K1 = cmx_div((((p[1] * extracellular_potassium_concentration__K_o) * (R(1) / (extracellular_potassium_concentration__K_o + time_independent_potassium_current__K_mk1))) * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07)));
    const R transient_outward_current__i_to = (transient_outward_current__g_to * (transient_outward_current__g_tos + (y[10] * (R(1.0) - transient_outward_current__g_tos))) * y[11] * (V - reversal_potentials__E_K));
    res[0] = g[0] - y[8] - dt * ((R(1.0) - (R(1.0) * (cmx_div(g[6], (L_type_Ca_channel__Km_f2 + g[6])) + g[0]))) * (R(1) / R(1000.0)));
    res[1] = g[1] - y[9] - dt * ((L_type_Ca_channel__R_decay * (R(1.0) - (cmx_div(g[7], (L_type_Ca_channel__Km_f2ds + g[7])) + g[1]))) * (R(1) / R(1000.0)));
    res[2] = g[2] - y[12] - dt * (((calcium_release__PrecFrac * calcium_release__SpeedRel * calcium_release__ActRate) - (g[2] * calcium_release__SpeedRel * calcium_release__InactRate)) * (R(1) / R(1000.0)));
    res[3] = g[3] - y[13] - dt * (((g[2] * calcium_release__SpeedRel * calcium_release__InactRate) - (calcium_release__SpeedRel * R(1.0) * g[3])) * (R(1) / R(1000.0)));
    res[4] = g[4] - y[14] - dt * ((((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))) * (fast_sodium_current__i_Na + persistent_sodium_current__i_p_Na + sodium_background_current__i_b_Na + (R(3.0) * sodium_potassium_pump__i_NaK) + (R(3.0) * sodium_calcium_exchanger__i_NaCa_cyt) + L_type_Ca_channel__i_Ca_L_Na_cyt + L_type_Ca_channel__i_Ca_L_Na_ds)) * (R(1) / R(1000.0)));
    res[5] = g[5] - y[15] - dt * ((((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))) * ((time_independent_potassium_current__i_K1 + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + L_type_Ca_channel__i_Ca_L_K_cyt + L_type_Ca_channel__i_Ca_L_K_ds + transient_outward_current__i_to) - (R(2.0) * sodium_potassium_pump__i_NaK))) * (R(1) / R(1000.0)));
    res[6] = g[6] - y[16] - dt * ((((((((-R(1.0)) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))) * (((L_type_Ca_channel__i_Ca_L_Ca_cyt + calcium_background_current__i_b_Ca) - (R(2.0) * sodium_calcium_exchanger__i_NaCa_cyt)) - (R(2.0) * sodium_calcium_exchanger__i_NaCa_ds))) + (g[7] * intracellular_calcium_concentration__V_ds_ratio * intracellular_calcium_concentration__Kdecay) + ((calcium_release__i_rel * intracellular_calcium_concentration__V_rel_ratio) * (R(1) / intracellular_calcium_concentration__V_i_ratio))) - intracellular_calcium_concentration__d_Ca_Calmod_dt) - intracellular_calcium_concentration__d_Ca_Trop_dt) - sarcoplasmic_reticulum_calcium_pump__i_up) * (R(1) / R(1000.0)));
    res[7] = g[7] - y[17] - dt * (((((-R(1.0)) * L_type_Ca_channel__i_Ca_L_Ca_ds) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_ds_ratio * intracellular_calcium_concentration__V_i * membrane__F))) - (g[7] * intracellular_calcium_concentration__Kdecay)) * (R(1) / R(1000.0)));
    res[8] = g[8] - y[18] - dt * ((((intracellular_calcium_concentration__V_i_ratio * (R(1) / intracellular_calcium_concentration__V_up_ratio)) * sarcoplasmic_reticulum_calcium_pump__i_up) - calcium_translocation__i_trans) * (R(1) / R(1000.0)));
    res[9] = g[9] - y[19] - dt * ((((intracellular_calcium_concentration__V_up_ratio * (R(1) / intracellular_calcium_concentration__V_rel_ratio)) * calcium_translocation__i_trans) - calcium_release__i_rel) * (R(1) / R(1000.0)));
    res[10] = g[10] - y[20] - dt * (((intracellular_calcium_concentration__alpha_Calmod * g[6] * (intracellular_calcium_concentration__Calmod - g[10])) - (intracellular_calcium_concentration__beta_Calmod * g[10])) * (R(1) / R(1000.0)));
    res[11] = g[11] - y[21] - dt * (((intracellular_calcium_concentration__alpha_Trop * g[6] * (intracellular_calcium_concentration__Trop - g[11])) - (intracellular_calcium_concentration__beta_Trop * g[11])) * (R(1) / R(1000.0)));
}

// Jacobian of be_residual in g: J[i*N + j] = delta_ij - dt d f_i / d g_j.
template <class R>
CMX_HD inline void be_jacobian(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* J)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8314.472);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.3415);
    const R reversal_potentials__P_kna = R(0.03);
    const R time_independent_potassium_current__K_mk1 = R(10.0);
    const R rapid_delayed_rectifier_potassium_current__g_Kr2 = R(0.0013);
    const R persistent_sodium_current__g_pna = R(0.004);
    const R sodium_background_current__g_bna = R(0.0006);
    const R L_type_Ca_channel__P_CaK = R(0.002);
    const R L_type_Ca_channel__P_CaNa = R(0.01);
    const R L_type_Ca_channel__Km_f2 = R(100000.0);
    const R L_type_Ca_channel__Km_f2ds = R(0.001);
    const R L_type_Ca_channel__R_decay = R(20.0);
    const R L_type_Ca_channel__FrICa = R(1.0);
    const R calcium_background_current__g_bca = R(0.00025);
    const R transient_outward_current__g_to = R(0.005);
    const R transient_outward_current__g_tos = R(0.0);
    const R sodium_potassium_pump__i_NaK_max = R(0.7);
    const R sodium_potassium_pump__K_mK = R(1.0);
    const R sodium_potassium_pump__K_mNa = R(40.0);
    const R sodium_calcium_exchanger__k_NaCa = R(0.0005);
    const R sodium_calcium_exchanger__n_NaCa = R(3.0);
    const R sodium_calcium_exchanger__d_NaCa = R(0.0);
    const R sodium_calcium_exchanger__gamma = R(0.5);
    const R sodium_calcium_exchanger__FRiNaCa = R(0.001);
    const R sarcoplasmic_reticulum_calcium_pump__K_cyca = R(0.0003);
    const R sarcoplasmic_reticulum_calcium_pump__K_xcs = R(0.4);
    const R sarcoplasmic_reticulum_calcium_pump__K_srca = R(0.5);
    const R sarcoplasmic_reticulum_calcium_pump__alpha_up = R(0.4);
    const R sarcoplasmic_reticulum_calcium_pump__beta_up = R(0.03);
    const R calcium_release__K_leak_rate = R(0.05);
    const R calcium_release__K_m_rel = R(250.0);
    const R calcium_release__K_m_Ca_cyt = R(0.0005);
    const R calcium_release__K_m_Ca_ds = R(0.01);
    const R extracellular_sodium_concentration__Na_o = R(140.0);
    const R extracellular_calcium_concentration__Ca_o = R(2.0);
    const R extracellular_potassium_concentration__K_o = R(4.0);
    const R intracellular_calcium_concentration__Calmod = R(0.02);
    const R intracellular_calcium_concentration__Trop = R(0.05);
    const R intracellular_calcium_concentration__alpha_Calmod = R(100000.0);
    const R intracellular_calcium_concentration__beta_Calmod = R(50.0);
    const R intracellular_calcium_concentration__alpha_Trop = R(100000.0);
    const R intracellular_calcium_concentration__beta_Trop = R(200.0);
    const R intracellular_calcium_concentration__radius = R(0.012);
    const R intracellular_calcium_concentration__length = R(0.074);
    const R intracellular_calcium_concentration__V_ds_ratio = R(0.1);
    const R intracellular_calcium_concentration__V_rel_ratio = R(0.1);
    const R intracellular_calcium_concentration__V_e_ratio = R(0.4);
    const R intracellular_calcium_concentration__V_up_ratio = R(0.01);
    const R intracellular_calcium_concentration__Kdecay = R(10.0);
    const R i_stim = (i_stim_area / cm) * R(0.09499999999999999);
    (void)i_stim;
    const R L_type_Ca_channel__i_Ca_L_Ca_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * R(4.0) * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((g[6] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_Ca_ds = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * R(4.0) * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((g[6] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377)))); });
    const R L_type_Ca_channel__i_Ca_L_K_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[5] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_K_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[5] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_cyt = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[4] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R L_type_Ca_channel__i_Ca_L_Na_ds = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return (cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[4] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887)))); });
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_calcium_concentration__Ca_o, g[6])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_release__CadsReg = cmx_div(g[7], (g[7] + calcium_release__K_m_Ca_ds));
    const R calcium_release__CaiReg = cmx_div(g[6], (g[6] + calcium_release__K_m_Ca_cyt));
    const R calcium_release__RegBindSite = (calcium_release__CaiReg + ((R(1.0) - calcium_release__CaiReg) * calcium_release__CadsReg));
    const R calcium_release__VoltDep = cmx_exp_v((R(0.08) * (V - R(40.0))), V, 0.11541560327111708, -4.616624130844683);
    const R calcium_release__ActRate = ((R(0.0) * calcium_release__VoltDep) + (R(500.0) * cmx_ipow<2>(calcium_release__RegBindSite)));
    const R calcium_release__InactRate = (R(60.0) + (R(500.0) * cmx_ipow<2>(calcium_release__RegBindSite)));
    const R calcium_release__PrecFrac = ((R(1.0) - g[2]) - g[3]);
    const R calcium_release__SpeedRel = (((V < (-R(50.0)))) ? (R(5.0)) : (R(1.0)));
    const R calcium_release__i_rel = (((cmx_ipow<2>(cmx_div(g[2], (g[2] + R(0.25)))) * calcium_release__K_m_rel) + calcium_release__K_leak_rate) * g[9]);
    const R calcium_translocation__i_trans = (R(50.0) * (g[8] - g[9]));
    const R reversal_potentials__E_mh = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)), (g[4] + (R(0.12) * g[5])))));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * (V - reversal_potentials__E_mh));
    const R intracellular_calcium_concentration__V_Cell = (R(3.141592654) * cmx_ipow<2>(intracellular_calcium_concentration__radius) * intracellular_calcium_concentration__length);
    const R intracellular_calcium_concentration__V_i_ratio = (((R(1.0) - intracellular_calcium_concentration__V_e_ratio) - intracellular_calcium_concentration__V_up_ratio) - intracellular_calcium_concentration__V_rel_ratio);
    const R intracellular_calcium_concentration__V_i = (intracellular_calcium_concentration__V_Cell * intracellular_calcium_concentration__V_i_ratio);
    const R intracellular_calcium_concentration__d_Ca_Calmod_dt = ((intracellular_calcium_concentration__alpha_Calmod * g[6] * (intracellular_calcium_concentration__Calmod - g[10])) - (intracellular_calcium_concentration__beta_Calmod * g[10]));
    const R intracellular_calcium_concentration__d_Ca_Trop_dt = ((intracellular_calcium_concentration__alpha_Trop * g[6] * (intracellular_calcium_concentration__Trop - g[11])) - (intracellular_calcium_concentration__beta_Trop * g[11]));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_sodium_concentration__Na_o, g[4])));
    const R persistent_sodium_current__i_p_Na = (cmx_div((persistent_sodium_current__g_pna * R(1.0)), (R(1.0) + cmx_exp_v(((-(V + R(52.0))) * (R(1) / R(8.0))), V, -0.18033688011112042, -9.377517765778261))) * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(extracellular_potassium_concentration__K_o, g[5])));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (cmx_div((((p[4] * y[1]) + (rapid_delayed_rectifier_potassium_current__g_Kr2 * y[2])) * R(1.0)), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003))) * (V - reversal_potentials__E_K));
    const R sarcoplasmic_reticulum_calcium_pump__K_1 = ((sarcoplasmic_reticulum_calcium_pump__K_cyca * sarcoplasmic_reticulum_calcium_pump__K_xcs) * (R(1) / sarcoplasmic_reticulum_calcium_pump__K_srca));
    const R sarcoplasmic_reticulum_calcium_pump__K_2 = (g[6] + (g[8] * sarcoplasmic_reticulum_calcium_pump__K_1) + (sarcoplasmic_reticulum_calcium_pump__K_cyca * sarcoplasmic_reticulum_calcium_pump__K_xcs) + sarcoplasmic_reticulum_calcium_pump__K_cyca);
    const R sarcoplasmic_reticulum_calcium_pump__i_up = ((cmx_div(g[6], sarcoplasmic_reticulum_calcium_pump__K_2) * sarcoplasmic_reticulum_calcium_pump__alpha_up) - (cmx_div((g[8] * sarcoplasmic_reticulum_calcium_pump__K_1), sarcoplasmic_reticulum_calcium_pump__K_2) * sarcoplasmic_reticulum_calcium_pump__beta_up));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)), (g[5] + (reversal_potentials__P_kna * g[4])))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger__i_NaCa_cyt = cmx_div(((R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[6]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))));
    const R sodium_calcium_exchanger__i_NaCa_ds = cmx_div((sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[7]))), ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))));
    const R sodium_potassium_pump__i_NaK = cmx_div((((sodium_potassium_pump__i_NaK_max * extracellular_potassium_concentration__K_o) * (R(1) / (sodium_potassium_pump__K_mK + extracellular_potassium_concentration__K_o))) * g[4]), (sodium_potassium_pump__K_mNa + g[4]));
    const R time_independent_potassium_current__i_K1 = cmx_div((((p[1] * extracellular_potassium_concentration__K_o) * (R(1) / (extracellular_potassium_concentration__K_o + time_independent_potassium_current__K_mk1))) * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07)));
    const R transient_outward_current__i_to = (transient_outward_current__g_to * (transient_outward_current__g_tos + (y[10] * (R(1.0) - transient_outward_current__g_tos))) * y[11] * (V - reversal_potentials__E_K));
    const R d_L_type_Ca_channel__i_Ca_L_Ca_cyt__d0 = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return ((cmx_div(cmx_div(((R(1) * (R(1.0) - L_type_Ca_channel__FrICa) * R(4.0) * p[2] * y[6] * y[7] * (V - R(50.0)) * membrane__F)), (membrane__R * membrane__T)), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((g[6] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377))))); });
    const R d_L_type_Ca_channel__i_Ca_L_Ca_cyt__d6 = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return ((((R(1) * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * R(4.0) * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))))); });
    const R d_L_type_Ca_channel__i_Ca_L_Ca_ds__d1 = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return ((cmx_div(cmx_div(((R(1) * L_type_Ca_channel__FrICa * R(4.0) * p[2] * y[6] * y[7] * (V - R(50.0)) * membrane__F)), (membrane__R * membrane__T)), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))) * ((g[6] * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_calcium_concentration__Ca_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T))), V, -0.10801137730230755, 5.400568865115377))))); });
    const R d_L_type_Ca_channel__i_Ca_L_Ca_ds__d6 = cmx_remove_singularity(V, R(49.99999999999981), R(1.3356880329847772e-06), [&](const R V) { return ((((R(1) * cmx_exp(((R(100.0) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * cmx_div(((L_type_Ca_channel__FrICa * R(4.0) * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F * R(2.0)) * (R(1) / (membrane__R * membrane__T)))))))); });
    const R d_L_type_Ca_channel__i_Ca_L_K_cyt__d0 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((cmx_div(cmx_div(((R(1) * (R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * (V - R(50.0)) * membrane__F)), (membrane__R * membrane__T)), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[5] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887))))); });
    const R d_L_type_Ca_channel__i_Ca_L_K_cyt__d5 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((((R(1) * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))))); });
    const R d_L_type_Ca_channel__i_Ca_L_K_ds__d1 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((cmx_div(cmx_div(((R(1) * L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * (V - R(50.0)) * membrane__F)), (membrane__R * membrane__T)), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[5] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_potassium_concentration__K_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887))))); });
    const R d_L_type_Ca_channel__i_Ca_L_K_ds__d5 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((((R(1) * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaK * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))))); });
    const R d_L_type_Ca_channel__i_Ca_L_Na_cyt__d0 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((cmx_div(cmx_div(((R(1) * (R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * (V - R(50.0)) * membrane__F)), (membrane__R * membrane__T)), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[4] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887))))); });
    const R d_L_type_Ca_channel__i_Ca_L_Na_cyt__d4 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((((R(1) * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * cmx_div((((R(1.0) - L_type_Ca_channel__FrICa) * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * g[0] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))))); });
    const R d_L_type_Ca_channel__i_Ca_L_Na_ds__d1 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((cmx_div(cmx_div(((R(1) * L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * (V - R(50.0)) * membrane__F)), (membrane__R * membrane__T)), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * ((g[4] * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) - (extracellular_sodium_concentration__Na_o * cmx_exp_v((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115377, 2.7002844325576887))))); });
    const R d_L_type_Ca_channel__i_Ca_L_Na_ds__d4 = cmx_remove_singularity(V, R(49.99999999999981), R(2.6713760659695544e-06), [&](const R V) { return ((((R(1) * cmx_exp(((R(50.0) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))) * cmx_div(((L_type_Ca_channel__FrICa * L_type_Ca_channel__P_CaNa * p[2] * y[6] * y[7] * g[1] * (V - R(50.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (-cmx_expm1((((-(V - R(50.0))) * membrane__F) * (R(1) / (membrane__R * membrane__T)))))))); });
    const R d_reversal_potentials__E_Ca__d6 = ((cmx_div(cmx_div((-(extracellular_calcium_concentration__Ca_o * R(1))), (g[6] * g[6])), cmx_div(extracellular_calcium_concentration__Ca_o, g[6])) * ((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_calcium_background_current__i_b_Ca__d6 = (((-d_reversal_potentials__E_Ca__d6) * calcium_background_current__g_bca));
    const R d_calcium_release__CadsReg__d7 = cmx_div((R(1) * (g[7] + calcium_release__K_m_Ca_ds) - g[7] * (R(1))), ((g[7] + calcium_release__K_m_Ca_ds) * (g[7] + calcium_release__K_m_Ca_ds)));
    const R d_calcium_release__CaiReg__d6 = cmx_div((R(1) * (g[6] + calcium_release__K_m_Ca_cyt) - g[6] * (R(1))), ((g[6] + calcium_release__K_m_Ca_cyt) * (g[6] + calcium_release__K_m_Ca_cyt)));
    const R d_calcium_release__RegBindSite__d6 = (d_calcium_release__CaiReg__d6 + (((-d_calcium_release__CaiReg__d6) * calcium_release__CadsReg)));
    const R d_calcium_release__RegBindSite__d7 = (((d_calcium_release__CadsReg__d7 * (R(1.0) - calcium_release__CaiReg))));
    const R d_calcium_release__ActRate__d6 = ((((R(2.0) * cmx_ipow<1>(calcium_release__RegBindSite) * d_calcium_release__RegBindSite__d6) * R(500.0))));
    const R d_calcium_release__ActRate__d7 = ((((R(2.0) * cmx_ipow<1>(calcium_release__RegBindSite) * d_calcium_release__RegBindSite__d7) * R(500.0))));
    const R d_calcium_release__InactRate__d6 = ((((R(2.0) * cmx_ipow<1>(calcium_release__RegBindSite) * d_calcium_release__RegBindSite__d6) * R(500.0))));
    const R d_calcium_release__InactRate__d7 = ((((R(2.0) * cmx_ipow<1>(calcium_release__RegBindSite) * d_calcium_release__RegBindSite__d7) * R(500.0))));
    const R d_calcium_release__PrecFrac__d2 = (-R(1));
    const R d_calcium_release__PrecFrac__d3 = (-R(1));
    const R d_calcium_release__i_rel__d2 = ((((((R(2.0) * cmx_ipow<1>(cmx_div(g[2], (g[2] + R(0.25)))) * cmx_div((R(1) * (g[2] + R(0.25)) - g[2] * (R(1))), ((g[2] + R(0.25)) * (g[2] + R(0.25))))) * calcium_release__K_m_rel))) * g[9]));
    const R d_calcium_release__i_rel__d9 = ((R(1) * ((cmx_ipow<2>(cmx_div(g[2], (g[2] + R(0.25)))) * calcium_release__K_m_rel) + calcium_release__K_leak_rate)));
    const R d_calcium_translocation__i_trans__d8 = ((R(1) * R(50.0)));
    const R d_calcium_translocation__i_trans__d9 = (((-R(1)) * R(50.0)));
    const R d_reversal_potentials__E_mh__d4 = ((cmx_div(cmx_div((-((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)) * (R(1)))), ((g[4] + (R(0.12) * g[5])) * (g[4] + (R(0.12) * g[5])))), cmx_div((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)), (g[4] + (R(0.12) * g[5])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_reversal_potentials__E_mh__d5 = ((cmx_div(cmx_div((-((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)) * (((R(1) * R(0.12)))))), ((g[4] + (R(0.12) * g[5])) * (g[4] + (R(0.12) * g[5])))), cmx_div((extracellular_sodium_concentration__Na_o + (R(0.12) * extracellular_potassium_concentration__K_o)), (g[4] + (R(0.12) * g[5])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_fast_sodium_current__i_Na__d4 = (((-d_reversal_potentials__E_mh__d4) * p[0] * cmx_ipow<3>(y[4]) * y[5]));
    const R d_fast_sodium_current__i_Na__d5 = (((-d_reversal_potentials__E_mh__d5) * p[0] * cmx_ipow<3>(y[4]) * y[5]));
    const R d_intracellular_calcium_concentration__d_Ca_Calmod_dt__d6 = ((R(1) * intracellular_calcium_concentration__alpha_Calmod * (intracellular_calcium_concentration__Calmod - g[10])));
    const R d_intracellular_calcium_concentration__d_Ca_Calmod_dt__d10 = ((((-R(1)) * intracellular_calcium_concentration__alpha_Calmod * g[6])) - ((R(1) * intracellular_calcium_concentration__beta_Calmod)));
    const R d_intracellular_calcium_concentration__d_Ca_Trop_dt__d6 = ((R(1) * intracellular_calcium_concentration__alpha_Trop * (intracellular_calcium_concentration__Trop - g[11])));
    const R d_intracellular_calcium_concentration__d_Ca_Trop_dt__d11 = ((((-R(1)) * intracellular_calcium_concentration__alpha_Trop * g[6])) - ((R(1) * intracellular_calcium_concentration__beta_Trop)));
    const R d_reversal_potentials__E_Na__d4 = ((cmx_div(cmx_div((-(extracellular_sodium_concentration__Na_o * R(1))), (g[4] * g[4])), cmx_div(extracellular_sodium_concentration__Na_o, g[4])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_persistent_sodium_current__i_p_Na__d4 = (((-d_reversal_potentials__E_Na__d4) * cmx_div((persistent_sodium_current__g_pna * R(1.0)), (R(1.0) + cmx_exp_v(((-(V + R(52.0))) * (R(1) / R(8.0))), V, -0.18033688011112042, -9.377517765778261)))));
    const R d_reversal_potentials__E_K__d5 = ((cmx_div(cmx_div((-(extracellular_potassium_concentration__K_o * R(1))), (g[5] * g[5])), cmx_div(extracellular_potassium_concentration__K_o, g[5])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_rapid_delayed_rectifier_potassium_current__i_Kr__d5 = (((-d_reversal_potentials__E_K__d5) * cmx_div((((p[4] * y[1]) + (rapid_delayed_rectifier_potassium_current__g_Kr2 * y[2])) * R(1.0)), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003)))));
    const R d_sarcoplasmic_reticulum_calcium_pump__K_2__d6 = (R(1));
    const R d_sarcoplasmic_reticulum_calcium_pump__K_2__d8 = (((R(1) * sarcoplasmic_reticulum_calcium_pump__K_1)));
    const R d_sarcoplasmic_reticulum_calcium_pump__i_up__d6 = (((cmx_div((R(1) * sarcoplasmic_reticulum_calcium_pump__K_2 - g[6] * d_sarcoplasmic_reticulum_calcium_pump__K_2__d6), (sarcoplasmic_reticulum_calcium_pump__K_2 * sarcoplasmic_reticulum_calcium_pump__K_2)) * sarcoplasmic_reticulum_calcium_pump__alpha_up)) - ((cmx_div((-((g[8] * sarcoplasmic_reticulum_calcium_pump__K_1) * d_sarcoplasmic_reticulum_calcium_pump__K_2__d6)), (sarcoplasmic_reticulum_calcium_pump__K_2 * sarcoplasmic_reticulum_calcium_pump__K_2)) * sarcoplasmic_reticulum_calcium_pump__beta_up)));
    const R d_sarcoplasmic_reticulum_calcium_pump__i_up__d8 = (((cmx_div((-(g[6] * d_sarcoplasmic_reticulum_calcium_pump__K_2__d8)), (sarcoplasmic_reticulum_calcium_pump__K_2 * sarcoplasmic_reticulum_calcium_pump__K_2)) * sarcoplasmic_reticulum_calcium_pump__alpha_up)) - ((cmx_div((((R(1) * sarcoplasmic_reticulum_calcium_pump__K_1)) * sarcoplasmic_reticulum_calcium_pump__K_2 - (g[8] * sarcoplasmic_reticulum_calcium_pump__K_1) * d_sarcoplasmic_reticulum_calcium_pump__K_2__d8), (sarcoplasmic_reticulum_calcium_pump__K_2 * sarcoplasmic_reticulum_calcium_pump__K_2)) * sarcoplasmic_reticulum_calcium_pump__beta_up)));
    const R d_reversal_potentials__E_Ks__d4 = ((cmx_div(cmx_div((-((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)) * (((R(1) * reversal_potentials__P_kna))))), ((g[5] + (reversal_potentials__P_kna * g[4])) * (g[5] + (reversal_potentials__P_kna * g[4])))), cmx_div((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)), (g[5] + (reversal_potentials__P_kna * g[4])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_reversal_potentials__E_Ks__d5 = ((cmx_div(cmx_div((-((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)) * (R(1)))), ((g[5] + (reversal_potentials__P_kna * g[4])) * (g[5] + (reversal_potentials__P_kna * g[4])))), cmx_div((extracellular_potassium_concentration__K_o + (reversal_potentials__P_kna * extracellular_sodium_concentration__Na_o)), (g[5] + (reversal_potentials__P_kna * g[4])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_slow_delayed_rectifier_potassium_current__i_Ks__d4 = (((-d_reversal_potentials__E_Ks__d4) * p[3] * cmx_ipow<2>(y[3])));
    const R d_slow_delayed_rectifier_potassium_current__i_Ks__d5 = (((-d_reversal_potentials__E_Ks__d5) * p[3] * cmx_ipow<2>(y[3])));
    const R d_sodium_background_current__i_b_Na__d4 = (((-d_reversal_potentials__E_Na__d4) * sodium_background_current__g_bna));
    const R d_sodium_calcium_exchanger__i_NaCa_cyt__d4 = cmx_div(((((((sodium_calcium_exchanger__n_NaCa * pow(g[4], sodium_calcium_exchanger__n_NaCa - R(1)) * R(1)) * cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * extracellular_calcium_concentration__Ca_o)) * (R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa)) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))) - ((R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[6]))) * (((((((((sodium_calcium_exchanger__n_NaCa * pow(g[4], sodium_calcium_exchanger__n_NaCa - R(1)) * R(1)) * extracellular_calcium_concentration__Ca_o))) * sodium_calcium_exchanger__d_NaCa))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))))), (((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069)))))));
    const R d_sodium_calcium_exchanger__i_NaCa_cyt__d6 = cmx_div(((((-((R(1) * cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)))) * (R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa)) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))) - ((R(1.0) - sodium_calcium_exchanger__FRiNaCa) * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[6]))) * ((((((((R(1) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)))) * sodium_calcium_exchanger__d_NaCa))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))) + ((cmx_div(R(1), R(0.0069))) * (R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa)))))))), (((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069))))) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[6] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[6] * (R(1) / R(0.0069)))))));
    const R d_sodium_calcium_exchanger__i_NaCa_ds__d4 = cmx_div(((((((sodium_calcium_exchanger__n_NaCa * pow(g[4], sodium_calcium_exchanger__n_NaCa - R(1)) * R(1)) * cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * extracellular_calcium_concentration__Ca_o)) * sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa)) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))) - (sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[7]))) * (((((((((sodium_calcium_exchanger__n_NaCa * pow(g[4], sodium_calcium_exchanger__n_NaCa - R(1)) * R(1)) * extracellular_calcium_concentration__Ca_o))) * sodium_calcium_exchanger__d_NaCa))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))))), (((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069)))))));
    const R d_sodium_calcium_exchanger__i_NaCa_ds__d7 = cmx_div(((((-((R(1) * cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)))) * sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa)) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))) - (sodium_calcium_exchanger__FRiNaCa * sodium_calcium_exchanger__k_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger__gamma * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.02700284432557688, 0.0) * pow(g[4], sodium_calcium_exchanger__n_NaCa) * extracellular_calcium_concentration__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger__gamma - R(1.0)) * (sodium_calcium_exchanger__n_NaCa - R(2.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.02700284432557688, -0.0) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa) * g[7]))) * ((((((((R(1) * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)))) * sodium_calcium_exchanger__d_NaCa))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))) + ((cmx_div(R(1), R(0.0069))) * (R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa)))))))), (((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069))))) * ((R(1.0) + (sodium_calcium_exchanger__d_NaCa * ((g[7] * pow(extracellular_sodium_concentration__Na_o, sodium_calcium_exchanger__n_NaCa)) + (extracellular_calcium_concentration__Ca_o * pow(g[4], sodium_calcium_exchanger__n_NaCa))))) * (R(1.0) + (g[7] * (R(1) / R(0.0069)))))));
    const R d_sodium_potassium_pump__i_NaK__d4 = cmx_div((((R(1) * ((sodium_potassium_pump__i_NaK_max * extracellular_potassium_concentration__K_o) * (R(1) / (sodium_potassium_pump__K_mK + extracellular_potassium_concentration__K_o))))) * (sodium_potassium_pump__K_mNa + g[4]) - (((sodium_potassium_pump__i_NaK_max * extracellular_potassium_concentration__K_o) * (R(1) / (sodium_potassium_pump__K_mK + extracellular_potassium_concentration__K_o))) * g[4]) * (R(1))), ((sodium_potassium_pump__K_mNa + g[4]) * (sodium_potassium_pump__K_mNa + g[4])));
    const R d_time_independent_potassium_current__i_K1__d5 = cmx_div(((((-d_reversal_potentials__E_K__d5) * ((p[1] * extracellular_potassium_concentration__K_o) * (R(1) / (extracellular_potassium_concentration__K_o + time_independent_potassium_current__K_mk1))))) * (R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07)) - (((p[1] * extracellular_potassium_concentration__K_o) * (R(1) / (extracellular_potassium_concentration__K_o + time_independent_potassium_current__K_mk1))) * (V - reversal_potentials__E_K)) * ((cmx_exp(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T)))) * cmx_div((((-d_reversal_potentials__E_K__d5) * membrane__F * R(1.25))), (membrane__R * membrane__T))))), ((R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07)) * (R(1.0) + cmx_exp_c(((((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)) * (R(1) / (membrane__R * membrane__T))), (((V - reversal_potentials__E_K) - R(10.0)) * membrane__F * R(1.25)), 5.597294657567624e-07))));
    const R d_transient_outward_current__i_to__d5 = (((-d_reversal_potentials__E_K__d5) * transient_outward_current__g_to * (transient_outward_current__g_tos + (y[10] * (R(1.0) - transient_outward_current__g_tos))) * y[11]));
    J[0] = R(1) - dt * cmx_div((-(((R(1)) * R(1.0)))), R(1000.0));
    J[1] = R(0);
    J[2] = R(0);
    J[3] = R(0);
    J[4] = R(0);
    J[5] = R(0);
    J[6] = R(0) - dt * cmx_div((-(((cmx_div((R(1) * (L_type_Ca_channel__Km_f2 + g[6]) - g[6] * (R(1))), ((L_type_Ca_channel__Km_f2 + g[6]) * (L_type_Ca_channel__Km_f2 + g[6])))) * R(1.0)))), R(1000.0));
    J[7] = R(0);
    J[8] = R(0);
    J[9] = R(0);
    J[10] = R(0);
    J[11] = R(0);
    J[12] = R(0);
    J[13] = R(1) - dt * cmx_div((((-(R(1))) * L_type_Ca_channel__R_decay)), R(1000.0));
    J[14] = R(0);
    J[15] = R(0);
    J[16] = R(0);
    J[17] = R(0);
    J[18] = R(0);
    J[19] = R(0) - dt * cmx_div((((-(cmx_div((R(1) * (L_type_Ca_channel__Km_f2ds + g[7]) - g[7] * (R(1))), ((L_type_Ca_channel__Km_f2ds + g[7]) * (L_type_Ca_channel__Km_f2ds + g[7]))))) * L_type_Ca_channel__R_decay)), R(1000.0));
    J[20] = R(0);
    J[21] = R(0);
    J[22] = R(0);
    J[23] = R(0);
    J[24] = R(0);
    J[25] = R(0);
    J[26] = R(1) - dt * cmx_div((((d_calcium_release__PrecFrac__d2 * calcium_release__SpeedRel * calcium_release__ActRate)) - ((R(1) * calcium_release__SpeedRel * calcium_release__InactRate))), R(1000.0));
    J[27] = R(0) - dt * cmx_div(((d_calcium_release__PrecFrac__d3 * calcium_release__SpeedRel * calcium_release__ActRate)), R(1000.0));
    J[28] = R(0);
    J[29] = R(0);
    J[30] = R(0) - dt * cmx_div((((d_calcium_release__ActRate__d6 * calcium_release__PrecFrac * calcium_release__SpeedRel)) - ((d_calcium_release__InactRate__d6 * g[2] * calcium_release__SpeedRel))), R(1000.0));
    J[31] = R(0) - dt * cmx_div((((d_calcium_release__ActRate__d7 * calcium_release__PrecFrac * calcium_release__SpeedRel)) - ((d_calcium_release__InactRate__d7 * g[2] * calcium_release__SpeedRel))), R(1000.0));
    J[32] = R(0);
    J[33] = R(0);
    J[34] = R(0);
    J[35] = R(0);
    J[36] = R(0);
    J[37] = R(0);
    J[38] = R(0) - dt * cmx_div(((R(1) * calcium_release__SpeedRel * calcium_release__InactRate)), R(1000.0));
    J[39] = R(1) - dt * cmx_div((-((R(1) * calcium_release__SpeedRel * R(1.0)))), R(1000.0));
    J[40] = R(0);
    J[41] = R(0);
    J[42] = R(0) - dt * cmx_div(((d_calcium_release__InactRate__d6 * g[2] * calcium_release__SpeedRel)), R(1000.0));
    J[43] = R(0) - dt * cmx_div(((d_calcium_release__InactRate__d7 * g[2] * calcium_release__SpeedRel)), R(1000.0));
    J[44] = R(0);
    J[45] = R(0);
    J[46] = R(0);
    J[47] = R(0);
    J[48] = R(0) - dt * cmx_div((((d_L_type_Ca_channel__i_Ca_L_Na_cyt__d0) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[49] = R(0) - dt * cmx_div((((d_L_type_Ca_channel__i_Ca_L_Na_ds__d1) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[50] = R(0);
    J[51] = R(0);
    J[52] = R(1) - dt * cmx_div((((d_fast_sodium_current__i_Na__d4 + d_persistent_sodium_current__i_p_Na__d4 + d_sodium_background_current__i_b_Na__d4 + ((d_sodium_potassium_pump__i_NaK__d4 * R(3.0))) + ((d_sodium_calcium_exchanger__i_NaCa_cyt__d4 * R(3.0))) + d_L_type_Ca_channel__i_Ca_L_Na_cyt__d4 + d_L_type_Ca_channel__i_Ca_L_Na_ds__d4) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[53] = R(0) - dt * cmx_div((((d_fast_sodium_current__i_Na__d5) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[54] = R(0) - dt * cmx_div((((((d_sodium_calcium_exchanger__i_NaCa_cyt__d6 * R(3.0)))) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[55] = R(0);
    J[56] = R(0);
    J[57] = R(0);
    J[58] = R(0);
    J[59] = R(0);
    J[60] = R(0) - dt * cmx_div((((d_L_type_Ca_channel__i_Ca_L_K_cyt__d0) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[61] = R(0) - dt * cmx_div((((d_L_type_Ca_channel__i_Ca_L_K_ds__d1) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[62] = R(0);
    J[63] = R(0);
    J[64] = R(0) - dt * cmx_div(((((d_slow_delayed_rectifier_potassium_current__i_Ks__d4) - ((d_sodium_potassium_pump__i_NaK__d4 * R(2.0)))) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[65] = R(1) - dt * cmx_div((((d_time_independent_potassium_current__i_K1__d5 + d_rapid_delayed_rectifier_potassium_current__i_Kr__d5 + d_slow_delayed_rectifier_potassium_current__i_Ks__d5 + d_L_type_Ca_channel__i_Ca_L_K_cyt__d5 + d_L_type_Ca_channel__i_Ca_L_K_ds__d5 + d_transient_outward_current__i_to__d5) * ((-R(1.0)) * (R(1) / (R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))), R(1000.0));
    J[66] = R(0);
    J[67] = R(0);
    J[68] = R(0);
    J[69] = R(0);
    J[70] = R(0);
    J[71] = R(0);
    J[72] = R(0) - dt * cmx_div(((((d_L_type_Ca_channel__i_Ca_L_Ca_cyt__d0) * ((-R(1.0)) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_i * membrane__F)))))), R(1000.0));
    J[73] = R(0);
    J[74] = R(0) - dt * cmx_div((cmx_div(((d_calcium_release__i_rel__d2 * intracellular_calcium_concentration__V_rel_ratio)), intracellular_calcium_concentration__V_i_ratio)), R(1000.0));
    J[75] = R(0);
    J[76] = R(0) - dt * cmx_div((((((-((d_sodium_calcium_exchanger__i_NaCa_cyt__d4 * R(2.0)))) - ((d_sodium_calcium_exchanger__i_NaCa_ds__d4 * R(2.0)))) * ((-R(1.0)) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_i * membrane__F)))))), R(1000.0));
    J[77] = R(0);
    J[78] = R(1) - dt * cmx_div(((((((((d_L_type_Ca_channel__i_Ca_L_Ca_cyt__d6 + d_calcium_background_current__i_b_Ca__d6) - ((d_sodium_calcium_exchanger__i_NaCa_cyt__d6 * R(2.0)))) * ((-R(1.0)) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_i * membrane__F)))))) - d_intracellular_calcium_concentration__d_Ca_Calmod_dt__d6) - d_intracellular_calcium_concentration__d_Ca_Trop_dt__d6) - d_sarcoplasmic_reticulum_calcium_pump__i_up__d6), R(1000.0));
    J[79] = R(0) - dt * cmx_div(((((-((d_sodium_calcium_exchanger__i_NaCa_ds__d7 * R(2.0)))) * ((-R(1.0)) * (R(1) / (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_i * membrane__F))))) + ((R(1) * intracellular_calcium_concentration__V_ds_ratio * intracellular_calcium_concentration__Kdecay))), R(1000.0));
    J[80] = R(0) - dt * cmx_div((-d_sarcoplasmic_reticulum_calcium_pump__i_up__d8), R(1000.0));
    J[81] = R(0) - dt * cmx_div((cmx_div(((d_calcium_release__i_rel__d9 * intracellular_calcium_concentration__V_rel_ratio)), intracellular_calcium_concentration__V_i_ratio)), R(1000.0));
    J[82] = R(0) - dt * cmx_div((-d_intracellular_calcium_concentration__d_Ca_Calmod_dt__d10), R(1000.0));
    J[83] = R(0) - dt * cmx_div((-d_intracellular_calcium_concentration__d_Ca_Trop_dt__d11), R(1000.0));
    J[84] = R(0);
    J[85] = R(0) - dt * cmx_div(cmx_div(((d_L_type_Ca_channel__i_Ca_L_Ca_ds__d1 * (-R(1.0)))), (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_ds_ratio * intracellular_calcium_concentration__V_i * membrane__F)), R(1000.0));
    J[86] = R(0);
    J[87] = R(0);
    J[88] = R(0);
    J[89] = R(0);
    J[90] = R(0) - dt * cmx_div(cmx_div(((d_L_type_Ca_channel__i_Ca_L_Ca_ds__d6 * (-R(1.0)))), (R(2.0) * R(1.0) * intracellular_calcium_concentration__V_ds_ratio * intracellular_calcium_concentration__V_i * membrane__F)), R(1000.0));
    J[91] = R(1) - dt * cmx_div((-((R(1) * intracellular_calcium_concentration__Kdecay))), R(1000.0));
    J[92] = R(0);
    J[93] = R(0);
    J[94] = R(0);
    J[95] = R(0);
    J[96] = R(0);
    J[97] = R(0);
    J[98] = R(0);
    J[99] = R(0);
    J[100] = R(0);
    J[101] = R(0);
    J[102] = R(0) - dt * cmx_div(((d_sarcoplasmic_reticulum_calcium_pump__i_up__d6 * (intracellular_calcium_concentration__V_i_ratio * (R(1) / intracellular_calcium_concentration__V_up_ratio)))), R(1000.0));
    J[103] = R(0);
    J[104] = R(1) - dt * cmx_div((((d_sarcoplasmic_reticulum_calcium_pump__i_up__d8 * (intracellular_calcium_concentration__V_i_ratio * (R(1) / intracellular_calcium_concentration__V_up_ratio)))) - d_calcium_translocation__i_trans__d8), R(1000.0));
    J[105] = R(0) - dt * cmx_div((-d_calcium_translocation__i_trans__d9), R(1000.0));
    J[106] = R(0);
    J[107] = R(0);
    J[108] = R(0);
    J[109] = R(0);
    J[110] = R(0) - dt * cmx_div((-d_calcium_release__i_rel__d2), R(1000.0));
    J[111] = R(0);
    J[112] = R(0);
    J[113] = R(0);
    J[114] = R(0);
    J[115] = R(0);
    J[116] = R(0) - dt * cmx_div(((d_calcium_translocation__i_trans__d8 * (intracellular_calcium_concentration__V_up_ratio * (R(1) / intracellular_calcium_concentration__V_rel_ratio)))), R(1000.0));
    J[117] = R(1) - dt * cmx_div((((d_calcium_translocation__i_trans__d9 * (intracellular_calcium_concentration__V_up_ratio * (R(1) / intracellular_calcium_concentration__V_rel_ratio)))) - d_calcium_release__i_rel__d9), R(1000.0));
    J[118] = R(0);
    J[119] = R(0);
    J[120] = R(0);
    J[121] = R(0);
    J[122] = R(0);
    J[123] = R(0);
    J[124] = R(0);
    J[125] = R(0);
    J[126] = R(0) - dt * cmx_div(((R(1) * intracellular_calcium_concentration__alpha_Calmod * (intracellular_calcium_concentration__Calmod - g[10]))), R(1000.0));
    J[127] = R(0);
    J[128] = R(0);
    J[129] = R(0);
    J[130] = R(1) - dt * cmx_div(((((-R(1)) * intracellular_calcium_concentration__alpha_Calmod * g[6])) - ((R(1) * intracellular_calcium_concentration__beta_Calmod))), R(1000.0));
    J[131] = R(0);
    J[132] = R(0);
    J[133] = R(0);
    J[134] = R(0);
    J[135] = R(0);
    J[136] = R(0);
    J[137] = R(0);
    J[138] = R(0) - dt * cmx_div(((R(1) * intracellular_calcium_concentration__alpha_Trop * (intracellular_calcium_concentration__Trop - g[11]))), R(1000.0));
    J[139] = R(0);
    J[140] = R(0);
    J[141] = R(0);
    J[142] = R(0);
    J[143] = R(1) - dt * cmx_div(((((-R(1)) * intracellular_calcium_concentration__alpha_Trop * g[6])) - ((R(1) * intracellular_calcium_concentration__beta_Trop))), R(1000.0));
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

}}  // namespace cmx::n98
