// Generated by chaste_amd/codegen/cellml2cpp.py from Maleckar2008.cellml -- do not edit.
// Model: Maleckar2008
#pragma once
#include "model_common.h"

namespace cmx { namespace maleckar2008 {

constexpr int kNumStates = 30;
constexpr int kVIndex = 0;
constexpr int kNumParams = 6;
constexpr bool kPerCapacitanceUnits = true;

// State variables (CellML component__variable), V first then document order.
static const char* const kStateNames[kNumStates] = {
    "membrane__V",
    "sodium_current_m_gate__m",
    "sodium_current_h1_gate__h1",
    "sodium_current_h2_gate__h2",
    "L_type_Ca_channel_d_L_gate__d_L",
    "L_type_Ca_channel_f_L1_gate__f_L1",
    "L_type_Ca_channel_f_L2_gate__f_L2",
    "Ca_independent_transient_outward_K_current_r_gate__r",
    "Ca_independent_transient_outward_K_current_s_gate__s",
    "ultra_rapid_K_current_aur_gate__a_ur",
    "ultra_rapid_K_current_iur_gate__i_ur",
    "delayed_rectifier_K_currents_n_gate__n",
    "delayed_rectifier_K_currents_pa_gate__pa",
    "intracellular_ion_concentrations__Na_i",
    "intracellular_ion_concentrations__Ca_i",
    "intracellular_ion_concentrations__K_i",
    "intracellular_ion_concentrations__Ca_d",
    "intracellular_Ca_buffering__O_C",
    "intracellular_Ca_buffering__O_TC",
    "intracellular_Ca_buffering__O_TMgC",
    "intracellular_Ca_buffering__O_TMgMg",
    "intracellular_Ca_buffering__O",
    "cleft_space_ion_concentrations__Na_c",
    "cleft_space_ion_concentrations__Ca_c",
    "cleft_space_ion_concentrations__K_c",
    "Ca_handling_by_the_SR__Ca_rel",
    "Ca_handling_by_the_SR__Ca_up",
    "Ca_handling_by_the_SR__O_Calse",
    "Ca_handling_by_the_SR__F1",
    "Ca_handling_by_the_SR__F2",
};
static constexpr double kInitialState[kNumStates] = {
    -73.941851,  // membrane__V
    0.003325,  // sodium_current_m_gate__m
    0.875262,  // sodium_current_h1_gate__h1
    0.870692,  // sodium_current_h2_gate__h2
    1.4e-05,  // L_type_Ca_channel_d_L_gate__d_L
    0.998578,  // L_type_Ca_channel_f_L1_gate__f_L1
    0.998561,  // L_type_Ca_channel_f_L2_gate__f_L2
    0.001098,  // Ca_independent_transient_outward_K_current_r_gate__r
    0.948202,  // Ca_independent_transient_outward_K_current_s_gate__s
    0.000371,  // ultra_rapid_K_current_aur_gate__a_ur
    0.966869,  // ultra_rapid_K_current_iur_gate__i_ur
    0.004661,  // delayed_rectifier_K_currents_n_gate__n
    5.4e-05,  // delayed_rectifier_K_currents_pa_gate__pa
    8.488527,  // intracellular_ion_concentrations__Na_i
    6.5e-05,  // intracellular_ion_concentrations__Ca_i
    129.502075,  // intracellular_ion_concentrations__K_i
    7.1e-05,  // intracellular_ion_concentrations__Ca_d
    0.026604,  // intracellular_Ca_buffering__O_C
    0.012843,  // intracellular_Ca_buffering__O_TC
    0.190077,  // intracellular_Ca_buffering__O_TMgC
    0.714719,  // intracellular_Ca_buffering__O_TMgMg
    1.38222,  // intracellular_Ca_buffering__O
    130.019282,  // cleft_space_ion_concentrations__Na_c
    1.814418,  // cleft_space_ion_concentrations__Ca_c
    5.588239,  // cleft_space_ion_concentrations__K_c
    0.630471,  // Ca_handling_by_the_SR__Ca_rel
    0.646226,  // Ca_handling_by_the_SR__Ca_up
    0.43071,  // Ca_handling_by_the_SR__O_Calse
    0.45453,  // Ca_handling_by_the_SR__F1
    0.002665,  // Ca_handling_by_the_SR__F2
};
static const char* const kParamNames[6] = {
    "sodium_current__ScaleFactorGna",
    "L_type_Ca_channel__ScaleFactorGCaL",
    "delayed_rectifier_K_currents__ScaleFactorGkr",
    "delayed_rectifier_K_currents__ScaleFactorGks",
    "Ca_independent_transient_outward_K_current__ScaleFactorIto",
    "ultra_rapid_K_current__ScaleFactorGKur",
};
// The names the reference gives the same parameters (the CellML oxmeta term,
// else the variable's cmeta:id; empty when the file annotates neither).
static const char* const kParamAliases[6] = {
    "ScaleFactorGna",
    "ScaleFactorGCaL",
    "ScaleFactorGkr",
    "ScaleFactorGks",
    "ScaleFactorIto",
    "ScaleFactorGKur",
};
static constexpr double kParamDefaults[6] = {
    1.0,
    1.0,
    1.0,
    1.0,
    1.0,
    1.0,
};
// 1 = integrated by Rush-Larsen (affine gate, V-dependent coefficients only)
static constexpr int kIsGate[kNumStates] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

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
    const R membrane__R = R(8314.0);
    const R membrane__T = R(306.15);
    const R membrane__F = R(96487.0);
    const R membrane__Cm = R(0.05);
    const R sodium_current__P_Na = R(0.0018);
    const R L_type_Ca_channel__g_Ca_L = R(6.75);
    const R L_type_Ca_channel__E_Ca_app = R(60.0);
    const R L_type_Ca_channel__k_Ca = R(0.025);
    const R Ca_independent_transient_outward_K_current__g_t = R(8.25);
    const R ultra_rapid_K_current__g_kur = R(2.25);
    const R inward_rectifier__g_K1 = R(3.1);
    const R inward_rectifier__ScaleFactorGK1 = R(1.0);
    const R delayed_rectifier_K_currents__g_Ks = R(1.0);
    const R delayed_rectifier_K_currents__g_Kr = R(0.5);
    const R delayed_rectifier_K_currents_pa_gate__AZD = R(0.0);
    const R background_currents__g_B_Na = R(0.060599);
    const R background_currents__g_B_Ca = R(0.078681);
    const R sodium_potassium_pump__K_NaK_K = R(1.0);
    const R sodium_potassium_pump__i_NaK_max = R(68.55);
    const R sodium_potassium_pump__ScaleFactorGNaK = R(1.0);
    const R sodium_potassium_pump__pow_K_NaK_Na_15 = R(36.4829);
    const R sarcolemmal_calcium_pump_current__i_CaP_max = R(4.0);
    const R sarcolemmal_calcium_pump_current__k_CaP = R(0.0002);
    const R Na_Ca_ion_exchanger_current__ScaleFactorGNaCa = R(1.0);
    const R Na_Ca_ion_exchanger_current__K_NaCa = R(0.0374842);
    const R Na_Ca_ion_exchanger_current__d_NaCa = R(0.0003);
    const R Na_Ca_ion_exchanger_current__gamma_Na = R(0.45);
    const R ACh_dependent_K_current__ACh = R(1e-24);
    const R intracellular_ion_concentrations__phi_Na_en = R(0.0);
    const R intracellular_ion_concentrations__Vol_i = R(0.005884);
    const R intracellular_ion_concentrations__Vol_d = R(0.00011768);
    const R intracellular_ion_concentrations__tau_di = R(0.01);
    const R intracellular_Ca_buffering__Mg_i = R(2.5);
    const R cleft_space_ion_concentrations__Vol_c = R(0.000800224);
    const R cleft_space_ion_concentrations__tau_Na = R(14.3);
    const R cleft_space_ion_concentrations__tau_K = R(10.0);
    const R cleft_space_ion_concentrations__tau_Ca = R(24.7);
    const R cleft_space_ion_concentrations__Na_b = R(130.0);
    const R cleft_space_ion_concentrations__Ca_b = R(1.8);
    const R cleft_space_ion_concentrations__K_b = R(5.4);
    const R Ca_handling_by_the_SR__I_up_max = R(2800.0);
    const R Ca_handling_by_the_SR__k_cyca = R(0.0003);
    const R Ca_handling_by_the_SR__k_srca = R(0.5);
    const R Ca_handling_by_the_SR__k_xcs = R(0.4);
    const R Ca_handling_by_the_SR__alpha_rel = R(200000.0);
    const R Ca_handling_by_the_SR__Vol_up = R(0.0003969);
    const R Ca_handling_by_the_SR__Vol_rel = R(4.41e-05);
    const R Ca_handling_by_the_SR__r_recov = R(0.815);
    const R Ca_handling_by_the_SR__tau_tr = R(0.01);
    const R Ca_handling_by_the_SR__k_rel_i = R(0.0003);
    const R Ca_handling_by_the_SR__k_rel_d = R(0.003);
    const R i_stim = (i_stim_area / cm) * R(50.00000000000001);
    (void)i_stim;
    const R Ca_independent_transient_outward_K_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(y[24], y[15])));
    const R ACh_dependent_K_current__i_KACh = ((R(10000.0) * (R(1) / (R(1.0) + ((R(9.13652) * pow(R(1.0), R(0.477811))) * (R(1) / pow(ACh_dependent_K_current__ACh, R(0.477811))))))) * (R(0.0517) + cmx_div(R(0.4516), (R(1.0) + cmx_exp_v(((V + R(59.53)) * (R(1) / R(17.18))), V, 0.08397526431251243, 4.999047484523865)))) * (V - Ca_independent_transient_outward_K_current__E_K) * membrane__Cm);
    const R Ca_handling_by_the_SR__J_O_Calse = ((R(480.0) * y[25] * (R(1.0) - y[27])) - (R(400.0) * y[27]));
    const R Ca_handling_by_the_SR__i_rel_f2 = cmx_div(y[29], (y[29] + R(0.25)));
    const R Ca_handling_by_the_SR__i_rel_factor = (Ca_handling_by_the_SR__i_rel_f2 * Ca_handling_by_the_SR__i_rel_f2);
    const R Ca_handling_by_the_SR__i_rel = (Ca_handling_by_the_SR__alpha_rel * Ca_handling_by_the_SR__i_rel_factor * (y[25] - y[14]));
    const R Ca_handling_by_the_SR__i_tr = (((y[26] - y[25]) * R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F) * (R(1) / Ca_handling_by_the_SR__tau_tr));
    const R Ca_handling_by_the_SR__i_up = cmx_div((Ca_handling_by_the_SR__I_up_max * ((y[14] * (R(1) / Ca_handling_by_the_SR__k_cyca)) - ((Ca_handling_by_the_SR__k_xcs * Ca_handling_by_the_SR__k_xcs * y[26]) * (R(1) / Ca_handling_by_the_SR__k_srca)))), (((y[14] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (y[26] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca))));
    const R Ca_handling_by_the_SR__r_Ca_d_term = cmx_div(y[16], (y[16] + Ca_handling_by_the_SR__k_rel_d));
    const R Ca_handling_by_the_SR__r_Ca_d_factor = (Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term);
    const R Ca_handling_by_the_SR__r_Ca_i_term = cmx_div(y[14], (y[14] + Ca_handling_by_the_SR__k_rel_i));
    const R Ca_handling_by_the_SR__r_Ca_i_factor = (Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term);
    const R Ca_handling_by_the_SR__r_act = (R(203.8) * (Ca_handling_by_the_SR__r_Ca_i_factor + Ca_handling_by_the_SR__r_Ca_d_factor));
    const R Ca_handling_by_the_SR__r_inact = (R(33.96) + (R(339.6) * Ca_handling_by_the_SR__r_Ca_i_factor));
    const R Ca_independent_transient_outward_K_current__i_t = (p[4] * Ca_independent_transient_outward_K_current__g_t * y[7] * y[8] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R Ca_independent_transient_outward_K_current_r_gate__r_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(1.0)) * (R(1) / (-R(11.0)))), V, -0.1311540946262694, 0.1311540946262694)));
    const R Ca_independent_transient_outward_K_current_r_gate__tau_r = ((R(0.0035) * cmx_exp_c(((((-V) * V) * (R(1) / R(30.0))) * (R(1) / R(30.0))), (((-V) * V) * (R(1) / R(30.0))), 0.04808983469629878)) + R(0.0015));
    const R Ca_independent_transient_outward_K_current_s_gate__s_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(40.5)) * (R(1) / R(11.5))), V, 0.12545174268599682, 5.080795578782872)));
    const R Ca_independent_transient_outward_K_current_s_gate__s_factor = ((V + R(52.45)) * (R(1) / R(15.8827)));
    const R Ca_independent_transient_outward_K_current_s_gate__tau_s = ((R(0.025635) * cmx_exp(((-Ca_independent_transient_outward_K_current_s_gate__s_factor) * Ca_independent_transient_outward_K_current_s_gate__s_factor))) + R(0.01414));
    const R L_type_Ca_channel__f_Ca = cmx_div(y[16], (y[16] + L_type_Ca_channel__k_Ca));
    const R L_type_Ca_channel__i_Ca_L = (p[1] * L_type_Ca_channel__g_Ca_L * y[4] * ((L_type_Ca_channel__f_Ca * y[5]) + ((R(1.0) - L_type_Ca_channel__f_Ca) * y[6])) * (V - L_type_Ca_channel__E_Ca_app));
    const R L_type_Ca_channel_d_L_gate__d_L_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / (-R(5.8)))), V, -0.2487405242912006, -2.238664718620805)));
    const R L_type_Ca_channel_d_L_gate__d_L_factor = ((V + R(35.0)) * (R(1) / R(30.0)));
    const R L_type_Ca_channel_d_L_gate__tau_d_L = ((R(0.0027) * cmx_exp(((-L_type_Ca_channel_d_L_gate__d_L_factor) * L_type_Ca_channel_d_L_gate__d_L_factor))) + R(0.002));
    const R L_type_Ca_channel_f_L1_gate__f_L_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(27.4)) * (R(1) / R(7.1))), V, 0.20319648463224838, 5.567583678923605)));
    const R L_type_Ca_channel_f_L1_gate__f_L_factor = (V + R(40.0));
    const R L_type_Ca_channel_f_L1_gate__tau_f_L1 = ((R(0.161) * cmx_exp_c(((((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.4))) * (R(1) / R(14.4))), (((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.4))), 0.10018715561728912)) + R(0.01));
    const R L_type_Ca_channel_f_L2_gate__tau_f_L2 = ((R(1.3323) * cmx_exp_c(((((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.2))) * (R(1) / R(14.2))), (((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.2))), 0.10159824231612419)) + R(0.0626));
    const R Na_Ca_ion_exchanger_current__i_NaCa = (Na_Ca_ion_exchanger_current__ScaleFactorGNaCa * cmx_div((Na_Ca_ion_exchanger_current__K_NaCa * ((y[13] * y[13] * y[13] * y[23] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (y[22] * y[22] * y[22] * y[14] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))), (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((y[22] * y[22] * y[22] * y[14]) + (y[13] * y[13] * y[13] * y[23]))))));
    const R background_currents__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(y[23], y[14])));
    const R background_currents__i_B_Ca = (background_currents__g_B_Ca * (V - background_currents__E_Ca));
    const R sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(y[22], y[13])));
    const R background_currents__i_B_Na = (background_currents__g_B_Na * (V - sodium_current__E_Na));
    const R delayed_rectifier_K_currents__factorAZD = ((R(1.0) - (R(1.0) * (R(1) / (R(1.0) + cmx_exp(((-(delayed_rectifier_K_currents_pa_gate__AZD - R(0.6))) * (R(1) / R(0.12)))))))) * (R(1) / R(0.9933)));
    const R delayed_rectifier_K_currents_pi_gate__pip = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(55.0)) * (R(1) / R(24.0))), V, 0.06011229337037347, 3.306176135370541)));
    const R delayed_rectifier_K_currents__i_Kr = (p[2] * delayed_rectifier_K_currents__g_Kr * delayed_rectifier_K_currents__factorAZD * y[12] * delayed_rectifier_K_currents_pi_gate__pip * (V - Ca_independent_transient_outward_K_current__E_K));
    const R delayed_rectifier_K_currents__i_Ks = (p[3] * delayed_rectifier_K_currents__g_Ks * y[11] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R delayed_rectifier_K_currents_n_gate__n_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(19.9)) * (R(1) / (-R(12.7)))), V, -0.1135980347156664, 2.2606008908417614)));
    const R delayed_rectifier_K_currents_n_gate__n_factor = ((V - R(20.0)) * (R(1) / R(20.0)));
    const R delayed_rectifier_K_currents_n_gate__tau_n = (R(0.7) + (R(0.4) * cmx_exp(((-delayed_rectifier_K_currents_n_gate__n_factor) * delayed_rectifier_K_currents_n_gate__n_factor))));
    const R delayed_rectifier_K_currents_pa_gate__p_a_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(15.0) + (R(10.0) * delayed_rectifier_K_currents_pa_gate__AZD)) * (R(1) / (-R(6.0)))), V, -0.2404491734814939, -3.6067376022224087)));
    const R delayed_rectifier_K_currents_pa_gate__pa_factor = ((V + R(20.1376)) * (R(1) / R(22.1996)));
    const R delayed_rectifier_K_currents_pa_gate__tau_pa = (R(0.03118) + (R(0.21718) * cmx_exp(((-delayed_rectifier_K_currents_pa_gate__pa_factor) * delayed_rectifier_K_currents_pa_gate__pa_factor))));
    const R intracellular_Ca_buffering__J_O_C = ((R(200000.0) * y[14] * (R(1.0) - y[17])) - (R(476.0) * y[17]));
    const R intracellular_Ca_buffering__J_O_TC = ((R(78400.0) * y[14] * (R(1.0) - y[18])) - (R(392.0) * y[18]));
    const R intracellular_Ca_buffering__J_O_TMgC = ((R(200000.0) * y[14] * ((R(1.0) - y[19]) - y[20])) - (R(6.6) * y[19]));
    const R intracellular_Ca_buffering__J_O = ((R(0.08) * intracellular_Ca_buffering__J_O_TC) + (R(0.16) * intracellular_Ca_buffering__J_O_TMgC) + (R(0.045) * intracellular_Ca_buffering__J_O_C));
    const R intracellular_Ca_buffering__J_O_TMgMg = ((R(2000.0) * intracellular_Ca_buffering__Mg_i * ((R(1.0) - y[19]) - y[20])) - (R(666.0) * y[20]));
    const R intracellular_ion_concentrations__i_di = (((y[16] - y[14]) * R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F) * (R(1) / intracellular_ion_concentrations__tau_di));
    const R inward_rectifier__i_K1 = cmx_div((inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((y[24] * (R(1) / R(1.0))), R(0.4457)) * (V - Ca_independent_transient_outward_K_current__E_K)), (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)));
    const R sarcolemmal_calcium_pump_current__i_CaP = cmx_div((sarcolemmal_calcium_pump_current__i_CaP_max * y[14]), (y[14] + sarcolemmal_calcium_pump_current__k_CaP));
    const R sodium_current__i_Na = cmx_remove_singularity(V, R(-0.0), R(2.6380041870925613e-06), [&](const R V) { return cmx_div((((p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * y[22] * V * membrane__F * membrane__F) * (R(1) / (membrane__R * membrane__T))) * cmx_expm1((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R sodium_current_h1_gate__h_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(63.6)) * (R(1) / R(5.3))), V, 0.27220661148848363, 17.31234049066756)));
    const R sodium_current_h1_gate__h_factor = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(35.1)) * (R(1) / R(3.2))), V, 0.4508422002778011, 15.824561229750817)));
    const R sodium_current_h1_gate__tau_h1 = ((R(0.03) * sodium_current_h1_gate__h_factor) + R(0.0003));
    const R sodium_current_h2_gate__tau_h2 = ((R(0.12) * sodium_current_h1_gate__h_factor) + R(0.003));
    const R sodium_current_m_gate__m_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(27.12)) * (R(1) / (-R(8.21)))), V, -0.1757241219109578, -4.765638186225175)));
    const R sodium_current_m_gate__m_factor = ((V + R(25.57)) * (R(1) / R(28.8)));
    const R sodium_current_m_gate__tau_m = ((R(4.2e-05) * cmx_exp(((-sodium_current_m_gate__m_factor) * sodium_current_m_gate__m_factor))) + R(2.4e-05));
    const R sodium_potassium_pump__pow_Na_i_15 = pow(y[13], R(1.5));
    const R sodium_potassium_pump__i_NaK = cmx_div((cmx_div((cmx_div((sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * y[24]), (y[24] + sodium_potassium_pump__K_NaK_K)) * sodium_potassium_pump__pow_Na_i_15), (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15)) * (V + R(150.0))), (V + R(200.0)));
    const R ultra_rapid_K_current__i_Kur = (p[5] * ultra_rapid_K_current__g_kur * y[9] * y[10] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R ultra_rapid_K_current_aur_gate__a_ur_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(6.0))) * (R(1) / R(8.6))), V, -0.16775523731267017, -1.006531423876021)));
    const R ultra_rapid_K_current_aur_gate__tau_a_ur = (cmx_div(R(0.009), (R(1.0) + cmx_exp_v(((V + R(5.0)) * (R(1) / R(12.0))), V, 0.12022458674074694, 0.6011229337037348))) + R(0.0005));
    const R ultra_rapid_K_current_iur_gate__i_ur_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(7.5)) * (R(1) / R(10.0))), V, 0.14426950408889636, 1.0820212806667224)));
    const R ultra_rapid_K_current_iur_gate__tau_i_ur = (cmx_div(R(0.59), (R(1.0) + cmx_exp_v(((V + R(60.0)) * (R(1) / R(10.0))), V, 0.14426950408889636, 8.65617024533378))) + R(3.05));
    if (RL) {
        g_k[1] = cmx_div(((-R(1)) / sodium_current_m_gate__tau_m), R(1000.0));
        g_inf[1] = cmx_div(-((sodium_current_m_gate__m_infinity / sodium_current_m_gate__tau_m)), ((-R(1)) / sodium_current_m_gate__tau_m));
    } else {
        dy[1] = (cmx_div((sodium_current_m_gate__m_infinity - y[1]), sodium_current_m_gate__tau_m) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[2] = cmx_div(((-R(1)) / sodium_current_h1_gate__tau_h1), R(1000.0));
        g_inf[2] = cmx_div(-((sodium_current_h1_gate__h_infinity / sodium_current_h1_gate__tau_h1)), ((-R(1)) / sodium_current_h1_gate__tau_h1));
    } else {
        dy[2] = (cmx_div((sodium_current_h1_gate__h_infinity - y[2]), sodium_current_h1_gate__tau_h1) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[3] = cmx_div(((-R(1)) / sodium_current_h2_gate__tau_h2), R(1000.0));
        g_inf[3] = cmx_div(-((sodium_current_h1_gate__h_infinity / sodium_current_h2_gate__tau_h2)), ((-R(1)) / sodium_current_h2_gate__tau_h2));
    } else {
        dy[3] = (cmx_div((sodium_current_h1_gate__h_infinity - y[3]), sodium_current_h2_gate__tau_h2) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[4] = cmx_div(((-R(1)) / L_type_Ca_channel_d_L_gate__tau_d_L), R(1000.0));
        g_inf[4] = cmx_div(-((L_type_Ca_channel_d_L_gate__d_L_infinity / L_type_Ca_channel_d_L_gate__tau_d_L)), ((-R(1)) / L_type_Ca_channel_d_L_gate__tau_d_L));
    } else {
        dy[4] = (cmx_div((L_type_Ca_channel_d_L_gate__d_L_infinity - y[4]), L_type_Ca_channel_d_L_gate__tau_d_L) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[5] = cmx_div(((-R(1)) / L_type_Ca_channel_f_L1_gate__tau_f_L1), R(1000.0));
        g_inf[5] = cmx_div(-((L_type_Ca_channel_f_L1_gate__f_L_infinity / L_type_Ca_channel_f_L1_gate__tau_f_L1)), ((-R(1)) / L_type_Ca_channel_f_L1_gate__tau_f_L1));
    } else {
        dy[5] = (cmx_div((L_type_Ca_channel_f_L1_gate__f_L_infinity - y[5]), L_type_Ca_channel_f_L1_gate__tau_f_L1) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[6] = cmx_div(((-R(1)) / L_type_Ca_channel_f_L2_gate__tau_f_L2), R(1000.0));
        g_inf[6] = cmx_div(-((L_type_Ca_channel_f_L1_gate__f_L_infinity / L_type_Ca_channel_f_L2_gate__tau_f_L2)), ((-R(1)) / L_type_Ca_channel_f_L2_gate__tau_f_L2));
    } else {
        dy[6] = (cmx_div((L_type_Ca_channel_f_L1_gate__f_L_infinity - y[6]), L_type_Ca_channel_f_L2_gate__tau_f_L2) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[7] = cmx_div(((-R(1)) / Ca_independent_transient_outward_K_current_r_gate__tau_r), R(1000.0));
        g_inf[7] = cmx_div(-((Ca_independent_transient_outward_K_current_r_gate__r_infinity / Ca_independent_transient_outward_K_current_r_gate__tau_r)), ((-R(1)) / Ca_independent_transient_outward_K_current_r_gate__tau_r));
    } else {
        dy[7] = (cmx_div((Ca_independent_transient_outward_K_current_r_gate__r_infinity - y[7]), Ca_independent_transient_outward_K_current_r_gate__tau_r) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[8] = cmx_div(((-R(1)) / Ca_independent_transient_outward_K_current_s_gate__tau_s), R(1000.0));
        g_inf[8] = cmx_div(-((Ca_independent_transient_outward_K_current_s_gate__s_infinity / Ca_independent_transient_outward_K_current_s_gate__tau_s)), ((-R(1)) / Ca_independent_transient_outward_K_current_s_gate__tau_s));
    } else {
        dy[8] = (cmx_div((Ca_independent_transient_outward_K_current_s_gate__s_infinity - y[8]), Ca_independent_transient_outward_K_current_s_gate__tau_s) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[9] = cmx_div(((-R(1)) / ultra_rapid_K_current_aur_gate__tau_a_ur), R(1000.0));
        g_inf[9] = cmx_div(-((ultra_rapid_K_current_aur_gate__a_ur_infinity / ultra_rapid_K_current_aur_gate__tau_a_ur)), ((-R(1)) / ultra_rapid_K_current_aur_gate__tau_a_ur));
    } else {
        dy[9] = (cmx_div((ultra_rapid_K_current_aur_gate__a_ur_infinity - y[9]), ultra_rapid_K_current_aur_gate__tau_a_ur) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[10] = cmx_div(((-R(1)) / ultra_rapid_K_current_iur_gate__tau_i_ur), R(1000.0));
        g_inf[10] = cmx_div(-((ultra_rapid_K_current_iur_gate__i_ur_infinity / ultra_rapid_K_current_iur_gate__tau_i_ur)), ((-R(1)) / ultra_rapid_K_current_iur_gate__tau_i_ur));
    } else {
        dy[10] = (cmx_div((ultra_rapid_K_current_iur_gate__i_ur_infinity - y[10]), ultra_rapid_K_current_iur_gate__tau_i_ur) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[11] = cmx_div(((-R(1)) / delayed_rectifier_K_currents_n_gate__tau_n), R(1000.0));
        g_inf[11] = cmx_div(-((delayed_rectifier_K_currents_n_gate__n_infinity / delayed_rectifier_K_currents_n_gate__tau_n)), ((-R(1)) / delayed_rectifier_K_currents_n_gate__tau_n));
    } else {
        dy[11] = (cmx_div((delayed_rectifier_K_currents_n_gate__n_infinity - y[11]), delayed_rectifier_K_currents_n_gate__tau_n) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[12] = cmx_div(((-R(1)) / delayed_rectifier_K_currents_pa_gate__tau_pa), R(1000.0));
        g_inf[12] = cmx_div(-((delayed_rectifier_K_currents_pa_gate__p_a_infinity / delayed_rectifier_K_currents_pa_gate__tau_pa)), ((-R(1)) / delayed_rectifier_K_currents_pa_gate__tau_pa));
    } else {
        dy[12] = (cmx_div((delayed_rectifier_K_currents_pa_gate__p_a_infinity - y[12]), delayed_rectifier_K_currents_pa_gate__tau_pa) * (R(1) / R(1000.0)));
    }
    dy[13] = (((-(sodium_current__i_Na + background_currents__i_B_Na + (R(3.0) * Na_Ca_ion_exchanger_current__i_NaCa) + (R(3.0) * sodium_potassium_pump__i_NaK) + intracellular_ion_concentrations__phi_Na_en)) * (R(1) / (intracellular_ion_concentrations__Vol_i * membrane__F))) * (R(1) / R(1000.0)));
    dy[14] = ((((-((background_currents__i_B_Ca + sarcolemmal_calcium_pump_current__i_CaP + Ca_handling_by_the_SR__i_up) - (intracellular_ion_concentrations__i_di + Ca_handling_by_the_SR__i_rel + (R(2.0) * Na_Ca_ion_exchanger_current__i_NaCa)))) * (R(1) / (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F))) - (R(1.0) * intracellular_Ca_buffering__J_O)) * (R(1) / R(1000.0)));
    dy[15] = (((-(((Ca_independent_transient_outward_K_current__i_t + ultra_rapid_K_current__i_Kur + inward_rectifier__i_K1 + delayed_rectifier_K_currents__i_Ks + delayed_rectifier_K_currents__i_Kr + ACh_dependent_K_current__i_KACh) - (R(2.0) * sodium_potassium_pump__i_NaK)) + i_stim)) * (R(1) / (intracellular_ion_concentrations__Vol_i * membrane__F))) * (R(1) / R(1000.0)));
    dy[16] = (((-(L_type_Ca_channel__i_Ca_L + intracellular_ion_concentrations__i_di)) * (R(1) / (R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F))) * (R(1) / R(1000.0)));
    dy[17] = (intracellular_Ca_buffering__J_O_C * (R(1) / R(1000.0)));
    dy[18] = (intracellular_Ca_buffering__J_O_TC * (R(1) / R(1000.0)));
    dy[19] = (intracellular_Ca_buffering__J_O_TMgC * (R(1) / R(1000.0)));
    dy[20] = (intracellular_Ca_buffering__J_O_TMgMg * (R(1) / R(1000.0)));
    dy[21] = (intracellular_Ca_buffering__J_O * (R(1) / R(1000.0)));
    dy[22] = ((((cleft_space_ion_concentrations__Na_b - y[22]) * (R(1) / cleft_space_ion_concentrations__tau_Na)) + ((sodium_current__i_Na + background_currents__i_B_Na + (R(3.0) * Na_Ca_ion_exchanger_current__i_NaCa) + (R(3.0) * sodium_potassium_pump__i_NaK) + intracellular_ion_concentrations__phi_Na_en) * (R(1) / (cleft_space_ion_concentrations__Vol_c * membrane__F)))) * (R(1) / R(1000.0)));
    dy[23] = ((((cleft_space_ion_concentrations__Ca_b - y[23]) * (R(1) / cleft_space_ion_concentrations__tau_Ca)) + (((L_type_Ca_channel__i_Ca_L + background_currents__i_B_Ca + sarcolemmal_calcium_pump_current__i_CaP) - (R(2.0) * Na_Ca_ion_exchanger_current__i_NaCa)) * (R(1) / (R(2.0) * cleft_space_ion_concentrations__Vol_c * membrane__F)))) * (R(1) / R(1000.0)));
    dy[24] = ((((cleft_space_ion_concentrations__K_b - y[24]) * (R(1) / cleft_space_ion_concentrations__tau_K)) + (((Ca_independent_transient_outward_K_current__i_t + ultra_rapid_K_current__i_Kur + inward_rectifier__i_K1 + delayed_rectifier_K_currents__i_Ks + delayed_rectifier_K_currents__i_Kr) - (R(2.0) * sodium_potassium_pump__i_NaK)) * (R(1) / (cleft_space_ion_concentrations__Vol_c * membrane__F)))) * (R(1) / R(1000.0)));
    dy[25] = ((((Ca_handling_by_the_SR__i_tr - Ca_handling_by_the_SR__i_rel) * (R(1) / (R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F))) - (R(31.0) * Ca_handling_by_the_SR__J_O_Calse)) * (R(1) / R(1000.0)));
    dy[26] = (((Ca_handling_by_the_SR__i_up - Ca_handling_by_the_SR__i_tr) * (R(1) / (R(2.0) * Ca_handling_by_the_SR__Vol_up * membrane__F))) * (R(1) / R(1000.0)));
    dy[27] = (Ca_handling_by_the_SR__J_O_Calse * (R(1) / R(1000.0)));
    dy[28] = (((Ca_handling_by_the_SR__r_recov * ((R(1.0) - y[28]) - y[29])) - (Ca_handling_by_the_SR__r_act * y[28])) * (R(1) / R(1000.0)));
    dy[29] = (((Ca_handling_by_the_SR__r_act * y[28]) - (Ca_handling_by_the_SR__r_inact * y[29])) * (R(1) / R(1000.0)));
}

// dV/dt including the stimulus (single-cell integration).
template <class R>
CMX_HD inline R dvdt(const R* y, R V, R i_stim_area, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(306.15);
    const R membrane__F = R(96487.0);
    const R membrane__Cm = R(0.05);
    const R sodium_current__P_Na = R(0.0018);
    const R L_type_Ca_channel__g_Ca_L = R(6.75);
    const R L_type_Ca_channel__E_Ca_app = R(60.0);
    const R L_type_Ca_channel__k_Ca = R(0.025);
    const R Ca_independent_transient_outward_K_current__g_t = R(8.25);
    const R ultra_rapid_K_current__g_kur = R(2.25);
    const R inward_rectifier__g_K1 = R(3.1);
    const R inward_rectifier__ScaleFactorGK1 = R(1.0);
    const R delayed_rectifier_K_currents__g_Ks = R(1.0);
    const R delayed_rectifier_K_currents__g_Kr = R(0.5);
    const R delayed_rectifier_K_currents_pa_gate__AZD = R(0.0);
    const R background_currents__g_B_Na = R(0.060599);
    const R background_currents__g_B_Ca = R(0.078681);
    const R sodium_potassium_pump__K_NaK_K = R(1.0);
    const R sodium_potassium_pump__i_NaK_max = R(68.55);
    const R sodium_potassium_pump__ScaleFactorGNaK = R(1.0);
    const R sodium_potassium_pump__pow_K_NaK_Na_15 = R(36.4829);
    const R sarcolemmal_calcium_pump_current__i_CaP_max = R(4.0);
    const R sarcolemmal_calcium_pump_current__k_CaP = R(0.0002);
    const R Na_Ca_ion_exchanger_current__ScaleFactorGNaCa = R(1.0);
    const R Na_Ca_ion_exchanger_current__K_NaCa = R(0.0374842);
    const R Na_Ca_ion_exchanger_current__d_NaCa = R(0.0003);
    const R Na_Ca_ion_exchanger_current__gamma_Na = R(0.45);
    const R ACh_dependent_K_current__ACh = R(1e-24);
    const R i_stim = (i_stim_area / cm) * R(50.00000000000001);
    (void)i_stim;
    const R Ca_independent_transient_outward_K_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(y[24], y[15])));
    const R ACh_dependent_K_current__i_KACh = ((R(10000.0) * (R(1) / (R(1.0) + ((R(9.13652) * pow(R(1.0), R(0.477811))) * (R(1) / pow(ACh_dependent_K_current__ACh, R(0.477811))))))) * (R(0.0517) + cmx_div(R(0.4516), (R(1.0) + cmx_exp_v(((V + R(59.53)) * (R(1) / R(17.18))), V, 0.08397526431251243, 4.999047484523865)))) * (V - Ca_independent_transient_outward_K_current__E_K) * membrane__Cm);
    const R Ca_independent_transient_outward_K_current__i_t = (p[4] * Ca_independent_transient_outward_K_current__g_t * y[7] * y[8] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R L_type_Ca_channel__f_Ca = cmx_div(y[16], (y[16] + L_type_Ca_channel__k_Ca));
    const R L_type_Ca_channel__i_Ca_L = (p[1] * L_type_Ca_channel__g_Ca_L * y[4] * ((L_type_Ca_channel__f_Ca * y[5]) + ((R(1.0) - L_type_Ca_channel__f_Ca) * y[6])) * (V - L_type_Ca_channel__E_Ca_app));
    const R Na_Ca_ion_exchanger_current__i_NaCa = (Na_Ca_ion_exchanger_current__ScaleFactorGNaCa * cmx_div((Na_Ca_ion_exchanger_current__K_NaCa * ((y[13] * y[13] * y[13] * y[23] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (y[22] * y[22] * y[22] * y[14] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))), (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((y[22] * y[22] * y[22] * y[14]) + (y[13] * y[13] * y[13] * y[23]))))));
    const R background_currents__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(y[23], y[14])));
    const R background_currents__i_B_Ca = (background_currents__g_B_Ca * (V - background_currents__E_Ca));
    const R sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(y[22], y[13])));
    const R background_currents__i_B_Na = (background_currents__g_B_Na * (V - sodium_current__E_Na));
    const R delayed_rectifier_K_currents__factorAZD = ((R(1.0) - (R(1.0) * (R(1) / (R(1.0) + cmx_exp(((-(delayed_rectifier_K_currents_pa_gate__AZD - R(0.6))) * (R(1) / R(0.12)))))))) * (R(1) / R(0.9933)));
    const R delayed_rectifier_K_currents_pi_gate__pip = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(55.0)) * (R(1) / R(24.0))), V, 0.06011229337037347, 3.306176135370541)));
    const R delayed_rectifier_K_currents__i_Kr = (p[2] * delayed_rectifier_K_currents__g_Kr * delayed_rectifier_K_currents__factorAZD * y[12] * delayed_rectifier_K_currents_pi_gate__pip * (V - Ca_independent_transient_outward_K_current__E_K));
    const R delayed_rectifier_K_currents__i_Ks = (p[3] * delayed_rectifier_K_currents__g_Ks * y[11] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R inward_rectifier__i_K1 = cmx_div((inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((y[24] * (R(1) / R(1.0))), R(0.4457)) * (V - Ca_independent_transient_outward_K_current__E_K)), (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)));
    const R sarcolemmal_calcium_pump_current__i_CaP = cmx_div((sarcolemmal_calcium_pump_current__i_CaP_max * y[14]), (y[14] + sarcolemmal_calcium_pump_current__k_CaP));
    const R sodium_current__i_Na = cmx_remove_singularity(V, R(-0.0), R(2.6380041870925613e-06), [&](const R V) { return cmx_div((((p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * y[22] * V * membrane__F * membrane__F) * (R(1) / (membrane__R * membrane__T))) * cmx_expm1((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R sodium_potassium_pump__pow_Na_i_15 = pow(y[13], R(1.5));
    const R sodium_potassium_pump__i_NaK = cmx_div((cmx_div((cmx_div((sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * y[24]), (y[24] + sodium_potassium_pump__K_NaK_K)) * sodium_potassium_pump__pow_Na_i_15), (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15)) * (V + R(150.0))), (V + R(200.0)));
    const R ultra_rapid_K_current__i_Kur = (p[5] * ultra_rapid_K_current__g_kur * y[9] * y[10] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R membrane__I = (sodium_current__i_Na + L_type_Ca_channel__i_Ca_L + Ca_independent_transient_outward_K_current__i_t + ultra_rapid_K_current__i_Kur + inward_rectifier__i_K1 + delayed_rectifier_K_currents__i_Kr + delayed_rectifier_K_currents__i_Ks + background_currents__i_B_Na + background_currents__i_B_Ca + sodium_potassium_pump__i_NaK + sarcolemmal_calcium_pump_current__i_CaP + Na_Ca_ion_exchanger_current__i_NaCa + ACh_dependent_K_current__i_KACh + i_stim);
    return (((-membrane__I) * (R(1) / membrane__Cm)) * (R(1) / R(1000.0)));
}

// Total ionic current in uA/cm^2 (GetIIonic): minus dV/dt at zero stimulus,
// times the membrane capacitance that converts it to a current density.
template <class R>
CMX_HD inline R iionic(const R* y, R V, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(306.15);
    const R membrane__F = R(96487.0);
    const R membrane__Cm = R(0.05);
    const R sodium_current__P_Na = R(0.0018);
    const R L_type_Ca_channel__g_Ca_L = R(6.75);
    const R L_type_Ca_channel__E_Ca_app = R(60.0);
    const R L_type_Ca_channel__k_Ca = R(0.025);
    const R Ca_independent_transient_outward_K_current__g_t = R(8.25);
    const R ultra_rapid_K_current__g_kur = R(2.25);
    const R inward_rectifier__g_K1 = R(3.1);
    const R inward_rectifier__ScaleFactorGK1 = R(1.0);
    const R delayed_rectifier_K_currents__g_Ks = R(1.0);
    const R delayed_rectifier_K_currents__g_Kr = R(0.5);
    const R delayed_rectifier_K_currents_pa_gate__AZD = R(0.0);
    const R background_currents__g_B_Na = R(0.060599);
    const R background_currents__g_B_Ca = R(0.078681);
    const R sodium_potassium_pump__K_NaK_K = R(1.0);
    const R sodium_potassium_pump__i_NaK_max = R(68.55);
    const R sodium_potassium_pump__ScaleFactorGNaK = R(1.0);
    const R sodium_potassium_pump__pow_K_NaK_Na_15 = R(36.4829);
    const R sarcolemmal_calcium_pump_current__i_CaP_max = R(4.0);
    const R sarcolemmal_calcium_pump_current__k_CaP = R(0.0002);
    const R Na_Ca_ion_exchanger_current__ScaleFactorGNaCa = R(1.0);
    const R Na_Ca_ion_exchanger_current__K_NaCa = R(0.0374842);
    const R Na_Ca_ion_exchanger_current__d_NaCa = R(0.0003);
    const R Na_Ca_ion_exchanger_current__gamma_Na = R(0.45);
    const R ACh_dependent_K_current__ACh = R(1e-24);
    const R i_stim = R(0);
    (void)i_stim;
    const R Ca_independent_transient_outward_K_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(y[24], y[15])));
    const R ACh_dependent_K_current__i_KACh = ((R(10000.0) * (R(1) / (R(1.0) + ((R(9.13652) * pow(R(1.0), R(0.477811))) * (R(1) / pow(ACh_dependent_K_current__ACh, R(0.477811))))))) * (R(0.0517) + cmx_div(R(0.4516), (R(1.0) + cmx_exp_v(((V + R(59.53)) * (R(1) / R(17.18))), V, 0.08397526431251243, 4.999047484523865)))) * (V - Ca_independent_transient_outward_K_current__E_K) * membrane__Cm);
    const R Ca_independent_transient_outward_K_current__i_t = (p[4] * Ca_independent_transient_outward_K_current__g_t * y[7] * y[8] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R L_type_Ca_channel__f_Ca = cmx_div(y[16], (y[16] + L_type_Ca_channel__k_Ca));
    const R L_type_Ca_channel__i_Ca_L = (p[1] * L_type_Ca_channel__g_Ca_L * y[4] * ((L_type_Ca_channel__f_Ca * y[5]) + ((R(1.0) - L_type_Ca_channel__f_Ca) * y[6])) * (V - L_type_Ca_channel__E_Ca_app));
    const R Na_Ca_ion_exchanger_current__i_NaCa = (Na_Ca_ion_exchanger_current__ScaleFactorGNaCa * cmx_div((Na_Ca_ion_exchanger_current__K_NaCa * ((y[13] * y[13] * y[13] * y[23] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (y[22] * y[22] * y[22] * y[14] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))), (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((y[22] * y[22] * y[22] * y[14]) + (y[13] * y[13] * y[13] * y[23]))))));
    const R background_currents__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(y[23], y[14])));
    const R background_currents__i_B_Ca = (background_currents__g_B_Ca * (V - background_currents__E_Ca));
    const R sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(y[22], y[13])));
    const R background_currents__i_B_Na = (background_currents__g_B_Na * (V - sodium_current__E_Na));
    const R delayed_rectifier_K_currents__factorAZD = ((R(1.0) - (R(1.0) * (R(1) / (R(1.0) + cmx_exp(((-(delayed_rectifier_K_currents_pa_gate__AZD - R(0.6))) * (R(1) / R(0.12)))))))) * (R(1) / R(0.9933)));
    const R delayed_rectifier_K_currents_pi_gate__pip = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(55.0)) * (R(1) / R(24.0))), V, 0.06011229337037347, 3.306176135370541)));
    const R delayed_rectifier_K_currents__i_Kr = (p[2] * delayed_rectifier_K_currents__g_Kr * delayed_rectifier_K_currents__factorAZD * y[12] * delayed_rectifier_K_currents_pi_gate__pip * (V - Ca_independent_transient_outward_K_current__E_K));
    const R delayed_rectifier_K_currents__i_Ks = (p[3] * delayed_rectifier_K_currents__g_Ks * y[11] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R inward_rectifier__i_K1 = cmx_div((inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((y[24] * (R(1) / R(1.0))), R(0.4457)) * (V - Ca_independent_transient_outward_K_current__E_K)), (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)));
    const R sarcolemmal_calcium_pump_current__i_CaP = cmx_div((sarcolemmal_calcium_pump_current__i_CaP_max * y[14]), (y[14] + sarcolemmal_calcium_pump_current__k_CaP));
    const R sodium_current__i_Na = cmx_remove_singularity(V, R(-0.0), R(2.6380041870925613e-06), [&](const R V) { return cmx_div((((p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * y[22] * V * membrane__F * membrane__F) * (R(1) / (membrane__R * membrane__T))) * cmx_expm1((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R sodium_potassium_pump__pow_Na_i_15 = pow(y[13], R(1.5));
    const R sodium_potassium_pump__i_NaK = cmx_div((cmx_div((cmx_div((sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * y[24]), (y[24] + sodium_potassium_pump__K_NaK_K)) * sodium_potassium_pump__pow_Na_i_15), (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15)) * (V + R(150.0))), (V + R(200.0)));
    const R ultra_rapid_K_current__i_Kur = (p[5] * ultra_rapid_K_current__g_kur * y[9] * y[10] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R membrane__I = (sodium_current__i_Na + L_type_Ca_channel__i_Ca_L + Ca_independent_transient_outward_K_current__i_t + ultra_rapid_K_current__i_Kur + inward_rectifier__i_K1 + delayed_rectifier_K_currents__i_Kr + delayed_rectifier_K_currents__i_Ks + background_currents__i_B_Na + background_currents__i_B_Ca + sodium_potassium_pump__i_NaK + sarcolemmal_calcium_pump_current__i_CaP + Na_Ca_ion_exchanger_current__i_NaCa + ACh_dependent_K_current__i_KACh + i_stim);
    const R dv = (((-membrane__I) * (R(1) / membrane__Cm)) * (R(1) / R(1000.0)));
    return -dv * cm;
}

// Backward Euler: the states solved by Newton (all but the linear gates).
constexpr int kNumNonlinear = 17;
static constexpr int kNonlinear[17] = {13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29};

// Linear states: y <- (y + a dt) / (1 - b dt) for dy/dt = a + b y, a and b
// functions of V (ComputeOneStepExceptVoltage, closed-form part).
template <class R>
CMX_HD inline void be_linear(R* y, R V, const R* p, R cm, R dt)
{
    CMX_MATH_USING
    (void)p; (void)cm; (void)V;
    const R delayed_rectifier_K_currents_pa_gate__AZD = R(0.0);
    const R Ca_independent_transient_outward_K_current_r_gate__r_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(1.0)) * (R(1) / (-R(11.0)))), V, -0.1311540946262694, 0.1311540946262694)));
    const R Ca_independent_transient_outward_K_current_r_gate__tau_r = ((R(0.0035) * cmx_exp_c(((((-V) * V) * (R(1) / R(30.0))) * (R(1) / R(30.0))), (((-V) * V) * (R(1) / R(30.0))), 0.04808983469629878)) + R(0.0015));
    const R Ca_independent_transient_outward_K_current_s_gate__s_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(40.5)) * (R(1) / R(11.5))), V, 0.12545174268599682, 5.080795578782872)));
    const R Ca_independent_transient_outward_K_current_s_gate__s_factor = ((V + R(52.45)) * (R(1) / R(15.8827)));
    const R Ca_independent_transient_outward_K_current_s_gate__tau_s = ((R(0.025635) * cmx_exp(((-Ca_independent_transient_outward_K_current_s_gate__s_factor) * Ca_independent_transient_outward_K_current_s_gate__s_factor))) + R(0.01414));
    const R L_type_Ca_channel_d_L_gate__d_L_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / (-R(5.8)))), V, -0.2487405242912006, -2.238664718620805)));
    const R L_type_Ca_channel_d_L_gate__d_L_factor = ((V + R(35.0)) * (R(1) / R(30.0)));
    const R L_type_Ca_channel_d_L_gate__tau_d_L = ((R(0.0027) * cmx_exp(((-L_type_Ca_channel_d_L_gate__d_L_factor) * L_type_Ca_channel_d_L_gate__d_L_factor))) + R(0.002));
    const R L_type_Ca_channel_f_L1_gate__f_L_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(27.4)) * (R(1) / R(7.1))), V, 0.20319648463224838, 5.567583678923605)));
    const R L_type_Ca_channel_f_L1_gate__f_L_factor = (V + R(40.0));
    const R L_type_Ca_channel_f_L1_gate__tau_f_L1 = ((R(0.161) * cmx_exp_c(((((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.4))) * (R(1) / R(14.4))), (((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.4))), 0.10018715561728912)) + R(0.01));
    const R L_type_Ca_channel_f_L2_gate__tau_f_L2 = ((R(1.3323) * cmx_exp_c(((((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.2))) * (R(1) / R(14.2))), (((-L_type_Ca_channel_f_L1_gate__f_L_factor) * L_type_Ca_channel_f_L1_gate__f_L_factor) * (R(1) / R(14.2))), 0.10159824231612419)) + R(0.0626));
    const R delayed_rectifier_K_currents_n_gate__n_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(19.9)) * (R(1) / (-R(12.7)))), V, -0.1135980347156664, 2.2606008908417614)));
    const R delayed_rectifier_K_currents_n_gate__n_factor = ((V - R(20.0)) * (R(1) / R(20.0)));
    const R delayed_rectifier_K_currents_n_gate__tau_n = (R(0.7) + (R(0.4) * cmx_exp(((-delayed_rectifier_K_currents_n_gate__n_factor) * delayed_rectifier_K_currents_n_gate__n_factor))));
    const R delayed_rectifier_K_currents_pa_gate__p_a_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(15.0) + (R(10.0) * delayed_rectifier_K_currents_pa_gate__AZD)) * (R(1) / (-R(6.0)))), V, -0.2404491734814939, -3.6067376022224087)));
    const R delayed_rectifier_K_currents_pa_gate__pa_factor = ((V + R(20.1376)) * (R(1) / R(22.1996)));
    const R delayed_rectifier_K_currents_pa_gate__tau_pa = (R(0.03118) + (R(0.21718) * cmx_exp(((-delayed_rectifier_K_currents_pa_gate__pa_factor) * delayed_rectifier_K_currents_pa_gate__pa_factor))));
    const R sodium_current_h1_gate__h_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(63.6)) * (R(1) / R(5.3))), V, 0.27220661148848363, 17.31234049066756)));
    const R sodium_current_h1_gate__h_factor = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(35.1)) * (R(1) / R(3.2))), V, 0.4508422002778011, 15.824561229750817)));
    const R sodium_current_h1_gate__tau_h1 = ((R(0.03) * sodium_current_h1_gate__h_factor) + R(0.0003));
    const R sodium_current_h2_gate__tau_h2 = ((R(0.12) * sodium_current_h1_gate__h_factor) + R(0.003));
    const R sodium_current_m_gate__m_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(27.12)) * (R(1) / (-R(8.21)))), V, -0.1757241219109578, -4.765638186225175)));
    const R sodium_current_m_gate__m_factor = ((V + R(25.57)) * (R(1) / R(28.8)));
    const R sodium_current_m_gate__tau_m = ((R(4.2e-05) * cmx_exp(((-sodium_current_m_gate__m_factor) * sodium_current_m_gate__m_factor))) + R(2.4e-05));
    const R ultra_rapid_K_current_aur_gate__a_ur_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(6.0))) * (R(1) / R(8.6))), V, -0.16775523731267017, -1.006531423876021)));
    const R ultra_rapid_K_current_aur_gate__tau_a_ur = (cmx_div(R(0.009), (R(1.0) + cmx_exp_v(((V + R(5.0)) * (R(1) / R(12.0))), V, 0.12022458674074694, 0.6011229337037348))) + R(0.0005));
    const R ultra_rapid_K_current_iur_gate__i_ur_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(7.5)) * (R(1) / R(10.0))), V, 0.14426950408889636, 1.0820212806667224)));
    const R ultra_rapid_K_current_iur_gate__tau_i_ur = (cmx_div(R(0.59), (R(1.0) + cmx_exp_v(((V + R(60.0)) * (R(1) / R(10.0))), V, 0.14426950408889636, 8.65617024533378))) + R(3.05));
    y[1] = (y[1] + (((sodium_current_m_gate__m_infinity / sodium_current_m_gate__tau_m) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / sodium_current_m_gate__tau_m) / R(1000.0))) * dt);
    y[2] = (y[2] + (((sodium_current_h1_gate__h_infinity / sodium_current_h1_gate__tau_h1) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / sodium_current_h1_gate__tau_h1) / R(1000.0))) * dt);
    y[3] = (y[3] + (((sodium_current_h1_gate__h_infinity / sodium_current_h2_gate__tau_h2) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / sodium_current_h2_gate__tau_h2) / R(1000.0))) * dt);
    y[4] = (y[4] + (((L_type_Ca_channel_d_L_gate__d_L_infinity / L_type_Ca_channel_d_L_gate__tau_d_L) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / L_type_Ca_channel_d_L_gate__tau_d_L) / R(1000.0))) * dt);
    y[5] = (y[5] + (((L_type_Ca_channel_f_L1_gate__f_L_infinity / L_type_Ca_channel_f_L1_gate__tau_f_L1) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / L_type_Ca_channel_f_L1_gate__tau_f_L1) / R(1000.0))) * dt);
    y[6] = (y[6] + (((L_type_Ca_channel_f_L1_gate__f_L_infinity / L_type_Ca_channel_f_L2_gate__tau_f_L2) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / L_type_Ca_channel_f_L2_gate__tau_f_L2) / R(1000.0))) * dt);
    y[7] = (y[7] + (((Ca_independent_transient_outward_K_current_r_gate__r_infinity / Ca_independent_transient_outward_K_current_r_gate__tau_r) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / Ca_independent_transient_outward_K_current_r_gate__tau_r) / R(1000.0))) * dt);
    y[8] = (y[8] + (((Ca_independent_transient_outward_K_current_s_gate__s_infinity / Ca_independent_transient_outward_K_current_s_gate__tau_s) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / Ca_independent_transient_outward_K_current_s_gate__tau_s) / R(1000.0))) * dt);
    y[9] = (y[9] + (((ultra_rapid_K_current_aur_gate__a_ur_infinity / ultra_rapid_K_current_aur_gate__tau_a_ur) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / ultra_rapid_K_current_aur_gate__tau_a_ur) / R(1000.0))) * dt);
    y[10] = (y[10] + (((ultra_rapid_K_current_iur_gate__i_ur_infinity / ultra_rapid_K_current_iur_gate__tau_i_ur) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / ultra_rapid_K_current_iur_gate__tau_i_ur) / R(1000.0))) * dt);
    y[11] = (y[11] + (((delayed_rectifier_K_currents_n_gate__n_infinity / delayed_rectifier_K_currents_n_gate__tau_n) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / delayed_rectifier_K_currents_n_gate__tau_n) / R(1000.0))) * dt);
    y[12] = (y[12] + (((delayed_rectifier_K_currents_pa_gate__p_a_infinity / delayed_rectifier_K_currents_pa_gate__tau_pa) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / delayed_rectifier_K_currents_pa_gate__tau_pa) / R(1000.0))) * dt);
}

// Newton residual of the nonlinear states at guess g (y holds the other
// states, linear ones already advanced): g - y_old - dt f(g).
template <class R>
CMX_HD inline void be_residual(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* res)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(306.15);
    const R membrane__F = R(96487.0);
    const R membrane__Cm = R(0.05);
    const R sodium_current__P_Na = R(0.0018);
    const R L_type_Ca_channel__g_Ca_L = R(6.75);
    const R L_type_Ca_channel__E_Ca_app = R(60.0);
    const R L_type_Ca_channel__k_Ca = R(0.025);
    const R Ca_independent_transient_outward_K_current__g_t = R(8.25);
    const R ultra_rapid_K_current__g_kur = R(2.25);
    const R inward_rectifier__g_K1 = R(3.1);
    const R inward_rectifier__ScaleFactorGK1 = R(1.0);
    const R delayed_rectifier_K_currents__g_Ks = R(1.0);
    const R delayed_rectifier_K_currents__g_Kr = R(0.5);
    const R delayed_rectifier_K_currents_pa_gate__AZD = R(0.0);
    const R background_currents__g_B_Na = R(0.060599);
    const R background_currents__g_B_Ca = R(0.078681);
    const R sodium_potassium_pump__K_NaK_K = R(1.0);
    const R sodium_potassium_pump__i_NaK_max = R(68.55);
    const R sodium_potassium_pump__ScaleFactorGNaK = R(1.0);
    const R sodium_potassium_pump__pow_K_NaK_Na_15 = R(36.4829);
    const R sarcolemmal_calcium_pump_current__i_CaP_max = R(4.0);
    const R sarcolemmal_calcium_pump_current__k_CaP = R(0.0002);
    const R Na_Ca_ion_exchanger_current__ScaleFactorGNaCa = R(1.0);
    const R Na_Ca_ion_exchanger_current__K_NaCa = R(0.0374842);
    const R Na_Ca_ion_exchanger_current__d_NaCa = R(0.0003);
    const R Na_Ca_ion_exchanger_current__gamma_Na = R(0.45);
    const R ACh_dependent_K_current__ACh = R(1e-24);
    const R intracellular_ion_concentrations__phi_Na_en = R(0.0);
    const R intracellular_ion_concentrations__Vol_i = R(0.005884);
    const R intracellular_ion_concentrations__Vol_d = R(0.00011768);
    const R intracellular_ion_concentrations__tau_di = R(0.01);
    const R intracellular_Ca_buffering__Mg_i = R(2.5);
    const R cleft_space_ion_concentrations__Vol_c = R(0.000800224);
    const R cleft_space_ion_concentrations__tau_Na = R(14.3);
    const R cleft_space_ion_concentrations__tau_K = R(10.0);
    const R cleft_space_ion_concentrations__tau_Ca = R(24.7);
    const R cleft_space_ion_concentrations__Na_b = R(130.0);
    const R cleft_space_ion_concentrations__Ca_b = R(1.8);
    const R cleft_space_ion_concentrations__K_b = R(5.4);
    const R Ca_handling_by_the_SR__I_up_max = R(2800.0);
    const R Ca_handling_by_the_SR__k_cyca = R(0.0003);
    const R Ca_handling_by_the_SR__k_srca = R(0.5);
    const R Ca_handling_by_the_SR__k_xcs = R(0.4);
    const R Ca_handling_by_the_SR__alpha_rel = R(200000.0);
    const R Ca_handling_by_the_SR__Vol_up = R(0.0003969);
    const R Ca_handling_by_the_SR__Vol_rel = R(4.41e-05);
    const R Ca_handling_by_the_SR__r_recov = R(0.815);
    const R Ca_handling_by_the_SR__tau_tr = R(0.01);
    const R Ca_handling_by_the_SR__k_rel_i = R(0.0003);
    const R Ca_handling_by_the_SR__k_rel_d = R(0.003);
    const R i_stim = (i_stim_area / cm) * R(50.00000000000001);
    (void)i_stim;
    const R Ca_independent_transient_outward_K_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(g[11], g[2])));
    const R ACh_dependent_K_current__i_KACh = ((R(10000.0) * (R(1) / (R(1.0) + ((R(9.13652) * pow(R(1.0), R(0.477811))) * (R(1) / pow(ACh_dependent_K_current__ACh, R(0.477811))))))) * (R(0.0517) + cmx_div(R(0.4516), (R(1.0) + cmx_exp_v(((V + R(59.53)) * (R(1) / R(17.18))), V, 0.08397526431251243, 4.999047484523865)))) * (V - Ca_independent_transient_outward_K_current__E_K) * membrane__Cm);
    const R Ca_handling_by_the_SR__J_O_Calse = ((R(480.0) * g[12] * (R(1.0) - g[14])) - (R(400.0) * g[14]));
    const R Ca_handling_by_the_SR__i_rel_f2 = cmx_div(g[16], (g[16] + R(0.25)));
    const R Ca_handling_by_the_SR__i_rel_factor = (Ca_handling_by_the_SR__i_rel_f2 * Ca_handling_by_the_SR__i_rel_f2);
    const R Ca_handling_by_the_SR__i_rel = (Ca_handling_by_the_SR__alpha_rel * Ca_handling_by_the_SR__i_rel_factor * (g[12] - g[1]));
    const R Ca_handling_by_the_SR__i_tr = (((g[13] - g[12]) * R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F) * (R(1) / Ca_handling_by_the_SR__tau_tr));
    const R Ca_handling_by_the_SR__i_up = cmx_div((Ca_handling_by_the_SR__I_up_max * ((g[1] * (R(1) / Ca_handling_by_the_SR__k_cyca)) - ((Ca_handling_by_the_SR__k_xcs * Ca_handling_by_the_SR__k_xcs * g[13]) * (R(1) / Ca_handling_by_the_SR__k_srca)))), (((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca))));
    const R Ca_handling_by_the_SR__r_Ca_d_term = cmx_div(g[3], (g[3] + Ca_handling_by_the_SR__k_rel_d));
    const R Ca_handling_by_the_SR__r_Ca_d_factor = (Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term);
    const R Ca_handling_by_the_SR__r_Ca_i_term = cmx_div(g[1], (g[1] + Ca_handling_by_the_SR__k_rel_i));
    const R Ca_handling_by_the_SR__r_Ca_i_factor = (Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term);
    const R Ca_handling_by_the_SR__r_act = (R(203.8) * (Ca_handling_by_the_SR__r_Ca_i_factor + Ca_handling_by_the_SR__r_Ca_d_factor));
    const R Ca_handling_by_the_SR__r_inact = (R(33.96) + (R(339.6) * Ca_handling_by_the_SR__r_Ca_i_factor));
    const R Ca_independent_transient_outward_K_current__i_t = (p[4] * Ca_independent_transient_outward_K_current__g_t * y[7] * y[8] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R L_type_Ca_channel__f_Ca = cmx_div(g[3], (g[3] + L_type_Ca_channel__k_Ca));
    const R L_type_Ca_channel__i_Ca_L = (p[1] * L_type_Ca_channel__g_Ca_L * y[4] * ((L_type_Ca_channel__f_Ca * y[5]) + ((R(1.0) - L_type_Ca_channel__f_Ca) * y[6])) * (V - L_type_Ca_channel__E_Ca_app));
    const R Na_Ca_ion_exchanger_current__i_NaCa = (Na_Ca_ion_exchanger_current__ScaleFactorGNaCa * cmx_div((Na_Ca_ion_exchanger_current__K_NaCa * ((g[0] * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (g[9] * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))), (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10]))))));
    const R background_currents__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(g[10], g[1])));
    const R background_currents__i_B_Ca = (background_currents__g_B_Ca * (V - background_currents__E_Ca));
    const R sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(g[9], g[0])));
    const R background_currents__i_B_Na = (background_currents__g_B_Na * (V - sodium_current__E_Na));
    const R delayed_rectifier_K_currents__factorAZD = ((R(1.0) - (R(1.0) * (R(1) / (R(1.0) + cmx_exp(((-(delayed_rectifier_K_currents_pa_gate__AZD - R(0.6))) * (R(1) / R(0.12)))))))) * (R(1) / R(0.9933)));
    const R delayed_rectifier_K_currents_pi_gate__pip = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(55.0)) * (R(1) / R(24.0))), V, 0.06011229337037347, 3.306176135370541)));
    const R delayed_rectifier_K_currents__i_Kr = (p[2] * delayed_rectifier_K_currents__g_Kr * delayed_rectifier_K_currents__factorAZD * y[12] * delayed_rectifier_K_currents_pi_gate__pip * (V - Ca_independent_transient_outward_K_current__E_K));
    const R delayed_rectifier_K_currents__i_Ks = (p[3] * delayed_rectifier_K_currents__g_Ks * y[11] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R intracellular_Ca_buffering__J_O_C = ((R(200000.0) * g[1] * (R(1.0) - g[4])) - (R(476.0) * g[4]));
    const R intracellular_Ca_buffering__J_O_TC = ((R(78400.0) * g[1] * (R(1.0) - g[5])) - (R(392.0) * g[5]));
    const R intracellular_Ca_buffering__J_O_TMgC = ((R(200000.0) * g[1] * ((R(1.0) - g[6]) - g[7])) - (R(6.6) * g[6]));
    const R intracellular_Ca_buffering__J_O = ((R(0.08) * intracellular_Ca_buffering__J_O_TC) + (R(0.16) * intracellular_Ca_buffering__J_O_TMgC) + (R(0.045) * intracellular_Ca_buffering__J_O_C));
    const R intracellular_Ca_buffering__J_O_TMgMg = ((R(2000.0) * intracellular_Ca_buffering__Mg_i * ((R(1.0) - g[6]) - g[7])) - (R(666.0) * g[7]));
    const R intracellular_ion_concentrations__i_di = (((g[3] - g[1]) * R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F) * (R(1) / intracellular_ion_concentrations__tau_di));
    const R inward_rectifier__i_K1 = cmx_div((inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((g[11] * (R(1) / R(1.0))), R(0.4457)) * (V - Ca_independent_transient_outward_K_current__E_K)), (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)));
    const R sarcolemmal_calcium_pump_current__i_CaP = cmx_div((sarcolemmal_calcium_pump_current__i_CaP_max * g[1]), (g[1] + sarcolemmal_calcium_pump_current__k_CaP));
    const R sodium_current__i_Na = cmx_remove_singularity(V, R(-0.0), R(2.6380041870925613e-06), [&](const R V) { return cmx_div((((p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * g[9] * V * membrane__F * membrane__F) * (R(1) / (membrane__R * membrane__T))) * cmx_expm1((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R sodium_potassium_pump__pow_Na_i_15 = pow(g[0], R(1.5));
    const R sodium_potassium_pump__i_NaK = cmx_div((cmx_div((cmx_div((sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * g[11]), (g[11] + sodium_potassium_pump__K_NaK_K)) * sodium_potassium_pump__pow_Na_i_15), (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15)) * (V + R(150.0))), (V + R(200.0)));
    const R ultra_rapid_K_current__i_Kur = (p[5] * ultra_rapid_K_current__g_kur * y[9] * y[10] * (V - Ca_independent_transient_outward_K_current__E_K));
    res[0] = g[0] - y[13] - dt * (((-(sodium_current__i_Na + background_currents__i_B_Na + (R(3.0) * Na_Ca_ion_exchanger_current__i_NaCa) + (R(3.0) * sodium_potassium_pump__i_NaK) + intracellular_ion_concentrations__phi_Na_en)) * (R(1) / (intracellular_ion_concentrations__Vol_i * membrane__F))) * (R(1) / R(1000.0)));
    res[1] = g[1] - y[14] - dt * ((((-((background_currents__i_B_Ca + sarcolemmal_calcium_pump_current__i_CaP + Ca_handling_by_the_SR__i_up) - (intracellular_ion_concentrations__i_di + Ca_handling_by_the_SR__i_rel + (R(2.0) * Na_Ca_ion_exchanger_current__i_NaCa)))) * (R(1) / (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F))) - (R(1.0) * intracellular_Ca_buffering__J_O)) * (R(1) / R(1000.0)));
    res[2] = g[2] - y[15] - dt * (((-(((Ca_independent_transient_outward_K_current__i_t + ultra_rapid_K_current__i_Kur + inward_rectifier__i_K1 + delayed_rectifier_K_currents__i_Ks + delayed_rectifier_K_currents__i_Kr + ACh_dependent_K_current__i_KACh) - (R(2.0) * sodium_potassium_pump__i_NaK)) + i_stim)) * (R(1) / (intracellular_ion_concentrations__Vol_i * membrane__F))) * (R(1) / R(1000.0)));
    res[3] = g[3] - y[16] - dt * (((-(L_type_Ca_channel__i_Ca_L + intracellular_ion_concentrations__i_di)) * (R(1) / (R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F))) * (R(1) / R(1000.0)));
    res[4] = g[4] - y[17] - dt * (intracellular_Ca_buffering__J_O_C * (R(1) / R(1000.0)));
    res[5] = g[5] - y[18] - dt * (intracellular_Ca_buffering__J_O_TC * (R(1) / R(1000.0)));
    res[6] = g[6] - y[19] - dt * (intracellular_Ca_buffering__J_O_TMgC * (R(1) / R(1000.0)));
    res[7] = g[7] - y[20] - dt * (intracellular_Ca_buffering__J_O_TMgMg * (R(1) / R(1000.0)));
    res[8] = g[8] - y[21] - dt * (intracellular_Ca_buffering__J_O * (R(1) / R(1000.0)));
    res[9] = g[9] - y[22] - dt * ((((cleft_space_ion_concentrations__Na_b - g[9]) * (R(1) / cleft_space_ion_concentrations__tau_Na)) + ((sodium_current__i_Na + background_currents__i_B_Na + (R(3.0) * Na_Ca_ion_exchanger_current__i_NaCa) + (R(3.0) * sodium_potassium_pump__i_NaK) + intracellular_ion_concentrations__phi_Na_en) * (R(1) / (cleft_space_ion_concentrations__Vol_c * membrane__F)))) * (R(1) / R(1000.0)));
    res[10] = g[10] - y[23] - dt * ((((cleft_space_ion_concentrations__Ca_b - g[10]) * (R(1) / cleft_space_ion_concentrations__tau_Ca)) + (((L_type_Ca_channel__i_Ca_L + background_currents__i_B_Ca + sarcolemmal_calcium_pump_current__i_CaP) - (R(2.0) * Na_Ca_ion_exchanger_current__i_NaCa)) * (R(1) / (R(2.0) * cleft_space_ion_concentrations__Vol_c * membrane__F)))) * (R(1) / R(1000.0)));
    res[11] = g[11] - y[24] - dt * ((((cleft_space_ion_concentrations__K_b - g[11]) * (R(1) / cleft_space_ion_concentrations__tau_K)) + (((Ca_independent_transient_outward_K_current__i_t + ultra_rapid_K_current__i_Kur + inward_rectifier__i_K1 + delayed_rectifier_K_currents__i_Ks + delayed_rectifier_K_currents__i_Kr) - (R(2.0) * sodium_potassium_pump__i_NaK)) * (R(1) / (cleft_space_ion_concentrations__Vol_c * membrane__F)))) * (R(1) / R(1000.0)));
    res[12] = g[12] - y[25] - dt * ((((Ca_handling_by_the_SR__i_tr - Ca_handling_by_the_SR__i_rel) * (R(1) / (R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F))) - (R(31.0) * Ca_handling_by_the_SR__J_O_Calse)) * (R(1) / R(1000.0)));
    res[13] = g[13] - y[26] - dt * (((Ca_handling_by_the_SR__i_up - Ca_handling_by_the_SR__i_tr) * (R(1) / (R(2.0) * Ca_handling_by_the_SR__Vol_up * membrane__F))) * (R(1) / R(1000.0)));
    res[14] = g[14] - y[27] - dt * (Ca_handling_by_the_SR__J_O_Calse * (R(1) / R(1000.0)));
    res[15] = g[15] - y[28] - dt * (((Ca_handling_by_the_SR__r_recov * ((R(1.0) - g[15]) - g[16])) - (Ca_handling_by_the_SR__r_act * g[15])) * (R(1) / R(1000.0)));
    res[16] = g[16] - y[29] - dt * (((Ca_handling_by_the_SR__r_act * g[15]) - (Ca_handling_by_the_SR__r_inact * g[16])) * (R(1) / R(1000.0)));
}

// Jacobian of be_residual in g: J[i*N + j] = delta_ij - dt d f_i / d g_j.
template <class R>
CMX_HD inline void be_jacobian(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* J)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(306.15);
    const R membrane__F = R(96487.0);
    const R membrane__Cm = R(0.05);
    const R sodium_current__P_Na = R(0.0018);
    const R L_type_Ca_channel__g_Ca_L = R(6.75);
    const R L_type_Ca_channel__E_Ca_app = R(60.0);
    const R L_type_Ca_channel__k_Ca = R(0.025);
    const R Ca_independent_transient_outward_K_current__g_t = R(8.25);
    const R ultra_rapid_K_current__g_kur = R(2.25);
    const R inward_rectifier__g_K1 = R(3.1);
    const R inward_rectifier__ScaleFactorGK1 = R(1.0);
    const R delayed_rectifier_K_currents__g_Ks = R(1.0);
    const R delayed_rectifier_K_currents__g_Kr = R(0.5);
    const R delayed_rectifier_K_currents_pa_gate__AZD = R(0.0);
    const R background_currents__g_B_Na = R(0.060599);
    const R background_currents__g_B_Ca = R(0.078681);
    const R sodium_potassium_pump__K_NaK_K = R(1.0);
    const R sodium_potassium_pump__i_NaK_max = R(68.55);
    const R sodium_potassium_pump__ScaleFactorGNaK = R(1.0);
    const R sodium_potassium_pump__pow_K_NaK_Na_15 = R(36.4829);
    const R sarcolemmal_calcium_pump_current__i_CaP_max = R(4.0);
    const R sarcolemmal_calcium_pump_current__k_CaP = R(0.0002);
    const R Na_Ca_ion_exchanger_current__ScaleFactorGNaCa = R(1.0);
    const R Na_Ca_ion_exchanger_current__K_NaCa = R(0.0374842);
    const R Na_Ca_ion_exchanger_current__d_NaCa = R(0.0003);
    const R Na_Ca_ion_exchanger_current__gamma_Na = R(0.45);
    const R ACh_dependent_K_current__ACh = R(1e-24);
    const R intracellular_ion_concentrations__phi_Na_en = R(0.0);
    const R intracellular_ion_concentrations__Vol_i = R(0.005884);
    const R intracellular_ion_concentrations__Vol_d = R(0.00011768);
    const R intracellular_ion_concentrations__tau_di = R(0.01);
    const R intracellular_Ca_buffering__Mg_i = R(2.5);
    const R cleft_space_ion_concentrations__Vol_c = R(0.000800224);
    const R cleft_space_ion_concentrations__tau_Na = R(14.3);
    const R cleft_space_ion_concentrations__tau_K = R(10.0);
    const R cleft_space_ion_concentrations__tau_Ca = R(24.7);
    const R cleft_space_ion_concentrations__Na_b = R(130.0);
    const R cleft_space_ion_concentrations__Ca_b = R(1.8);
    const R cleft_space_ion_concentrations__K_b = R(5.4);
    const R Ca_handling_by_the_SR__I_up_max = R(2800.0);
    const R Ca_handling_by_the_SR__k_cyca = R(0.0003);
    const R Ca_handling_by_the_SR__k_srca = R(0.5);
    const R Ca_handling_by_the_SR__k_xcs = R(0.4);
    const R Ca_handling_by_the_SR__alpha_rel = R(200000.0);
    const R Ca_handling_by_the_SR__Vol_up = R(0.0003969);
    const R Ca_handling_by_the_SR__Vol_rel = R(4.41e-05);
    const R Ca_handling_by_the_SR__r_recov = R(0.815);
    const R Ca_handling_by_the_SR__tau_tr = R(0.01);
    const R Ca_handling_by_the_SR__k_rel_i = R(0.0003);
    const R Ca_handling_by_the_SR__k_rel_d = R(0.003);
    const R i_stim = (i_stim_area / cm) * R(50.00000000000001);
    (void)i_stim;
    const R Ca_independent_transient_outward_K_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(g[11], g[2])));
    const R ACh_dependent_K_current__i_KACh = ((R(10000.0) * (R(1) / (R(1.0) + ((R(9.13652) * pow(R(1.0), R(0.477811))) * (R(1) / pow(ACh_dependent_K_current__ACh, R(0.477811))))))) * (R(0.0517) + cmx_div(R(0.4516), (R(1.0) + cmx_exp_v(((V + R(59.53)) * (R(1) / R(17.18))), V, 0.08397526431251243, 4.999047484523865)))) * (V - Ca_independent_transient_outward_K_current__E_K) * membrane__Cm);
    const R Ca_handling_by_the_SR__J_O_Calse = ((R(480.0) * g[12] * (R(1.0) - g[14])) - (R(400.0) * g[14]));
    const R Ca_handling_by_the_SR__i_rel_f2 = cmx_div(g[16], (g[16] + R(0.25)));
    const R Ca_handling_by_the_SR__i_rel_factor = (Ca_handling_by_the_SR__i_rel_f2 * Ca_handling_by_the_SR__i_rel_f2);
    const R Ca_handling_by_the_SR__i_rel = (Ca_handling_by_the_SR__alpha_rel * Ca_handling_by_the_SR__i_rel_factor * (g[12] - g[1]));
    const R Ca_handling_by_the_SR__i_tr = (((g[13] - g[12]) * R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F) * (R(1) / Ca_handling_by_the_SR__tau_tr));
    const R Ca_handling_by_the_SR__i_up = cmx_div((Ca_handling_by_the_SR__I_up_max * ((g[1] * (R(1) / Ca_handling_by_the_SR__k_cyca)) - ((Ca_handling_by_the_SR__k_xcs * Ca_handling_by_the_SR__k_xcs * g[13]) * (R(1) / Ca_handling_by_the_SR__k_srca)))), (((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca))));
    const R Ca_handling_by_the_SR__r_Ca_d_term = cmx_div(g[3], (g[3] + Ca_handling_by_the_SR__k_rel_d));
    const R Ca_handling_by_the_SR__r_Ca_d_factor = (Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term);
    const R Ca_handling_by_the_SR__r_Ca_i_term = cmx_div(g[1], (g[1] + Ca_handling_by_the_SR__k_rel_i));
    const R Ca_handling_by_the_SR__r_Ca_i_factor = (Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term);
    const R Ca_handling_by_the_SR__r_act = (R(203.8) * (Ca_handling_by_the_SR__r_Ca_i_factor + Ca_handling_by_the_SR__r_Ca_d_factor));
    const R Ca_handling_by_the_SR__r_inact = (R(33.96) + (R(339.6) * Ca_handling_by_the_SR__r_Ca_i_factor));
    const R Ca_independent_transient_outward_K_current__i_t = (p[4] * Ca_independent_transient_outward_K_current__g_t * y[7] * y[8] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R L_type_Ca_channel__f_Ca = cmx_div(g[3], (g[3] + L_type_Ca_channel__k_Ca));
    const R L_type_Ca_channel__i_Ca_L = (p[1] * L_type_Ca_channel__g_Ca_L * y[4] * ((L_type_Ca_channel__f_Ca * y[5]) + ((R(1.0) - L_type_Ca_channel__f_Ca) * y[6])) * (V - L_type_Ca_channel__E_Ca_app));
    const R Na_Ca_ion_exchanger_current__i_NaCa = (Na_Ca_ion_exchanger_current__ScaleFactorGNaCa * cmx_div((Na_Ca_ion_exchanger_current__K_NaCa * ((g[0] * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (g[9] * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))), (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10]))))));
    const R background_currents__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(g[10], g[1])));
    const R background_currents__i_B_Ca = (background_currents__g_B_Ca * (V - background_currents__E_Ca));
    const R sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(g[9], g[0])));
    const R background_currents__i_B_Na = (background_currents__g_B_Na * (V - sodium_current__E_Na));
    const R delayed_rectifier_K_currents__factorAZD = ((R(1.0) - (R(1.0) * (R(1) / (R(1.0) + cmx_exp(((-(delayed_rectifier_K_currents_pa_gate__AZD - R(0.6))) * (R(1) / R(0.12)))))))) * (R(1) / R(0.9933)));
    const R delayed_rectifier_K_currents_pi_gate__pip = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(55.0)) * (R(1) / R(24.0))), V, 0.06011229337037347, 3.306176135370541)));
    const R delayed_rectifier_K_currents__i_Kr = (p[2] * delayed_rectifier_K_currents__g_Kr * delayed_rectifier_K_currents__factorAZD * y[12] * delayed_rectifier_K_currents_pi_gate__pip * (V - Ca_independent_transient_outward_K_current__E_K));
    const R delayed_rectifier_K_currents__i_Ks = (p[3] * delayed_rectifier_K_currents__g_Ks * y[11] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R intracellular_Ca_buffering__J_O_C = ((R(200000.0) * g[1] * (R(1.0) - g[4])) - (R(476.0) * g[4]));
    const R intracellular_Ca_buffering__J_O_TC = ((R(78400.0) * g[1] * (R(1.0) - g[5])) - (R(392.0) * g[5]));
    const R intracellular_Ca_buffering__J_O_TMgC = ((R(200000.0) * g[1] * ((R(1.0) - g[6]) - g[7])) - (R(6.6) * g[6]));
    const R intracellular_Ca_buffering__J_O = ((R(0.08) * intracellular_Ca_buffering__J_O_TC) + (R(0.16) * intracellular_Ca_buffering__J_O_TMgC) + (R(0.045) * intracellular_Ca_buffering__J_O_C));
    const R intracellular_Ca_buffering__J_O_TMgMg = ((R(2000.0) * intracellular_Ca_buffering__Mg_i * ((R(1.0) - g[6]) - g[7])) - (R(666.0) * g[7]));
    const R intracellular_ion_concentrations__i_di = (((g[3] - g[1]) * R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F) * (R(1) / intracellular_ion_concentrations__tau_di));
    const R inward_rectifier__i_K1 = cmx_div((inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((g[11] * (R(1) / R(1.0))), R(0.4457)) * (V - Ca_independent_transient_outward_K_current__E_K)), (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)));
    const R sarcolemmal_calcium_pump_current__i_CaP = cmx_div((sarcolemmal_calcium_pump_current__i_CaP_max * g[1]), (g[1] + sarcolemmal_calcium_pump_current__k_CaP));
    const R sodium_current__i_Na = cmx_remove_singularity(V, R(-0.0), R(2.6380041870925613e-06), [&](const R V) { return cmx_div((((p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * g[9] * V * membrane__F * membrane__F) * (R(1) / (membrane__R * membrane__T))) * cmx_expm1((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R sodium_potassium_pump__pow_Na_i_15 = pow(g[0], R(1.5));
    const R sodium_potassium_pump__i_NaK = cmx_div((cmx_div((cmx_div((sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * g[11]), (g[11] + sodium_potassium_pump__K_NaK_K)) * sodium_potassium_pump__pow_Na_i_15), (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15)) * (V + R(150.0))), (V + R(200.0)));
    const R ultra_rapid_K_current__i_Kur = (p[5] * ultra_rapid_K_current__g_kur * y[9] * y[10] * (V - Ca_independent_transient_outward_K_current__E_K));
    const R d_Ca_independent_transient_outward_K_current__E_K__d2 = ((cmx_div(cmx_div((-(g[11] * R(1))), (g[2] * g[2])), cmx_div(g[11], g[2])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_Ca_independent_transient_outward_K_current__E_K__d11 = ((cmx_div(cmx_div(R(1), g[2]), cmx_div(g[11], g[2])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_ACh_dependent_K_current__i_KACh__d2 = (((-d_Ca_independent_transient_outward_K_current__E_K__d2) * (R(10000.0) * (R(1) / (R(1.0) + ((R(9.13652) * pow(R(1.0), R(0.477811))) * (R(1) / pow(ACh_dependent_K_current__ACh, R(0.477811))))))) * (R(0.0517) + cmx_div(R(0.4516), (R(1.0) + cmx_exp_v(((V + R(59.53)) * (R(1) / R(17.18))), V, 0.08397526431251243, 4.999047484523865)))) * membrane__Cm));
    const R d_ACh_dependent_K_current__i_KACh__d11 = (((-d_Ca_independent_transient_outward_K_current__E_K__d11) * (R(10000.0) * (R(1) / (R(1.0) + ((R(9.13652) * pow(R(1.0), R(0.477811))) * (R(1) / pow(ACh_dependent_K_current__ACh, R(0.477811))))))) * (R(0.0517) + cmx_div(R(0.4516), (R(1.0) + cmx_exp_v(((V + R(59.53)) * (R(1) / R(17.18))), V, 0.08397526431251243, 4.999047484523865)))) * membrane__Cm));
    const R d_Ca_handling_by_the_SR__J_O_Calse__d12 = ((R(1) * R(480.0) * (R(1.0) - g[14])));
    const R d_Ca_handling_by_the_SR__J_O_Calse__d14 = ((((-R(1)) * R(480.0) * g[12])) - ((R(1) * R(400.0))));
    const R d_Ca_handling_by_the_SR__i_rel_f2__d16 = cmx_div((R(1) * (g[16] + R(0.25)) - g[16] * (R(1))), ((g[16] + R(0.25)) * (g[16] + R(0.25))));
    const R d_Ca_handling_by_the_SR__i_rel_factor__d16 = ((d_Ca_handling_by_the_SR__i_rel_f2__d16 * Ca_handling_by_the_SR__i_rel_f2) + (d_Ca_handling_by_the_SR__i_rel_f2__d16 * Ca_handling_by_the_SR__i_rel_f2));
    const R d_Ca_handling_by_the_SR__i_rel__d1 = (((-R(1)) * Ca_handling_by_the_SR__alpha_rel * Ca_handling_by_the_SR__i_rel_factor));
    const R d_Ca_handling_by_the_SR__i_rel__d12 = ((R(1) * Ca_handling_by_the_SR__alpha_rel * Ca_handling_by_the_SR__i_rel_factor));
    const R d_Ca_handling_by_the_SR__i_rel__d16 = ((d_Ca_handling_by_the_SR__i_rel_factor__d16 * Ca_handling_by_the_SR__alpha_rel * (g[12] - g[1])));
    const R d_Ca_handling_by_the_SR__i_tr__d12 = cmx_div((((-R(1)) * R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F)), Ca_handling_by_the_SR__tau_tr);
    const R d_Ca_handling_by_the_SR__i_tr__d13 = cmx_div(((R(1) * R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F)), Ca_handling_by_the_SR__tau_tr);
    const R d_Ca_handling_by_the_SR__i_up__d1 = cmx_div((((cmx_div(R(1), Ca_handling_by_the_SR__k_cyca) * Ca_handling_by_the_SR__I_up_max)) * (((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca))) - (Ca_handling_by_the_SR__I_up_max * ((g[1] * (R(1) / Ca_handling_by_the_SR__k_cyca)) - ((Ca_handling_by_the_SR__k_xcs * Ca_handling_by_the_SR__k_xcs * g[13]) * (R(1) / Ca_handling_by_the_SR__k_srca)))) * (cmx_div((R(1)), Ca_handling_by_the_SR__k_cyca))), ((((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca))) * (((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca)))));
    const R d_Ca_handling_by_the_SR__i_up__d13 = cmx_div(((((-cmx_div(((R(1) * Ca_handling_by_the_SR__k_xcs * Ca_handling_by_the_SR__k_xcs)), Ca_handling_by_the_SR__k_srca)) * Ca_handling_by_the_SR__I_up_max)) * (((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca))) - (Ca_handling_by_the_SR__I_up_max * ((g[1] * (R(1) / Ca_handling_by_the_SR__k_cyca)) - ((Ca_handling_by_the_SR__k_xcs * Ca_handling_by_the_SR__k_xcs * g[13]) * (R(1) / Ca_handling_by_the_SR__k_srca)))) * (cmx_div((((R(1)) * Ca_handling_by_the_SR__k_xcs)), Ca_handling_by_the_SR__k_srca))), ((((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca))) * (((g[1] + Ca_handling_by_the_SR__k_cyca) * (R(1) / Ca_handling_by_the_SR__k_cyca)) + ((Ca_handling_by_the_SR__k_xcs * (g[13] + Ca_handling_by_the_SR__k_srca)) * (R(1) / Ca_handling_by_the_SR__k_srca)))));
    const R d_Ca_handling_by_the_SR__r_Ca_d_term__d3 = cmx_div((R(1) * (g[3] + Ca_handling_by_the_SR__k_rel_d) - g[3] * (R(1))), ((g[3] + Ca_handling_by_the_SR__k_rel_d) * (g[3] + Ca_handling_by_the_SR__k_rel_d)));
    const R d_Ca_handling_by_the_SR__r_Ca_d_factor__d3 = ((d_Ca_handling_by_the_SR__r_Ca_d_term__d3 * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term) + (d_Ca_handling_by_the_SR__r_Ca_d_term__d3 * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term) + (d_Ca_handling_by_the_SR__r_Ca_d_term__d3 * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term) + (d_Ca_handling_by_the_SR__r_Ca_d_term__d3 * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term * Ca_handling_by_the_SR__r_Ca_d_term));
    const R d_Ca_handling_by_the_SR__r_Ca_i_term__d1 = cmx_div((R(1) * (g[1] + Ca_handling_by_the_SR__k_rel_i) - g[1] * (R(1))), ((g[1] + Ca_handling_by_the_SR__k_rel_i) * (g[1] + Ca_handling_by_the_SR__k_rel_i)));
    const R d_Ca_handling_by_the_SR__r_Ca_i_factor__d1 = ((d_Ca_handling_by_the_SR__r_Ca_i_term__d1 * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term) + (d_Ca_handling_by_the_SR__r_Ca_i_term__d1 * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term) + (d_Ca_handling_by_the_SR__r_Ca_i_term__d1 * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term) + (d_Ca_handling_by_the_SR__r_Ca_i_term__d1 * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term * Ca_handling_by_the_SR__r_Ca_i_term));
    const R d_Ca_handling_by_the_SR__r_act__d1 = (((d_Ca_handling_by_the_SR__r_Ca_i_factor__d1) * R(203.8)));
    const R d_Ca_handling_by_the_SR__r_act__d3 = (((d_Ca_handling_by_the_SR__r_Ca_d_factor__d3) * R(203.8)));
    const R d_Ca_handling_by_the_SR__r_inact__d1 = (((d_Ca_handling_by_the_SR__r_Ca_i_factor__d1 * R(339.6))));
    const R d_Ca_independent_transient_outward_K_current__i_t__d2 = (((-d_Ca_independent_transient_outward_K_current__E_K__d2) * p[4] * Ca_independent_transient_outward_K_current__g_t * y[7] * y[8]));
    const R d_Ca_independent_transient_outward_K_current__i_t__d11 = (((-d_Ca_independent_transient_outward_K_current__E_K__d11) * p[4] * Ca_independent_transient_outward_K_current__g_t * y[7] * y[8]));
    const R d_L_type_Ca_channel__f_Ca__d3 = cmx_div((R(1) * (g[3] + L_type_Ca_channel__k_Ca) - g[3] * (R(1))), ((g[3] + L_type_Ca_channel__k_Ca) * (g[3] + L_type_Ca_channel__k_Ca)));
    const R d_L_type_Ca_channel__i_Ca_L__d3 = (((((d_L_type_Ca_channel__f_Ca__d3 * y[5])) + (((-d_L_type_Ca_channel__f_Ca__d3) * y[6]))) * p[1] * L_type_Ca_channel__g_Ca_L * y[4] * (V - L_type_Ca_channel__E_Ca_app)));
    const R d_Na_Ca_ion_exchanger_current__i_NaCa__d0 = ((cmx_div((((((R(1) * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) + (R(1) * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) + (R(1) * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0))) * Na_Ca_ion_exchanger_current__K_NaCa)) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) - (Na_Ca_ion_exchanger_current__K_NaCa * ((g[0] * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (g[9] * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))) * ((((((R(1) * g[0] * g[0] * g[10]) + (R(1) * g[0] * g[0] * g[10]) + (R(1) * g[0] * g[0] * g[10]))) * Na_Ca_ion_exchanger_current__d_NaCa)))), ((R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))))) * Na_Ca_ion_exchanger_current__ScaleFactorGNaCa));
    const R d_Na_Ca_ion_exchanger_current__i_NaCa__d1 = ((cmx_div(((((-((R(1) * g[9] * g[9] * g[9] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))) * Na_Ca_ion_exchanger_current__K_NaCa)) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) - (Na_Ca_ion_exchanger_current__K_NaCa * ((g[0] * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (g[9] * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))) * ((((((R(1) * g[9] * g[9] * g[9]))) * Na_Ca_ion_exchanger_current__d_NaCa)))), ((R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))))) * Na_Ca_ion_exchanger_current__ScaleFactorGNaCa));
    const R d_Na_Ca_ion_exchanger_current__i_NaCa__d9 = ((cmx_div(((((-((R(1) * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)) + (R(1) * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)) + (R(1) * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))) * Na_Ca_ion_exchanger_current__K_NaCa)) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) - (Na_Ca_ion_exchanger_current__K_NaCa * ((g[0] * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (g[9] * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))) * ((((((R(1) * g[9] * g[9] * g[1]) + (R(1) * g[9] * g[9] * g[1]) + (R(1) * g[9] * g[9] * g[1]))) * Na_Ca_ion_exchanger_current__d_NaCa)))), ((R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))))) * Na_Ca_ion_exchanger_current__ScaleFactorGNaCa));
    const R d_Na_Ca_ion_exchanger_current__i_NaCa__d10 = ((cmx_div((((((R(1) * g[0] * g[0] * g[0] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0))) * Na_Ca_ion_exchanger_current__K_NaCa)) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) - (Na_Ca_ion_exchanger_current__K_NaCa * ((g[0] * g[0] * g[0] * g[10] * cmx_exp_v(((membrane__F * V * Na_Ca_ion_exchanger_current__gamma_Na) * (R(1) / (membrane__R * membrane__T))), V, 0.024609997648091457, 0.0)) - (g[9] * g[9] * g[9] * g[1] * cmx_exp_v((((Na_Ca_ion_exchanger_current__gamma_Na - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.030078886014334005, -0.0)))) * ((((((R(1) * g[0] * g[0] * g[0]))) * Na_Ca_ion_exchanger_current__d_NaCa)))), ((R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))) * (R(1.0) + (Na_Ca_ion_exchanger_current__d_NaCa * ((g[9] * g[9] * g[9] * g[1]) + (g[0] * g[0] * g[0] * g[10])))))) * Na_Ca_ion_exchanger_current__ScaleFactorGNaCa));
    const R d_background_currents__E_Ca__d1 = ((cmx_div(cmx_div((-(g[10] * R(1))), (g[1] * g[1])), cmx_div(g[10], g[1])) * ((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F)))));
    const R d_background_currents__E_Ca__d10 = ((cmx_div(cmx_div(R(1), g[1]), cmx_div(g[10], g[1])) * ((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F)))));
    const R d_background_currents__i_B_Ca__d1 = (((-d_background_currents__E_Ca__d1) * background_currents__g_B_Ca));
    const R d_background_currents__i_B_Ca__d10 = (((-d_background_currents__E_Ca__d10) * background_currents__g_B_Ca));
    const R d_sodium_current__E_Na__d0 = ((cmx_div(cmx_div((-(g[9] * R(1))), (g[0] * g[0])), cmx_div(g[9], g[0])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_sodium_current__E_Na__d9 = ((cmx_div(cmx_div(R(1), g[0]), cmx_div(g[9], g[0])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_background_currents__i_B_Na__d0 = (((-d_sodium_current__E_Na__d0) * background_currents__g_B_Na));
    const R d_background_currents__i_B_Na__d9 = (((-d_sodium_current__E_Na__d9) * background_currents__g_B_Na));
    const R d_delayed_rectifier_K_currents__i_Kr__d2 = (((-d_Ca_independent_transient_outward_K_current__E_K__d2) * p[2] * delayed_rectifier_K_currents__g_Kr * delayed_rectifier_K_currents__factorAZD * y[12] * delayed_rectifier_K_currents_pi_gate__pip));
    const R d_delayed_rectifier_K_currents__i_Kr__d11 = (((-d_Ca_independent_transient_outward_K_current__E_K__d11) * p[2] * delayed_rectifier_K_currents__g_Kr * delayed_rectifier_K_currents__factorAZD * y[12] * delayed_rectifier_K_currents_pi_gate__pip));
    const R d_delayed_rectifier_K_currents__i_Ks__d2 = (((-d_Ca_independent_transient_outward_K_current__E_K__d2) * p[3] * delayed_rectifier_K_currents__g_Ks * y[11]));
    const R d_delayed_rectifier_K_currents__i_Ks__d11 = (((-d_Ca_independent_transient_outward_K_current__E_K__d11) * p[3] * delayed_rectifier_K_currents__g_Ks * y[11]));
    const R d_intracellular_Ca_buffering__J_O_C__d1 = ((R(1) * R(200000.0) * (R(1.0) - g[4])));
    const R d_intracellular_Ca_buffering__J_O_C__d4 = ((((-R(1)) * R(200000.0) * g[1])) - ((R(1) * R(476.0))));
    const R d_intracellular_Ca_buffering__J_O_TC__d1 = ((R(1) * R(78400.0) * (R(1.0) - g[5])));
    const R d_intracellular_Ca_buffering__J_O_TC__d5 = ((((-R(1)) * R(78400.0) * g[1])) - ((R(1) * R(392.0))));
    const R d_intracellular_Ca_buffering__J_O_TMgC__d1 = ((R(1) * R(200000.0) * ((R(1.0) - g[6]) - g[7])));
    const R d_intracellular_Ca_buffering__J_O_TMgC__d6 = ((((-R(1)) * R(200000.0) * g[1])) - ((R(1) * R(6.6))));
    const R d_intracellular_Ca_buffering__J_O_TMgC__d7 = (((-R(1)) * R(200000.0) * g[1]));
    const R d_intracellular_Ca_buffering__J_O__d1 = (((d_intracellular_Ca_buffering__J_O_TC__d1 * R(0.08))) + ((d_intracellular_Ca_buffering__J_O_TMgC__d1 * R(0.16))) + ((d_intracellular_Ca_buffering__J_O_C__d1 * R(0.045))));
    const R d_intracellular_Ca_buffering__J_O__d4 = (((d_intracellular_Ca_buffering__J_O_C__d4 * R(0.045))));
    const R d_intracellular_Ca_buffering__J_O__d5 = (((d_intracellular_Ca_buffering__J_O_TC__d5 * R(0.08))));
    const R d_intracellular_Ca_buffering__J_O__d6 = (((d_intracellular_Ca_buffering__J_O_TMgC__d6 * R(0.16))));
    const R d_intracellular_Ca_buffering__J_O__d7 = (((d_intracellular_Ca_buffering__J_O_TMgC__d7 * R(0.16))));
    const R d_intracellular_Ca_buffering__J_O_TMgMg__d6 = (((-R(1)) * R(2000.0) * intracellular_Ca_buffering__Mg_i));
    const R d_intracellular_Ca_buffering__J_O_TMgMg__d7 = ((((-R(1)) * R(2000.0) * intracellular_Ca_buffering__Mg_i)) - ((R(1) * R(666.0))));
    const R d_intracellular_ion_concentrations__i_di__d1 = cmx_div((((-R(1)) * R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F)), intracellular_ion_concentrations__tau_di);
    const R d_intracellular_ion_concentrations__i_di__d3 = cmx_div(((R(1) * R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F)), intracellular_ion_concentrations__tau_di);
    const R d_inward_rectifier__i_K1__d2 = cmx_div(((((-d_Ca_independent_transient_outward_K_current__E_K__d2) * inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((g[11] * (R(1) / R(1.0))), R(0.4457)))) * (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)) - (inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((g[11] * (R(1) / R(1.0))), R(0.4457)) * (V - Ca_independent_transient_outward_K_current__E_K)) * ((cmx_exp(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T)))) * cmx_div(((((-d_Ca_independent_transient_outward_K_current__E_K__d2)) * R(1.5) * membrane__F)), (membrane__R * membrane__T))))), ((R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)) * (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07))));
    const R d_inward_rectifier__i_K1__d11 = cmx_div(((((R(0.4457) * pow((g[11] * (R(1) / R(1.0))), R(0.4457) - R(1)) * cmx_div(R(1), R(1.0))) * inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * (V - Ca_independent_transient_outward_K_current__E_K)) + ((-d_Ca_independent_transient_outward_K_current__E_K__d11) * inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((g[11] * (R(1) / R(1.0))), R(0.4457)))) * (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)) - (inward_rectifier__ScaleFactorGK1 * inward_rectifier__g_K1 * pow((g[11] * (R(1) / R(1.0))), R(0.4457)) * (V - Ca_independent_transient_outward_K_current__E_K)) * ((cmx_exp(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T)))) * cmx_div(((((-d_Ca_independent_transient_outward_K_current__E_K__d11)) * R(1.5) * membrane__F)), (membrane__R * membrane__T))))), ((R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07)) * (R(1.0) + cmx_exp_c(((R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), (R(1.5) * ((V - Ca_independent_transient_outward_K_current__E_K) + R(3.6)) * membrane__F), 5.668005395796891e-07))));
    const R d_sarcolemmal_calcium_pump_current__i_CaP__d1 = cmx_div((((R(1) * sarcolemmal_calcium_pump_current__i_CaP_max)) * (g[1] + sarcolemmal_calcium_pump_current__k_CaP) - (sarcolemmal_calcium_pump_current__i_CaP_max * g[1]) * (R(1))), ((g[1] + sarcolemmal_calcium_pump_current__k_CaP) * (g[1] + sarcolemmal_calcium_pump_current__k_CaP)));
    const R d_sodium_current__i_Na__d0 = cmx_remove_singularity(V, R(-0.0), R(2.6380041870925613e-06), [&](const R V) { return cmx_div((((cmx_exp((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T)))) * cmx_div((((-d_sodium_current__E_Na__d0) * membrane__F)), (membrane__R * membrane__T))) * ((p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * g[9] * V * membrane__F * membrane__F) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_sodium_current__i_Na__d9 = cmx_remove_singularity(V, R(-0.0), R(2.6380041870925613e-06), [&](const R V) { return cmx_div(((cmx_div(((R(1) * p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * V * membrane__F * membrane__F)), (membrane__R * membrane__T)) * cmx_expm1((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T))))) + ((cmx_exp((((V - sodium_current__E_Na) * membrane__F) * (R(1) / (membrane__R * membrane__T)))) * cmx_div((((-d_sodium_current__E_Na__d9) * membrane__F)), (membrane__R * membrane__T))) * ((p[0] * sodium_current__P_Na * y[1] * y[1] * y[1] * ((R(0.9) * y[2]) + (R(0.1) * y[3])) * g[9] * V * membrane__F * membrane__F) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_sodium_potassium_pump__pow_Na_i_15__d0 = (R(1.5) * pow(g[0], R(1.5) - R(1)) * R(1));
    const R d_sodium_potassium_pump__i_NaK__d0 = cmx_div(((cmx_div((((d_sodium_potassium_pump__pow_Na_i_15__d0 * cmx_div((sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * g[11]), (g[11] + sodium_potassium_pump__K_NaK_K)))) * (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15) - (cmx_div((sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * g[11]), (g[11] + sodium_potassium_pump__K_NaK_K)) * sodium_potassium_pump__pow_Na_i_15) * (d_sodium_potassium_pump__pow_Na_i_15__d0)), ((sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15) * (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15))) * (V + R(150.0)))), (V + R(200.0)));
    const R d_sodium_potassium_pump__i_NaK__d11 = cmx_div(((cmx_div(((cmx_div((((R(1) * sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max)) * (g[11] + sodium_potassium_pump__K_NaK_K) - (sodium_potassium_pump__ScaleFactorGNaK * sodium_potassium_pump__i_NaK_max * g[11]) * (R(1))), ((g[11] + sodium_potassium_pump__K_NaK_K) * (g[11] + sodium_potassium_pump__K_NaK_K))) * sodium_potassium_pump__pow_Na_i_15)), (sodium_potassium_pump__pow_Na_i_15 + sodium_potassium_pump__pow_K_NaK_Na_15)) * (V + R(150.0)))), (V + R(200.0)));
    const R d_ultra_rapid_K_current__i_Kur__d2 = (((-d_Ca_independent_transient_outward_K_current__E_K__d2) * p[5] * ultra_rapid_K_current__g_kur * y[9] * y[10]));
    const R d_ultra_rapid_K_current__i_Kur__d11 = (((-d_Ca_independent_transient_outward_K_current__E_K__d11) * p[5] * ultra_rapid_K_current__g_kur * y[9] * y[10]));
    J[0] = R(1) - dt * cmx_div(cmx_div((-(d_sodium_current__i_Na__d0 + d_background_currents__i_B_Na__d0 + ((d_Na_Ca_ion_exchanger_current__i_NaCa__d0 * R(3.0))) + ((d_sodium_potassium_pump__i_NaK__d0 * R(3.0))))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[1] = R(0) - dt * cmx_div(cmx_div((-(((d_Na_Ca_ion_exchanger_current__i_NaCa__d1 * R(3.0))))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[2] = R(0);
    J[3] = R(0);
    J[4] = R(0);
    J[5] = R(0);
    J[6] = R(0);
    J[7] = R(0);
    J[8] = R(0);
    J[9] = R(0) - dt * cmx_div(cmx_div((-(d_sodium_current__i_Na__d9 + d_background_currents__i_B_Na__d9 + ((d_Na_Ca_ion_exchanger_current__i_NaCa__d9 * R(3.0))))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[10] = R(0) - dt * cmx_div(cmx_div((-(((d_Na_Ca_ion_exchanger_current__i_NaCa__d10 * R(3.0))))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[11] = R(0) - dt * cmx_div(cmx_div((-(((d_sodium_potassium_pump__i_NaK__d11 * R(3.0))))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[12] = R(0);
    J[13] = R(0);
    J[14] = R(0);
    J[15] = R(0);
    J[16] = R(0);
    J[17] = R(0) - dt * cmx_div(cmx_div((-(-(((d_Na_Ca_ion_exchanger_current__i_NaCa__d0 * R(2.0)))))), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[18] = R(1) - dt * cmx_div((cmx_div((-((d_background_currents__i_B_Ca__d1 + d_sarcolemmal_calcium_pump_current__i_CaP__d1 + d_Ca_handling_by_the_SR__i_up__d1) - (d_intracellular_ion_concentrations__i_di__d1 + d_Ca_handling_by_the_SR__i_rel__d1 + ((d_Na_Ca_ion_exchanger_current__i_NaCa__d1 * R(2.0)))))), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)) - ((d_intracellular_Ca_buffering__J_O__d1 * R(1.0)))), R(1000.0));
    J[19] = R(0);
    J[20] = R(0) - dt * cmx_div(cmx_div((-(-(d_intracellular_ion_concentrations__i_di__d3))), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[21] = R(0) - dt * cmx_div((-((d_intracellular_Ca_buffering__J_O__d4 * R(1.0)))), R(1000.0));
    J[22] = R(0) - dt * cmx_div((-((d_intracellular_Ca_buffering__J_O__d5 * R(1.0)))), R(1000.0));
    J[23] = R(0) - dt * cmx_div((-((d_intracellular_Ca_buffering__J_O__d6 * R(1.0)))), R(1000.0));
    J[24] = R(0) - dt * cmx_div((-((d_intracellular_Ca_buffering__J_O__d7 * R(1.0)))), R(1000.0));
    J[25] = R(0);
    J[26] = R(0) - dt * cmx_div(cmx_div((-(-(((d_Na_Ca_ion_exchanger_current__i_NaCa__d9 * R(2.0)))))), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[27] = R(0) - dt * cmx_div(cmx_div((-((d_background_currents__i_B_Ca__d10) - (((d_Na_Ca_ion_exchanger_current__i_NaCa__d10 * R(2.0)))))), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[28] = R(0);
    J[29] = R(0) - dt * cmx_div(cmx_div((-(-(d_Ca_handling_by_the_SR__i_rel__d12))), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[30] = R(0) - dt * cmx_div(cmx_div((-(d_Ca_handling_by_the_SR__i_up__d13)), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[31] = R(0);
    J[32] = R(0);
    J[33] = R(0) - dt * cmx_div(cmx_div((-(-(d_Ca_handling_by_the_SR__i_rel__d16))), (R(2.0) * intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[34] = R(0) - dt * cmx_div(cmx_div((-((-((d_sodium_potassium_pump__i_NaK__d0 * R(2.0)))))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[35] = R(0);
    J[36] = R(1) - dt * cmx_div(cmx_div((-((d_Ca_independent_transient_outward_K_current__i_t__d2 + d_ultra_rapid_K_current__i_Kur__d2 + d_inward_rectifier__i_K1__d2 + d_delayed_rectifier_K_currents__i_Ks__d2 + d_delayed_rectifier_K_currents__i_Kr__d2 + d_ACh_dependent_K_current__i_KACh__d2))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[37] = R(0);
    J[38] = R(0);
    J[39] = R(0);
    J[40] = R(0);
    J[41] = R(0);
    J[42] = R(0);
    J[43] = R(0);
    J[44] = R(0);
    J[45] = R(0) - dt * cmx_div(cmx_div((-(((d_Ca_independent_transient_outward_K_current__i_t__d11 + d_ultra_rapid_K_current__i_Kur__d11 + d_inward_rectifier__i_K1__d11 + d_delayed_rectifier_K_currents__i_Ks__d11 + d_delayed_rectifier_K_currents__i_Kr__d11 + d_ACh_dependent_K_current__i_KACh__d11) - ((d_sodium_potassium_pump__i_NaK__d11 * R(2.0)))))), (intracellular_ion_concentrations__Vol_i * membrane__F)), R(1000.0));
    J[46] = R(0);
    J[47] = R(0);
    J[48] = R(0);
    J[49] = R(0);
    J[50] = R(0);
    J[51] = R(0);
    J[52] = R(0) - dt * cmx_div(cmx_div((-(d_intracellular_ion_concentrations__i_di__d1)), (R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F)), R(1000.0));
    J[53] = R(0);
    J[54] = R(1) - dt * cmx_div(cmx_div((-(d_L_type_Ca_channel__i_Ca_L__d3 + d_intracellular_ion_concentrations__i_di__d3)), (R(2.0) * intracellular_ion_concentrations__Vol_d * membrane__F)), R(1000.0));
    J[55] = R(0);
    J[56] = R(0);
    J[57] = R(0);
    J[58] = R(0);
    J[59] = R(0);
    J[60] = R(0);
    J[61] = R(0);
    J[62] = R(0);
    J[63] = R(0);
    J[64] = R(0);
    J[65] = R(0);
    J[66] = R(0);
    J[67] = R(0);
    J[68] = R(0);
    J[69] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_C__d1, R(1000.0));
    J[70] = R(0);
    J[71] = R(0);
    J[72] = R(1) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_C__d4, R(1000.0));
    J[73] = R(0);
    J[74] = R(0);
    J[75] = R(0);
    J[76] = R(0);
    J[77] = R(0);
    J[78] = R(0);
    J[79] = R(0);
    J[80] = R(0);
    J[81] = R(0);
    J[82] = R(0);
    J[83] = R(0);
    J[84] = R(0);
    J[85] = R(0);
    J[86] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_TC__d1, R(1000.0));
    J[87] = R(0);
    J[88] = R(0);
    J[89] = R(0);
    J[90] = R(1) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_TC__d5, R(1000.0));
    J[91] = R(0);
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
    J[102] = R(0);
    J[103] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_TMgC__d1, R(1000.0));
    J[104] = R(0);
    J[105] = R(0);
    J[106] = R(0);
    J[107] = R(0);
    J[108] = R(1) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_TMgC__d6, R(1000.0));
    J[109] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_TMgC__d7, R(1000.0));
    J[110] = R(0);
    J[111] = R(0);
    J[112] = R(0);
    J[113] = R(0);
    J[114] = R(0);
    J[115] = R(0);
    J[116] = R(0);
    J[117] = R(0);
    J[118] = R(0);
    J[119] = R(0);
    J[120] = R(0);
    J[121] = R(0);
    J[122] = R(0);
    J[123] = R(0);
    J[124] = R(0);
    J[125] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_TMgMg__d6, R(1000.0));
    J[126] = R(1) - dt * cmx_div(d_intracellular_Ca_buffering__J_O_TMgMg__d7, R(1000.0));
    J[127] = R(0);
    J[128] = R(0);
    J[129] = R(0);
    J[130] = R(0);
    J[131] = R(0);
    J[132] = R(0);
    J[133] = R(0);
    J[134] = R(0);
    J[135] = R(0);
    J[136] = R(0);
    J[137] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O__d1, R(1000.0));
    J[138] = R(0);
    J[139] = R(0);
    J[140] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O__d4, R(1000.0));
    J[141] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O__d5, R(1000.0));
    J[142] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O__d6, R(1000.0));
    J[143] = R(0) - dt * cmx_div(d_intracellular_Ca_buffering__J_O__d7, R(1000.0));
    J[144] = R(1);
    J[145] = R(0);
    J[146] = R(0);
    J[147] = R(0);
    J[148] = R(0);
    J[149] = R(0);
    J[150] = R(0);
    J[151] = R(0);
    J[152] = R(0);
    J[153] = R(0) - dt * cmx_div((cmx_div((d_sodium_current__i_Na__d0 + d_background_currents__i_B_Na__d0 + ((d_Na_Ca_ion_exchanger_current__i_NaCa__d0 * R(3.0))) + ((d_sodium_potassium_pump__i_NaK__d0 * R(3.0)))), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[154] = R(0) - dt * cmx_div((cmx_div((((d_Na_Ca_ion_exchanger_current__i_NaCa__d1 * R(3.0)))), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[155] = R(0);
    J[156] = R(0);
    J[157] = R(0);
    J[158] = R(0);
    J[159] = R(0);
    J[160] = R(0);
    J[161] = R(0);
    J[162] = R(1) - dt * cmx_div((cmx_div((-R(1)), cleft_space_ion_concentrations__tau_Na) + cmx_div((d_sodium_current__i_Na__d9 + d_background_currents__i_B_Na__d9 + ((d_Na_Ca_ion_exchanger_current__i_NaCa__d9 * R(3.0)))), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[163] = R(0) - dt * cmx_div((cmx_div((((d_Na_Ca_ion_exchanger_current__i_NaCa__d10 * R(3.0)))), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[164] = R(0) - dt * cmx_div((cmx_div((((d_sodium_potassium_pump__i_NaK__d11 * R(3.0)))), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[165] = R(0);
    J[166] = R(0);
    J[167] = R(0);
    J[168] = R(0);
    J[169] = R(0);
    J[170] = R(0) - dt * cmx_div((cmx_div((-((d_Na_Ca_ion_exchanger_current__i_NaCa__d0 * R(2.0)))), (R(2.0) * cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[171] = R(0) - dt * cmx_div((cmx_div(((d_background_currents__i_B_Ca__d1 + d_sarcolemmal_calcium_pump_current__i_CaP__d1) - ((d_Na_Ca_ion_exchanger_current__i_NaCa__d1 * R(2.0)))), (R(2.0) * cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[172] = R(0);
    J[173] = R(0) - dt * cmx_div((cmx_div((d_L_type_Ca_channel__i_Ca_L__d3), (R(2.0) * cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[174] = R(0);
    J[175] = R(0);
    J[176] = R(0);
    J[177] = R(0);
    J[178] = R(0);
    J[179] = R(0) - dt * cmx_div((cmx_div((-((d_Na_Ca_ion_exchanger_current__i_NaCa__d9 * R(2.0)))), (R(2.0) * cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[180] = R(1) - dt * cmx_div((cmx_div((-R(1)), cleft_space_ion_concentrations__tau_Ca) + cmx_div(((d_background_currents__i_B_Ca__d10) - ((d_Na_Ca_ion_exchanger_current__i_NaCa__d10 * R(2.0)))), (R(2.0) * cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[181] = R(0);
    J[182] = R(0);
    J[183] = R(0);
    J[184] = R(0);
    J[185] = R(0);
    J[186] = R(0);
    J[187] = R(0) - dt * cmx_div((cmx_div((-((d_sodium_potassium_pump__i_NaK__d0 * R(2.0)))), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[188] = R(0);
    J[189] = R(0) - dt * cmx_div((cmx_div((d_Ca_independent_transient_outward_K_current__i_t__d2 + d_ultra_rapid_K_current__i_Kur__d2 + d_inward_rectifier__i_K1__d2 + d_delayed_rectifier_K_currents__i_Ks__d2 + d_delayed_rectifier_K_currents__i_Kr__d2), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[190] = R(0);
    J[191] = R(0);
    J[192] = R(0);
    J[193] = R(0);
    J[194] = R(0);
    J[195] = R(0);
    J[196] = R(0);
    J[197] = R(0);
    J[198] = R(1) - dt * cmx_div((cmx_div((-R(1)), cleft_space_ion_concentrations__tau_K) + cmx_div(((d_Ca_independent_transient_outward_K_current__i_t__d11 + d_ultra_rapid_K_current__i_Kur__d11 + d_inward_rectifier__i_K1__d11 + d_delayed_rectifier_K_currents__i_Ks__d11 + d_delayed_rectifier_K_currents__i_Kr__d11) - ((d_sodium_potassium_pump__i_NaK__d11 * R(2.0)))), (cleft_space_ion_concentrations__Vol_c * membrane__F))), R(1000.0));
    J[199] = R(0);
    J[200] = R(0);
    J[201] = R(0);
    J[202] = R(0);
    J[203] = R(0);
    J[204] = R(0);
    J[205] = R(0) - dt * cmx_div(cmx_div((-d_Ca_handling_by_the_SR__i_rel__d1), (R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F)), R(1000.0));
    J[206] = R(0);
    J[207] = R(0);
    J[208] = R(0);
    J[209] = R(0);
    J[210] = R(0);
    J[211] = R(0);
    J[212] = R(0);
    J[213] = R(0);
    J[214] = R(0);
    J[215] = R(0);
    J[216] = R(1) - dt * cmx_div((cmx_div((d_Ca_handling_by_the_SR__i_tr__d12 - d_Ca_handling_by_the_SR__i_rel__d12), (R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F)) - ((d_Ca_handling_by_the_SR__J_O_Calse__d12 * R(31.0)))), R(1000.0));
    J[217] = R(0) - dt * cmx_div(cmx_div(d_Ca_handling_by_the_SR__i_tr__d13, (R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F)), R(1000.0));
    J[218] = R(0) - dt * cmx_div((-((d_Ca_handling_by_the_SR__J_O_Calse__d14 * R(31.0)))), R(1000.0));
    J[219] = R(0);
    J[220] = R(0) - dt * cmx_div(cmx_div((-d_Ca_handling_by_the_SR__i_rel__d16), (R(2.0) * Ca_handling_by_the_SR__Vol_rel * membrane__F)), R(1000.0));
    J[221] = R(0);
    J[222] = R(0) - dt * cmx_div(cmx_div(d_Ca_handling_by_the_SR__i_up__d1, (R(2.0) * Ca_handling_by_the_SR__Vol_up * membrane__F)), R(1000.0));
    J[223] = R(0);
    J[224] = R(0);
    J[225] = R(0);
    J[226] = R(0);
    J[227] = R(0);
    J[228] = R(0);
    J[229] = R(0);
    J[230] = R(0);
    J[231] = R(0);
    J[232] = R(0);
    J[233] = R(0) - dt * cmx_div(cmx_div((-d_Ca_handling_by_the_SR__i_tr__d12), (R(2.0) * Ca_handling_by_the_SR__Vol_up * membrane__F)), R(1000.0));
    J[234] = R(1) - dt * cmx_div(cmx_div((d_Ca_handling_by_the_SR__i_up__d13 - d_Ca_handling_by_the_SR__i_tr__d13), (R(2.0) * Ca_handling_by_the_SR__Vol_up * membrane__F)), R(1000.0));
    J[235] = R(0);
    J[236] = R(0);
    J[237] = R(0);
    J[238] = R(0);
    J[239] = R(0);
    J[240] = R(0);
    J[241] = R(0);
    J[242] = R(0);
    J[243] = R(0);
    J[244] = R(0);
    J[245] = R(0);
    J[246] = R(0);
    J[247] = R(0);
    J[248] = R(0);
    J[249] = R(0);
    J[250] = R(0) - dt * cmx_div(d_Ca_handling_by_the_SR__J_O_Calse__d12, R(1000.0));
    J[251] = R(0);
    J[252] = R(1) - dt * cmx_div(d_Ca_handling_by_the_SR__J_O_Calse__d14, R(1000.0));
    J[253] = R(0);
    J[254] = R(0);
    J[255] = R(0);
    J[256] = R(0) - dt * cmx_div((-((d_Ca_handling_by_the_SR__r_act__d1 * g[15]))), R(1000.0));
    J[257] = R(0);
    J[258] = R(0) - dt * cmx_div((-((d_Ca_handling_by_the_SR__r_act__d3 * g[15]))), R(1000.0));
    J[259] = R(0);
    J[260] = R(0);
    J[261] = R(0);
    J[262] = R(0);
    J[263] = R(0);
    J[264] = R(0);
    J[265] = R(0);
    J[266] = R(0);
    J[267] = R(0);
    J[268] = R(0);
    J[269] = R(0);
    J[270] = R(1) - dt * cmx_div(((((-R(1)) * Ca_handling_by_the_SR__r_recov)) - ((R(1) * Ca_handling_by_the_SR__r_act))), R(1000.0));
    J[271] = R(0) - dt * cmx_div((((-R(1)) * Ca_handling_by_the_SR__r_recov)), R(1000.0));
    J[272] = R(0);
    J[273] = R(0) - dt * cmx_div((((d_Ca_handling_by_the_SR__r_act__d1 * g[15])) - ((d_Ca_handling_by_the_SR__r_inact__d1 * g[16]))), R(1000.0));
    J[274] = R(0);
    J[275] = R(0) - dt * cmx_div(((d_Ca_handling_by_the_SR__r_act__d3 * g[15])), R(1000.0));
    J[276] = R(0);
    J[277] = R(0);
    J[278] = R(0);
    J[279] = R(0);
    J[280] = R(0);
    J[281] = R(0);
    J[282] = R(0);
    J[283] = R(0);
    J[284] = R(0);
    J[285] = R(0);
    J[286] = R(0);
    J[287] = R(0) - dt * cmx_div(((R(1) * Ca_handling_by_the_SR__r_act)), R(1000.0));
    J[288] = R(1) - dt * cmx_div((-((R(1) * Ca_handling_by_the_SR__r_inact))), R(1000.0));
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

}}  // namespace cmx::maleckar2008
