// Generated by chaste_amd/codegen/cellml2cpp.py from FaberRudy2000.cellml -- do not edit.
// Model: LR_Dynamic_model_2000
#pragma once
#include "model_common.h"

namespace cmx { namespace fr2000 {

constexpr int kNumStates = 25;
constexpr int kVIndex = 0;
constexpr int kNumParams = 4;
constexpr bool kPerCapacitanceUnits = true;

// State variables (CellML component__variable), V first then document order.
static const char* const kStateNames[kNumStates] = {
    "membrane__V",
    "fast_sodium_current_m_gate__m",
    "fast_sodium_current_h_gate__h",
    "fast_sodium_current_j_gate__j",
    "L_type_Ca_channel_d_gate__d",
    "L_type_Ca_channel_f_gate__f",
    "T_type_Ca_channel_b_gate__b",
    "T_type_Ca_channel_g_gate__g",
    "rapid_delayed_rectifier_potassium_current_xr_gate__xr",
    "slow_delayed_rectifier_potassium_current_xs1_gate__xs1",
    "slow_delayed_rectifier_potassium_current_xs2_gate__xs2",
    "transient_outward_current_zdv_gate__zdv",
    "transient_outward_current_ydv_gate__ydv",
    "calcium_dynamics__Cai",
    "calcium_dynamics__Ca_JSR",
    "calcium_dynamics__Ca_NSR",
    "calcium_dynamics__APtrack",
    "calcium_dynamics__APtrack2",
    "calcium_dynamics__APtrack3",
    "calcium_dynamics__Cainfluxtrack",
    "calcium_dynamics__OVRLDtrack",
    "calcium_dynamics__OVRLDtrack2",
    "calcium_dynamics__OVRLDtrack3",
    "ionic_concentrations__Nai",
    "ionic_concentrations__Ki",
};
static constexpr double kInitialState[kNumStates] = {
    -90.0,  // membrane__V
    0.0008,  // fast_sodium_current_m_gate__m
    0.993771,  // fast_sodium_current_h_gate__h
    0.995727,  // fast_sodium_current_j_gate__j
    3.210618e-06,  // L_type_Ca_channel_d_gate__d
    0.999837,  // L_type_Ca_channel_f_gate__f
    0.000970231,  // T_type_Ca_channel_b_gate__b
    0.994305,  // T_type_Ca_channel_g_gate__g
    0.000124042,  // rapid_delayed_rectifier_potassium_current_xr_gate__xr
    0.00445683,  // slow_delayed_rectifier_potassium_current_xs1_gate__xs1
    0.00445683,  // slow_delayed_rectifier_potassium_current_xs2_gate__xs2
    0.5,  // transient_outward_current_zdv_gate__zdv
    0.5,  // transient_outward_current_ydv_gate__ydv
    6e-05,  // calcium_dynamics__Cai
    1.8,  // calcium_dynamics__Ca_JSR
    1.8,  // calcium_dynamics__Ca_NSR
    0.0,  // calcium_dynamics__APtrack
    0.0,  // calcium_dynamics__APtrack2
    0.0,  // calcium_dynamics__APtrack3
    0.0,  // calcium_dynamics__Cainfluxtrack
    0.0,  // calcium_dynamics__OVRLDtrack
    0.0,  // calcium_dynamics__OVRLDtrack2
    0.0,  // calcium_dynamics__OVRLDtrack3
    9.0,  // ionic_concentrations__Nai
    141.2,  // ionic_concentrations__Ki
};
static const char* const kParamNames[4] = {
    "slow_delayed_rectifier_potassium_current__ScaleFactorGks",
    "transient_outward_current__ScaleFactorIto",
    "rapid_delayed_rectifier_potassium_current__ScaleFactorGkr",
    "fast_sodium_current__g_Na",
};
// The names the reference gives the same parameters (the CellML oxmeta term,
// else the variable's cmeta:id; empty when the file annotates neither).
static const char* const kParamAliases[4] = {
    "ScaleFactorGks",
    "ScaleFactorIto",
    "ScaleFactorGkr",
    "",
};
static constexpr double kParamDefaults[4] = {
    1.0,
    0.0,
    1.0,
    16.0,
};
// 1 = integrated by Rush-Larsen (affine gate, V-dependent coefficients only)
static constexpr int kIsGate[kNumStates] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

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
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.0);
    const R membrane__Cm = R(0.001);
    const R fast_sodium_current_m_gate__delta_m = R(1e-05);
    const R L_type_Ca_channel__gamma_Nai = R(0.75);
    const R L_type_Ca_channel__gamma_Nao = R(0.75);
    const R L_type_Ca_channel__gamma_Ki = R(0.75);
    const R L_type_Ca_channel__gamma_Ko = R(0.75);
    const R L_type_Ca_channel__P_Ca = R(0.00054);
    const R L_type_Ca_channel__P_Na = R(6.75e-07);
    const R L_type_Ca_channel__P_K = R(1.93e-07);
    const R L_type_Ca_channel__gamma_Cai = R(1.0);
    const R L_type_Ca_channel__gamma_Cao = R(0.341);
    const R L_type_Ca_channel_f_Ca_gate__Km_Ca = R(0.0006);
    const R T_type_Ca_channel__g_CaT = R(0.05);
    const R slow_delayed_rectifier_potassium_current__PNaK = R(0.01833);
    const R plateau_potassium_current__g_Kp = R(0.00552);
    const R sodium_activated_potassium_current__nKNa = R(2.8);
    const R sodium_activated_potassium_current__kdKNa = R(66.0);
    const R ATP_sensitive_potassium_current__i_K_ATP_on = R(1.0);
    const R ATP_sensitive_potassium_current__nATP = R(0.24);
    const R ATP_sensitive_potassium_current__nicholsarea = R(5e-05);
    const R ATP_sensitive_potassium_current__ATPi = R(3.0);
    const R ATP_sensitive_potassium_current__hATP = R(2.0);
    const R ATP_sensitive_potassium_current__kATP = R(0.00025);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.0005);
    const R sarcolemmal_calcium_pump__I_pCa = R(1.15);
    const R sodium_background_current__g_Nab = R(0.004);
    const R calcium_background_current__g_Cab = R(0.003016);
    const R sodium_potassium_pump__I_NaK = R(2.25);
    const R sodium_potassium_pump__K_mNai = R(10.0);
    const R sodium_potassium_pump__K_mKo = R(1.5);
    const R non_specific_calcium_activated_current__K_m_ns_Ca = R(0.0012);
    const R Na_Ca_exchanger__c1 = R(0.00025);
    const R Na_Ca_exchanger__c2 = R(0.0001);
    const R Na_Ca_exchanger__gamma = R(0.15);
    const R calcium_dynamics__G_rel_max = R(60000.0);
    const R calcium_dynamics__G_rel_overload = R(4000.0);
    const R calcium_dynamics__tau_tr = R(0.18);
    const R calcium_dynamics__K_mrel = R(0.0008);
    const R calcium_dynamics__delta_Ca_ith = R(0.00018);
    const R calcium_dynamics__CSQN_max = R(10.0);
    const R calcium_dynamics__K_mCSQN = R(0.8);
    const R calcium_dynamics__K_mup = R(0.00092);
    const R calcium_dynamics__I_up = R(8.75);
    const R calcium_dynamics__Ca_NSR_max = R(15.0);
    const R calcium_dynamics__Cao = R(1.8);
    const R calcium_dynamics__K_mTn = R(0.0005);
    const R calcium_dynamics__K_mCMDN = R(0.00238);
    const R calcium_dynamics__Tn_max = R(0.07);
    const R calcium_dynamics__CMDN_max = R(0.05);
    const R calcium_dynamics__CSQNthresh = R(0.7);
    const R calcium_dynamics__Logicthresh = R(0.98);
    const R ionic_concentrations__Nao = R(132.0);
    const R ionic_concentrations__Ko = R(4.5);
    const R ionic_concentrations__A_cap = R(1.434e-07);
    const R ionic_concentrations__preplength = R(0.001);
    const R ionic_concentrations__radius = R(0.00011);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R ATP_sensitive_potassium_current__g_K_ATP = ((ATP_sensitive_potassium_current__i_K_ATP_on * R(0.000193)) * (R(1) / ATP_sensitive_potassium_current__nicholsarea));
    const R ATP_sensitive_potassium_current__pATP = (R(1.0) * (R(1) / (R(1.0) + pow((ATP_sensitive_potassium_current__ATPi * (R(1) / ATP_sensitive_potassium_current__kATP)), ATP_sensitive_potassium_current__hATP))));
    const R ATP_sensitive_potassium_current__GKbaraATP = (ATP_sensitive_potassium_current__g_K_ATP * ATP_sensitive_potassium_current__pATP * pow((ionic_concentrations__Ko * (R(1) / R(4.0))), ATP_sensitive_potassium_current__nATP));
    const R time_independent_potassium_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Ko, y[24])));
    const R ATP_sensitive_potassium_current__i_K_ATP = (ATP_sensitive_potassium_current__GKbaraATP * (V - time_independent_potassium_current__E_K));
    const R L_type_Ca_channel__I_CaCa = cmx_remove_singularity(V, R(-0.0), R(1.3356169352749133e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Ca * cmx_ipow<2>(R(2.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Cai * y[13] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801712697600754, 0.0)) - (L_type_Ca_channel__gamma_Cao * calcium_dynamics__Cao))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel_f_Ca_gate__f_Ca = cmx_div(R(1.0), (R(1.0) + (y[13] * (R(1) / L_type_Ca_channel_f_Ca_gate__Km_Ca))));
    const R L_type_Ca_channel__i_CaCa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaCa);
    const R L_type_Ca_channel__I_CaK = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_K * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * y[24] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaK = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaK);
    const R L_type_Ca_channel__I_CaNa = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Na * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * y[23] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaNa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaNa);
    const R L_type_Ca_channel_d_gate__E0_d = (V + R(10.0));
    const R L_type_Ca_channel_d_gate__d_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_c(((-L_type_Ca_channel_d_gate__E0_d) * (R(1) / R(6.24))), (-L_type_Ca_channel_d_gate__E0_d), 0.23120112834759027)));
    const R L_type_Ca_channel_d_gate__tau_d = (((fabs(L_type_Ca_channel_d_gate__E0_d) < R(1e-05))) ? ((R(0.001) * (R(1) / (R(0.035) * R(6.24))))) : (cmx_div((R(0.001) * L_type_Ca_channel_d_gate__d_infinity * (-cmx_expm1(((-L_type_Ca_channel_d_gate__E0_d) * (R(1) / R(6.24)))))), (R(0.035) * L_type_Ca_channel_d_gate__E0_d))));
    const R L_type_Ca_channel_d_gate__alpha_d = cmx_div(L_type_Ca_channel_d_gate__d_infinity, L_type_Ca_channel_d_gate__tau_d);
    const R L_type_Ca_channel_d_gate__beta_d = cmx_div((R(1.0) - L_type_Ca_channel_d_gate__d_infinity), L_type_Ca_channel_d_gate__tau_d);
    const R L_type_Ca_channel_f_gate__f_infinity = (cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(32.0)) * (R(1) / R(8.0))), V, 0.18033688011112042, 5.7707801635558535))) + cmx_div(R(0.6), (R(1.0) + cmx_exp_v(((R(50.0) - V) * (R(1) / R(20.0))), V, -0.07213475204444818, 3.6067376022224087))));
    const R L_type_Ca_channel_f_gate__tau_f = cmx_div(R(0.001), ((R(0.0197) * cmx_exp((-cmx_ipow<2>((R(0.0337) * (V + R(10.0))))))) + R(0.02)));
    const R L_type_Ca_channel_f_gate__alpha_f = cmx_div(L_type_Ca_channel_f_gate__f_infinity, L_type_Ca_channel_f_gate__tau_f);
    const R L_type_Ca_channel_f_gate__beta_f = cmx_div((R(1.0) - L_type_Ca_channel_f_gate__f_infinity), L_type_Ca_channel_f_gate__tau_f);
    const R Na_Ca_exchanger__i_NaCa = cmx_div((Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(y[23]) * calcium_dynamics__Cao) - (cmx_ipow<3>(ionic_concentrations__Nao) * y[13]))), (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(y[23]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * y[13])))));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(calcium_dynamics__Cao, y[13])));
    const R T_type_Ca_channel__i_Ca_T = (T_type_Ca_channel__g_CaT * y[6] * y[6] * y[7] * (V - calcium_background_current__E_Ca));
    const R T_type_Ca_channel_b_gate__b_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(14.0))) * (R(1) / R(10.8))), V, -0.13358287415638548, -1.870160238189397)));
    const R T_type_Ca_channel_b_gate__tau_b = (R(0.0037) + cmx_div(R(0.0061), (R(1.0) + cmx_exp_v(((V + R(25.0)) * (R(1) / R(4.5))), V, 0.3205988979753252, 8.01497244938313))));
    const R T_type_Ca_channel_g_gate__g_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(60.0)) * (R(1) / R(5.6))), V, 0.25762411444445776, 15.457446866667466)));
    const R T_type_Ca_channel_g_gate__tau_g = (((V <= R(0.0))) ? ((((-R(0.000875)) * V) + R(0.012))) : (R(0.012)));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R ionic_concentrations__volume = (R(3.14159265358979323846) * ionic_concentrations__preplength * cmx_ipow<2>(ionic_concentrations__radius));
    const R ionic_concentrations__V_myo = (R(0.68) * ionic_concentrations__volume);
    const R calcium_dynamics__V_JSR = ((R(0.0048) * (R(1) / R(0.68))) * ionic_concentrations__V_myo);
    const R calcium_dynamics__V_NSR = ((R(0.0552) * (R(1) / R(0.68))) * ionic_concentrations__V_myo);
    const R calcium_dynamics__K_leak = (calcium_dynamics__I_up * (R(1) / calcium_dynamics__Ca_NSR_max));
    const R calcium_dynamics__i_leak = (calcium_dynamics__K_leak * y[15]);
    const R calcium_dynamics__G_rel = (((y[19] > calcium_dynamics__delta_Ca_ith)) ? ((cmx_div((calcium_dynamics__G_rel_max * (y[19] - calcium_dynamics__delta_Ca_ith)), ((calcium_dynamics__K_mrel + y[19]) - calcium_dynamics__delta_Ca_ith)) * (R(1.0) - y[17]) * y[17])) : (((((y[19] <= calcium_dynamics__delta_Ca_ith) && (y[21] > R(0.0)))) ? ((calcium_dynamics__G_rel_overload * (R(1.0) - y[21]) * y[21])) : (R(0.0)))));
    const R calcium_dynamics__i_rel = (calcium_dynamics__G_rel * (y[14] - y[13]));
    const R calcium_dynamics__i_tr = ((y[15] - y[14]) * (R(1) / calcium_dynamics__tau_tr));
    const R calcium_dynamics__i_up = cmx_div((calcium_dynamics__I_up * y[13]), (y[13] + calcium_dynamics__K_mup));
    const R fast_sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Nao, y[23])));
    const R fast_sodium_current__i_Na = (p[3] * cmx_ipow<3>(y[1]) * y[2] * y[3] * (V - fast_sodium_current__E_Na));
    const R fast_sodium_current_h_gate__alpha_h = (((V < (-R(40.0)))) ? ((R(135.0) * cmx_exp_v(((R(80.0) + V) * (R(1) / (-R(6.8)))), V, -0.2121610354248476, -16.972882833987807))) : (R(0.0)));
    const R fast_sodium_current_h_gate__beta_h = (((V < (-R(40.0)))) ? (((R(3560.0) * cmx_exp_v((R(0.079) * V), V, 0.11397290823022813, 0.0)) + (R(310000000.0) * cmx_exp_v((R(0.35) * V), V, 0.5049432643111371, 0.0)))) : (cmx_div(R(1000.0), (R(0.13) * (R(1.0) + cmx_exp_v(((V + R(10.66)) * (R(1) / (-R(11.1)))), V, -0.12997252620621294, -1.3855071293582297))))));
    const R fast_sodium_current_j_gate__alpha_j = (((V < (-R(40.0)))) ? (cmx_div((R(1000.0) * (-((R(127140.0) * cmx_exp_v((R(0.2444) * V), V, 0.3525946679932627, 0.0)) + (R(3.474e-05) * cmx_exp_v(((-R(0.04391)) * V), V, -0.06334873924543438, -0.0)))) * (V + R(37.78))), (R(1.0) + cmx_exp_v((R(0.311) * (V + R(79.23))), V, 0.44867815771646763, 35.54877043587573)))) : (R(0.0)));
    const R fast_sodium_current_j_gate__beta_j = (((V < (-R(40.0)))) ? (cmx_div((R(121.2) * cmx_exp_v(((-R(0.01052)) * V), V, -0.015177151830151895, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.1378)) * (V + R(40.14))), V, -0.19880337663449915, -7.979967538108797)))) : (cmx_div((R(300.0) * cmx_exp_v(((-R(2.535e-07)) * V), V, -3.657231928653522e-07, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(32.0))), V, -0.14426950408889636, -4.616624130844683)))));
    const R fast_sodium_current_m_gate__E0_m = (V + R(47.13));
    const R fast_sodium_current_m_gate__alpha_m = (((fabs(fast_sodium_current_m_gate__E0_m) >= fast_sodium_current_m_gate__delta_m)) ? (cmx_div((R(320.0) * fast_sodium_current_m_gate__E0_m), (-cmx_expm1(((-R(0.1)) * fast_sodium_current_m_gate__E0_m))))) : (R(3200.0)));
    const R fast_sodium_current_m_gate__beta_m = (R(80.0) * cmx_exp_v(((-V) * (R(1) / R(11.0))), V, -0.1311540946262694, -0.0));
    const R L_type_Ca_channel__i_Ca_L = (L_type_Ca_channel__i_CaCa + L_type_Ca_channel__i_CaK + L_type_Ca_channel__i_CaNa);
    const R non_specific_calcium_activated_current__P_ns_Ca = (R(0.0) * R(1.75e-07));
    const R non_specific_calcium_activated_current__I_ns_K = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * y[24] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_K = cmx_div((non_specific_calcium_activated_current__I_ns_K * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, y[13]))));
    const R non_specific_calcium_activated_current__I_ns_Na = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * y[23] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_Na = cmx_div((non_specific_calcium_activated_current__I_ns_Na * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, y[13]))));
    const R non_specific_calcium_activated_current__i_ns_Ca = (non_specific_calcium_activated_current__i_ns_Na + non_specific_calcium_activated_current__i_ns_K);
    const R plateau_potassium_current__Kp = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(7.488) - V) * (R(1) / R(5.98))), V, -0.2412533513192246, 1.806505094678354)));
    const R plateau_potassium_current__i_Kp = (plateau_potassium_current__g_Kp * plateau_potassium_current__Kp * (V - time_independent_potassium_current__E_K));
    const R rapid_delayed_rectifier_potassium_current__Rect = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003)));
    const R rapid_delayed_rectifier_potassium_current__g_Kr = (p[2] * R(0.02614) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (rapid_delayed_rectifier_potassium_current__g_Kr * y[8] * rapid_delayed_rectifier_potassium_current__Rect * (V - time_independent_potassium_current__E_K));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__I_pCa * y[13]), (sarcolemmal_calcium_pump__K_mpCa + y[13]));
    const R slow_delayed_rectifier_potassium_current__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)), (y[24] + (slow_delayed_rectifier_potassium_current__PNaK * y[23])))));
    const R slow_delayed_rectifier_potassium_current__g_Ks = (p[0] * R(0.433) * (R(1.0) + cmx_div(R(0.6), (R(1.0) + pow(cmx_div(R(3.8e-05), y[13]), R(1.4))))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (slow_delayed_rectifier_potassium_current__g_Ks * y[9] * y[10] * (V - slow_delayed_rectifier_potassium_current__E_Ks));
    const R sodium_activated_potassium_current__g_K_Na = (R(0.0) * R(0.12848));
    const R sodium_activated_potassium_current__pona = cmx_div(R(0.85), (R(1.0) + pow(cmx_div(sodium_activated_potassium_current__kdKNa, y[23]), sodium_activated_potassium_current__nKNa)));
    const R sodium_activated_potassium_current__pov = (R(0.8) - cmx_div(R(0.65), (R(1.0) + cmx_exp_v(((V + R(125.0)) * (R(1) / R(15.0))), V, 0.09617966939259756, 12.022458674074695))));
    const R sodium_activated_potassium_current__i_K_Na = (sodium_activated_potassium_current__g_K_Na * sodium_activated_potassium_current__pona * sodium_activated_potassium_current__pov * (V - time_independent_potassium_current__E_K));
    const R sodium_background_current__i_Na_b = (sodium_background_current__g_Nab * (V - fast_sodium_current__E_Na));
    const R sodium_potassium_pump__sigma = ((R(1.0) * (R(1) / R(7.0))) * cmx_expm1((ionic_concentrations__Nao * (R(1) / R(67.3)))));
    const R sodium_potassium_pump__f_NaK = cmx_div(R(1.0), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400856348800377, -0.0)) + (R(0.0365) * sodium_potassium_pump__sigma * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400856348800377, -0.0))));
    const R sodium_potassium_pump__i_NaK = ((cmx_div((sodium_potassium_pump__I_NaK * sodium_potassium_pump__f_NaK * R(1.0)), (R(1.0) + cmx_ipow<2>(cmx_div(sodium_potassium_pump__K_mNai, y[23])))) * ionic_concentrations__Ko) * (R(1) / (ionic_concentrations__Ko + sodium_potassium_pump__K_mKo)));
    const R time_independent_potassium_current__g_K1 = (R(0.75) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R time_independent_potassium_current_K1_gate__alpha_K1 = cmx_div(R(1020.0), (R(1.0) + cmx_exp_c((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215))), (((V - time_independent_potassium_current__E_K) - R(59.215))), 0.34408276725201775)));
    const R time_independent_potassium_current_K1_gate__beta_K1 = cmx_div((R(1000.0) * ((R(0.49124) * cmx_exp_c((R(0.08032) * ((V - time_independent_potassium_current__E_K) + R(5.476))), (((V - time_independent_potassium_current__E_K) + R(5.476))), 0.11587726568420155)) + cmx_exp_c((R(0.06175) * ((V - time_independent_potassium_current__E_K) - R(594.31))), (((V - time_independent_potassium_current__E_K) - R(594.31))), 0.08908641877489348))), (R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938)));
    const R time_independent_potassium_current_K1_gate__K1_infinity = cmx_div(time_independent_potassium_current_K1_gate__alpha_K1, (time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1));
    const R time_independent_potassium_current__i_K1 = (time_independent_potassium_current__g_K1 * time_independent_potassium_current_K1_gate__K1_infinity * (V - time_independent_potassium_current__E_K));
    const R transient_outward_current__g_to = (p[1] * R(0.5));
    const R transient_outward_current__rvdv = cmx_exp_v((V * (R(1) / R(100.0))), V, 0.014426950408889635, 0.0);
    const R transient_outward_current__i_to = (transient_outward_current__g_to * cmx_ipow<3>(y[11]) * y[12] * transient_outward_current__rvdv * (V - time_independent_potassium_current__E_K));
    const R membrane__dVdt = (((-R(1.0)) * (R(1) / membrane__Cm)) * (fast_sodium_current__i_Na + L_type_Ca_channel__i_Ca_L + T_type_Ca_channel__i_Ca_T + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + sodium_activated_potassium_current__i_K_Na + ATP_sensitive_potassium_current__i_K_ATP + transient_outward_current__i_to + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + Na_Ca_exchanger__i_NaCa + sarcolemmal_calcium_pump__i_p_Ca + sodium_background_current__i_Na_b + calcium_background_current__i_Ca_b + sodium_potassium_pump__i_NaK + non_specific_calcium_activated_current__i_ns_Ca + i_stim));
    const R rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr = cmx_remove_singularity(V, R(-38.90000000000011), R(6.896551724137951e-07), [&](const R V) { return cmx_remove_singularity(V, R(-14.2), R(8.130081300813008e-07), [&](const R V) { return cmx_div(R(0.001), (cmx_div((R(0.00138) * (V + R(14.2))), (-cmx_expm1(((-R(0.123)) * (V + R(14.2)))))) + cmx_div((R(0.00061) * (V + R(38.9))), cmx_expm1((R(0.145) * (V + R(38.9))))))); }); });
    const R rapid_delayed_rectifier_potassium_current_xr_gate__xr_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(21.5))) * (R(1) / R(7.5))), V, -0.19235933878519512, -4.135725783881695)));
    const R slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1 = cmx_remove_singularity(V, R(-29.99999999999988), R(6.75675675675673e-07), [&](const R V) { return cmx_div(R(0.001), (cmx_div((R(7.19e-05) * (V + R(30.0))), (-cmx_expm1(((-R(0.148)) * (V + R(30.0)))))) + cmx_div((R(0.000131) * (V + R(30.0))), cmx_expm1((R(0.0687) * (V + R(30.0))))))); });
    const R slow_delayed_rectifier_potassium_current_xs1_gate__xs1_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V - R(1.5))) * (R(1) / R(16.7))), V, -0.08638892460412954, 0.1295833869061943)));
    const R slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2 = cmx_remove_singularity(V, R(-29.99999999999988), R(6.75675675675673e-07), [&](const R V) { return cmx_div((R(4.0) * R(0.001)), (cmx_div((R(7.19e-05) * (V + R(30.0))), (-cmx_expm1(((-R(0.148)) * (V + R(30.0)))))) + cmx_div((R(0.000131) * (V + R(30.0))), cmx_expm1((R(0.0687) * (V + R(30.0))))))); });
    const R slow_delayed_rectifier_potassium_current_xs2_gate__xs2_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V - R(1.5))) * (R(1) / R(16.7))), V, -0.08638892460412954, 0.1295833869061943)));
    const R transient_outward_current_ydv_gate__alpha_ydv = cmx_div(R(15.0), (R(1.0) + cmx_exp_v(((V + R(60.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 17.31234049066756)));
    const R transient_outward_current_ydv_gate__beta_ydv = cmx_div((R(100.0) * cmx_exp_v(((V + R(25.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 7.213475204444817)), (R(1.0) + cmx_exp_v(((V + R(25.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 7.213475204444817)));
    const R transient_outward_current_ydv_gate__tau_ydv = cmx_div(R(1.0), (transient_outward_current_ydv_gate__alpha_ydv + transient_outward_current_ydv_gate__beta_ydv));
    const R transient_outward_current_ydv_gate__ydv_ss = cmx_div(transient_outward_current_ydv_gate__alpha_ydv, (transient_outward_current_ydv_gate__alpha_ydv + transient_outward_current_ydv_gate__beta_ydv));
    const R transient_outward_current_zdv_gate__alpha_zdv = cmx_div((R(10000.0) * cmx_exp_v(((V - R(40.0)) * (R(1) / R(25.0))), V, 0.05770780163555854, -2.3083120654223417)), (R(1.0) + cmx_exp_v(((V - R(40.0)) * (R(1) / R(25.0))), V, 0.05770780163555854, -2.3083120654223417)));
    const R transient_outward_current_zdv_gate__beta_zdv = cmx_div((R(10000.0) * cmx_exp_v(((-(V + R(90.0))) * (R(1) / R(25.0))), V, -0.05770780163555854, -5.193702147200268)), (R(1.0) + cmx_exp_v(((-(V + R(90.0))) * (R(1) / R(25.0))), V, -0.05770780163555854, -5.193702147200268)));
    const R transient_outward_current_zdv_gate__tau_zdv = cmx_div(R(1.0), (transient_outward_current_zdv_gate__alpha_zdv + transient_outward_current_zdv_gate__beta_zdv));
    const R transient_outward_current_zdv_gate__zdv_ss = cmx_div(transient_outward_current_zdv_gate__alpha_zdv, (transient_outward_current_zdv_gate__alpha_zdv + transient_outward_current_zdv_gate__beta_zdv));
    if (RL) {
        g_k[1] = cmx_div(((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1))), R(1000.0));
        g_inf[1] = cmx_div(-((fast_sodium_current_m_gate__alpha_m * R(1.0))), ((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1))));
    } else {
        dy[1] = (((fast_sodium_current_m_gate__alpha_m * (R(1.0) - y[1])) - (fast_sodium_current_m_gate__beta_m * y[1])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[2] = cmx_div(((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1))), R(1000.0));
        g_inf[2] = cmx_div(-((fast_sodium_current_h_gate__alpha_h * R(1.0))), ((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1))));
    } else {
        dy[2] = (((fast_sodium_current_h_gate__alpha_h * (R(1.0) - y[2])) - (fast_sodium_current_h_gate__beta_h * y[2])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[3] = cmx_div(((fast_sodium_current_j_gate__alpha_j * (-R(1))) - (fast_sodium_current_j_gate__beta_j * R(1))), R(1000.0));
        g_inf[3] = cmx_div(-((fast_sodium_current_j_gate__alpha_j * R(1.0))), ((fast_sodium_current_j_gate__alpha_j * (-R(1))) - (fast_sodium_current_j_gate__beta_j * R(1))));
    } else {
        dy[3] = (((fast_sodium_current_j_gate__alpha_j * (R(1.0) - y[3])) - (fast_sodium_current_j_gate__beta_j * y[3])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[4] = cmx_div(((L_type_Ca_channel_d_gate__alpha_d * (-R(1))) - (L_type_Ca_channel_d_gate__beta_d * R(1))), R(1000.0));
        g_inf[4] = cmx_div(-((L_type_Ca_channel_d_gate__alpha_d * R(1.0))), ((L_type_Ca_channel_d_gate__alpha_d * (-R(1))) - (L_type_Ca_channel_d_gate__beta_d * R(1))));
    } else {
        dy[4] = (((L_type_Ca_channel_d_gate__alpha_d * (R(1.0) - y[4])) - (L_type_Ca_channel_d_gate__beta_d * y[4])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[5] = cmx_div(((L_type_Ca_channel_f_gate__alpha_f * (-R(1))) - (L_type_Ca_channel_f_gate__beta_f * R(1))), R(1000.0));
        g_inf[5] = cmx_div(-((L_type_Ca_channel_f_gate__alpha_f * R(1.0))), ((L_type_Ca_channel_f_gate__alpha_f * (-R(1))) - (L_type_Ca_channel_f_gate__beta_f * R(1))));
    } else {
        dy[5] = (((L_type_Ca_channel_f_gate__alpha_f * (R(1.0) - y[5])) - (L_type_Ca_channel_f_gate__beta_f * y[5])) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[6] = cmx_div(((-R(1)) / T_type_Ca_channel_b_gate__tau_b), R(1000.0));
        g_inf[6] = cmx_div(-((T_type_Ca_channel_b_gate__b_inf / T_type_Ca_channel_b_gate__tau_b)), ((-R(1)) / T_type_Ca_channel_b_gate__tau_b));
    } else {
        dy[6] = (cmx_div((T_type_Ca_channel_b_gate__b_inf - y[6]), T_type_Ca_channel_b_gate__tau_b) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[7] = cmx_div(((-R(1)) / T_type_Ca_channel_g_gate__tau_g), R(1000.0));
        g_inf[7] = cmx_div(-((T_type_Ca_channel_g_gate__g_inf / T_type_Ca_channel_g_gate__tau_g)), ((-R(1)) / T_type_Ca_channel_g_gate__tau_g));
    } else {
        dy[7] = (cmx_div((T_type_Ca_channel_g_gate__g_inf - y[7]), T_type_Ca_channel_g_gate__tau_g) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[8] = cmx_div(((-R(1)) / rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr), R(1000.0));
        g_inf[8] = cmx_div(-((rapid_delayed_rectifier_potassium_current_xr_gate__xr_infinity / rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr)), ((-R(1)) / rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr));
    } else {
        dy[8] = (cmx_div((rapid_delayed_rectifier_potassium_current_xr_gate__xr_infinity - y[8]), rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[9] = cmx_div(((-R(1)) / slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1), R(1000.0));
        g_inf[9] = cmx_div(-((slow_delayed_rectifier_potassium_current_xs1_gate__xs1_infinity / slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1)), ((-R(1)) / slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1));
    } else {
        dy[9] = (cmx_div((slow_delayed_rectifier_potassium_current_xs1_gate__xs1_infinity - y[9]), slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[10] = cmx_div(((-R(1)) / slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2), R(1000.0));
        g_inf[10] = cmx_div(-((slow_delayed_rectifier_potassium_current_xs2_gate__xs2_infinity / slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2)), ((-R(1)) / slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2));
    } else {
        dy[10] = (cmx_div((slow_delayed_rectifier_potassium_current_xs2_gate__xs2_infinity - y[10]), slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[11] = cmx_div(((-R(1)) / transient_outward_current_zdv_gate__tau_zdv), R(1000.0));
        g_inf[11] = cmx_div(-((transient_outward_current_zdv_gate__zdv_ss / transient_outward_current_zdv_gate__tau_zdv)), ((-R(1)) / transient_outward_current_zdv_gate__tau_zdv));
    } else {
        dy[11] = (cmx_div((transient_outward_current_zdv_gate__zdv_ss - y[11]), transient_outward_current_zdv_gate__tau_zdv) * (R(1) / R(1000.0)));
    }
    if (RL) {
        g_k[12] = cmx_div(((-R(1)) / transient_outward_current_ydv_gate__tau_ydv), R(1000.0));
        g_inf[12] = cmx_div(-((transient_outward_current_ydv_gate__ydv_ss / transient_outward_current_ydv_gate__tau_ydv)), ((-R(1)) / transient_outward_current_ydv_gate__tau_ydv));
    } else {
        dy[12] = (cmx_div((transient_outward_current_ydv_gate__ydv_ss - y[12]), transient_outward_current_ydv_gate__tau_ydv) * (R(1) / R(1000.0)));
    }
    dy[13] = ((cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + y[13]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + y[13]))))) * ((((-ionic_concentrations__A_cap) * (((L_type_Ca_channel__i_CaCa + T_type_Ca_channel__i_Ca_T) - (R(2.0) * Na_Ca_exchanger__i_NaCa)) + sarcolemmal_calcium_pump__i_p_Ca + calcium_background_current__i_Ca_b)) * (R(1) / (R(2.0) * ionic_concentrations__V_myo * membrane__F))) + ((calcium_dynamics__i_rel * calcium_dynamics__V_JSR) * (R(1) / ionic_concentrations__V_myo)) + (((calcium_dynamics__i_leak - calcium_dynamics__i_up) * calcium_dynamics__V_NSR) * (R(1) / ionic_concentrations__V_myo)))) * (R(1) / R(1000.0)));
    dy[14] = ((cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + y[14]))))) * (calcium_dynamics__i_tr - calcium_dynamics__i_rel)) * (R(1) / R(1000.0)));
    dy[15] = ((((((-calcium_dynamics__i_tr) * calcium_dynamics__V_JSR) * (R(1) / calcium_dynamics__V_NSR)) - calcium_dynamics__i_leak) + calcium_dynamics__i_up) * (R(1) / R(1000.0)));
    dy[16] = ((((membrane__dVdt > R(150000.0))) ? (((R(100000.0) * (R(1.0) - y[16])) - (R(500.0) * y[16]))) : (((-R(500.0)) * y[16]))) * (R(1) / R(1000.0)));
    dy[17] = (((((y[16] < R(0.2)) && (y[16] > R(0.18)))) ? (((R(100000.0) * (R(1.0) - y[17])) - (R(500.0) * y[17]))) : (((-R(500.0)) * y[17]))) * (R(1) / R(1000.0)));
    dy[18] = (((((y[16] < R(0.2)) && (y[16] > R(0.18)))) ? (((R(100000.0) * (R(1.0) - y[18])) - (R(500.0) * y[18]))) : (((-R(10.0)) * y[18]))) * (R(1) / R(1000.0)));
    dy[19] = ((((y[16] > R(0.2))) ? ((((-ionic_concentrations__A_cap) * (((L_type_Ca_channel__i_CaCa + T_type_Ca_channel__i_Ca_T) - Na_Ca_exchanger__i_NaCa) + sarcolemmal_calcium_pump__i_p_Ca + calcium_background_current__i_Ca_b)) * (R(1) / (R(2.0) * ionic_concentrations__V_myo * membrane__F)))) : (((((y[17] > R(0.01)) && (y[16] <= R(0.2)))) ? (R(0.0)) : (((-R(500.0)) * y[19]))))) * (R(1) / R(1000.0)));
    dy[20] = (((((cmx_div(R(1.0), (R(1.0) + cmx_div(calcium_dynamics__K_mCSQN, y[14]))) > calcium_dynamics__CSQNthresh) && (y[22] < R(0.37)) && (y[18] < R(0.37)))) ? ((R(50000.0) * (R(1.0) - y[20]))) : (((-R(500.0)) * y[20]))) * (R(1) / R(1000.0)));
    dy[21] = (((((y[20] > calcium_dynamics__Logicthresh) && (y[21] < calcium_dynamics__Logicthresh))) ? ((R(50000.0) * (R(1.0) - y[21]))) : (((-R(500.0)) * y[21]))) * (R(1) / R(1000.0)));
    dy[22] = (((((y[20] > calcium_dynamics__Logicthresh) && (y[22] < calcium_dynamics__Logicthresh))) ? ((R(50000.0) * (R(1.0) - y[22]))) : (((-R(10.0)) * y[22]))) * (R(1) / R(1000.0)));
    dy[23] = ((((-(fast_sodium_current__i_Na + L_type_Ca_channel__i_CaNa + sodium_background_current__i_Na_b + non_specific_calcium_activated_current__i_ns_Na + (Na_Ca_exchanger__i_NaCa * R(3.0)) + (sodium_potassium_pump__i_NaK * R(3.0)))) * ionic_concentrations__A_cap) * (R(1) / (ionic_concentrations__V_myo * membrane__F))) * (R(1) / R(1000.0)));
    dy[24] = ((((-(L_type_Ca_channel__i_CaK + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + sodium_activated_potassium_current__i_K_Na + ATP_sensitive_potassium_current__i_K_ATP + transient_outward_current__i_to + non_specific_calcium_activated_current__i_ns_K + ((-sodium_potassium_pump__i_NaK) * R(2.0)))) * ionic_concentrations__A_cap) * (R(1) / (ionic_concentrations__V_myo * membrane__F))) * (R(1) / R(1000.0)));
}

// dV/dt including the stimulus (single-cell integration).
template <class R>
CMX_HD inline R dvdt(const R* y, R V, R i_stim_area, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.0);
    const R membrane__Cm = R(0.001);
    const R L_type_Ca_channel__gamma_Nai = R(0.75);
    const R L_type_Ca_channel__gamma_Nao = R(0.75);
    const R L_type_Ca_channel__gamma_Ki = R(0.75);
    const R L_type_Ca_channel__gamma_Ko = R(0.75);
    const R L_type_Ca_channel__P_Ca = R(0.00054);
    const R L_type_Ca_channel__P_Na = R(6.75e-07);
    const R L_type_Ca_channel__P_K = R(1.93e-07);
    const R L_type_Ca_channel__gamma_Cai = R(1.0);
    const R L_type_Ca_channel__gamma_Cao = R(0.341);
    const R L_type_Ca_channel_f_Ca_gate__Km_Ca = R(0.0006);
    const R T_type_Ca_channel__g_CaT = R(0.05);
    const R slow_delayed_rectifier_potassium_current__PNaK = R(0.01833);
    const R plateau_potassium_current__g_Kp = R(0.00552);
    const R sodium_activated_potassium_current__nKNa = R(2.8);
    const R sodium_activated_potassium_current__kdKNa = R(66.0);
    const R ATP_sensitive_potassium_current__i_K_ATP_on = R(1.0);
    const R ATP_sensitive_potassium_current__nATP = R(0.24);
    const R ATP_sensitive_potassium_current__nicholsarea = R(5e-05);
    const R ATP_sensitive_potassium_current__ATPi = R(3.0);
    const R ATP_sensitive_potassium_current__hATP = R(2.0);
    const R ATP_sensitive_potassium_current__kATP = R(0.00025);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.0005);
    const R sarcolemmal_calcium_pump__I_pCa = R(1.15);
    const R sodium_background_current__g_Nab = R(0.004);
    const R calcium_background_current__g_Cab = R(0.003016);
    const R sodium_potassium_pump__I_NaK = R(2.25);
    const R sodium_potassium_pump__K_mNai = R(10.0);
    const R sodium_potassium_pump__K_mKo = R(1.5);
    const R non_specific_calcium_activated_current__K_m_ns_Ca = R(0.0012);
    const R Na_Ca_exchanger__c1 = R(0.00025);
    const R Na_Ca_exchanger__c2 = R(0.0001);
    const R Na_Ca_exchanger__gamma = R(0.15);
    const R calcium_dynamics__Cao = R(1.8);
    const R ionic_concentrations__Nao = R(132.0);
    const R ionic_concentrations__Ko = R(4.5);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R ATP_sensitive_potassium_current__g_K_ATP = ((ATP_sensitive_potassium_current__i_K_ATP_on * R(0.000193)) * (R(1) / ATP_sensitive_potassium_current__nicholsarea));
    const R ATP_sensitive_potassium_current__pATP = (R(1.0) * (R(1) / (R(1.0) + pow((ATP_sensitive_potassium_current__ATPi * (R(1) / ATP_sensitive_potassium_current__kATP)), ATP_sensitive_potassium_current__hATP))));
    const R ATP_sensitive_potassium_current__GKbaraATP = (ATP_sensitive_potassium_current__g_K_ATP * ATP_sensitive_potassium_current__pATP * pow((ionic_concentrations__Ko * (R(1) / R(4.0))), ATP_sensitive_potassium_current__nATP));
    const R time_independent_potassium_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Ko, y[24])));
    const R ATP_sensitive_potassium_current__i_K_ATP = (ATP_sensitive_potassium_current__GKbaraATP * (V - time_independent_potassium_current__E_K));
    const R L_type_Ca_channel__I_CaCa = cmx_remove_singularity(V, R(-0.0), R(1.3356169352749133e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Ca * cmx_ipow<2>(R(2.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Cai * y[13] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801712697600754, 0.0)) - (L_type_Ca_channel__gamma_Cao * calcium_dynamics__Cao))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel_f_Ca_gate__f_Ca = cmx_div(R(1.0), (R(1.0) + (y[13] * (R(1) / L_type_Ca_channel_f_Ca_gate__Km_Ca))));
    const R L_type_Ca_channel__i_CaCa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaCa);
    const R L_type_Ca_channel__I_CaK = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_K * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * y[24] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaK = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaK);
    const R L_type_Ca_channel__I_CaNa = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Na * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * y[23] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaNa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaNa);
    const R L_type_Ca_channel__i_Ca_L = (L_type_Ca_channel__i_CaCa + L_type_Ca_channel__i_CaK + L_type_Ca_channel__i_CaNa);
    const R Na_Ca_exchanger__i_NaCa = cmx_div((Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(y[23]) * calcium_dynamics__Cao) - (cmx_ipow<3>(ionic_concentrations__Nao) * y[13]))), (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(y[23]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * y[13])))));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(calcium_dynamics__Cao, y[13])));
    const R T_type_Ca_channel__i_Ca_T = (T_type_Ca_channel__g_CaT * y[6] * y[6] * y[7] * (V - calcium_background_current__E_Ca));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R fast_sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Nao, y[23])));
    const R fast_sodium_current__i_Na = (p[3] * cmx_ipow<3>(y[1]) * y[2] * y[3] * (V - fast_sodium_current__E_Na));
    const R non_specific_calcium_activated_current__P_ns_Ca = (R(0.0) * R(1.75e-07));
    const R non_specific_calcium_activated_current__I_ns_K = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * y[24] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_K = cmx_div((non_specific_calcium_activated_current__I_ns_K * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, y[13]))));
    const R non_specific_calcium_activated_current__I_ns_Na = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * y[23] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_Na = cmx_div((non_specific_calcium_activated_current__I_ns_Na * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, y[13]))));
    const R non_specific_calcium_activated_current__i_ns_Ca = (non_specific_calcium_activated_current__i_ns_Na + non_specific_calcium_activated_current__i_ns_K);
    const R plateau_potassium_current__Kp = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(7.488) - V) * (R(1) / R(5.98))), V, -0.2412533513192246, 1.806505094678354)));
    const R plateau_potassium_current__i_Kp = (plateau_potassium_current__g_Kp * plateau_potassium_current__Kp * (V - time_independent_potassium_current__E_K));
    const R rapid_delayed_rectifier_potassium_current__Rect = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003)));
    const R rapid_delayed_rectifier_potassium_current__g_Kr = (p[2] * R(0.02614) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (rapid_delayed_rectifier_potassium_current__g_Kr * y[8] * rapid_delayed_rectifier_potassium_current__Rect * (V - time_independent_potassium_current__E_K));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__I_pCa * y[13]), (sarcolemmal_calcium_pump__K_mpCa + y[13]));
    const R slow_delayed_rectifier_potassium_current__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)), (y[24] + (slow_delayed_rectifier_potassium_current__PNaK * y[23])))));
    const R slow_delayed_rectifier_potassium_current__g_Ks = (p[0] * R(0.433) * (R(1.0) + cmx_div(R(0.6), (R(1.0) + pow(cmx_div(R(3.8e-05), y[13]), R(1.4))))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (slow_delayed_rectifier_potassium_current__g_Ks * y[9] * y[10] * (V - slow_delayed_rectifier_potassium_current__E_Ks));
    const R sodium_activated_potassium_current__g_K_Na = (R(0.0) * R(0.12848));
    const R sodium_activated_potassium_current__pona = cmx_div(R(0.85), (R(1.0) + pow(cmx_div(sodium_activated_potassium_current__kdKNa, y[23]), sodium_activated_potassium_current__nKNa)));
    const R sodium_activated_potassium_current__pov = (R(0.8) - cmx_div(R(0.65), (R(1.0) + cmx_exp_v(((V + R(125.0)) * (R(1) / R(15.0))), V, 0.09617966939259756, 12.022458674074695))));
    const R sodium_activated_potassium_current__i_K_Na = (sodium_activated_potassium_current__g_K_Na * sodium_activated_potassium_current__pona * sodium_activated_potassium_current__pov * (V - time_independent_potassium_current__E_K));
    const R sodium_background_current__i_Na_b = (sodium_background_current__g_Nab * (V - fast_sodium_current__E_Na));
    const R sodium_potassium_pump__sigma = ((R(1.0) * (R(1) / R(7.0))) * cmx_expm1((ionic_concentrations__Nao * (R(1) / R(67.3)))));
    const R sodium_potassium_pump__f_NaK = cmx_div(R(1.0), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400856348800377, -0.0)) + (R(0.0365) * sodium_potassium_pump__sigma * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400856348800377, -0.0))));
    const R sodium_potassium_pump__i_NaK = ((cmx_div((sodium_potassium_pump__I_NaK * sodium_potassium_pump__f_NaK * R(1.0)), (R(1.0) + cmx_ipow<2>(cmx_div(sodium_potassium_pump__K_mNai, y[23])))) * ionic_concentrations__Ko) * (R(1) / (ionic_concentrations__Ko + sodium_potassium_pump__K_mKo)));
    const R time_independent_potassium_current__g_K1 = (R(0.75) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R time_independent_potassium_current_K1_gate__alpha_K1 = cmx_div(R(1020.0), (R(1.0) + cmx_exp_c((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215))), (((V - time_independent_potassium_current__E_K) - R(59.215))), 0.34408276725201775)));
    const R time_independent_potassium_current_K1_gate__beta_K1 = cmx_div((R(1000.0) * ((R(0.49124) * cmx_exp_c((R(0.08032) * ((V - time_independent_potassium_current__E_K) + R(5.476))), (((V - time_independent_potassium_current__E_K) + R(5.476))), 0.11587726568420155)) + cmx_exp_c((R(0.06175) * ((V - time_independent_potassium_current__E_K) - R(594.31))), (((V - time_independent_potassium_current__E_K) - R(594.31))), 0.08908641877489348))), (R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938)));
    const R time_independent_potassium_current_K1_gate__K1_infinity = cmx_div(time_independent_potassium_current_K1_gate__alpha_K1, (time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1));
    const R time_independent_potassium_current__i_K1 = (time_independent_potassium_current__g_K1 * time_independent_potassium_current_K1_gate__K1_infinity * (V - time_independent_potassium_current__E_K));
    const R transient_outward_current__g_to = (p[1] * R(0.5));
    const R transient_outward_current__rvdv = cmx_exp_v((V * (R(1) / R(100.0))), V, 0.014426950408889635, 0.0);
    const R transient_outward_current__i_to = (transient_outward_current__g_to * cmx_ipow<3>(y[11]) * y[12] * transient_outward_current__rvdv * (V - time_independent_potassium_current__E_K));
    return ((((-R(1.0)) * (R(1) / membrane__Cm)) * (fast_sodium_current__i_Na + L_type_Ca_channel__i_Ca_L + T_type_Ca_channel__i_Ca_T + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + sodium_activated_potassium_current__i_K_Na + ATP_sensitive_potassium_current__i_K_ATP + transient_outward_current__i_to + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + Na_Ca_exchanger__i_NaCa + sarcolemmal_calcium_pump__i_p_Ca + sodium_background_current__i_Na_b + calcium_background_current__i_Ca_b + sodium_potassium_pump__i_NaK + non_specific_calcium_activated_current__i_ns_Ca + i_stim)) * (R(1) / R(1000.0)));
}

// Total ionic current in uA/cm^2 (GetIIonic): minus dV/dt at zero stimulus,
// times the membrane capacitance that converts it to a current density.
template <class R>
CMX_HD inline R iionic(const R* y, R V, const R* p, R cm)
{
    CMX_MATH_USING
    (void)p; (void)cm;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.0);
    const R membrane__Cm = R(0.001);
    const R L_type_Ca_channel__gamma_Nai = R(0.75);
    const R L_type_Ca_channel__gamma_Nao = R(0.75);
    const R L_type_Ca_channel__gamma_Ki = R(0.75);
    const R L_type_Ca_channel__gamma_Ko = R(0.75);
    const R L_type_Ca_channel__P_Ca = R(0.00054);
    const R L_type_Ca_channel__P_Na = R(6.75e-07);
    const R L_type_Ca_channel__P_K = R(1.93e-07);
    const R L_type_Ca_channel__gamma_Cai = R(1.0);
    const R L_type_Ca_channel__gamma_Cao = R(0.341);
    const R L_type_Ca_channel_f_Ca_gate__Km_Ca = R(0.0006);
    const R T_type_Ca_channel__g_CaT = R(0.05);
    const R slow_delayed_rectifier_potassium_current__PNaK = R(0.01833);
    const R plateau_potassium_current__g_Kp = R(0.00552);
    const R sodium_activated_potassium_current__nKNa = R(2.8);
    const R sodium_activated_potassium_current__kdKNa = R(66.0);
    const R ATP_sensitive_potassium_current__i_K_ATP_on = R(1.0);
    const R ATP_sensitive_potassium_current__nATP = R(0.24);
    const R ATP_sensitive_potassium_current__nicholsarea = R(5e-05);
    const R ATP_sensitive_potassium_current__ATPi = R(3.0);
    const R ATP_sensitive_potassium_current__hATP = R(2.0);
    const R ATP_sensitive_potassium_current__kATP = R(0.00025);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.0005);
    const R sarcolemmal_calcium_pump__I_pCa = R(1.15);
    const R sodium_background_current__g_Nab = R(0.004);
    const R calcium_background_current__g_Cab = R(0.003016);
    const R sodium_potassium_pump__I_NaK = R(2.25);
    const R sodium_potassium_pump__K_mNai = R(10.0);
    const R sodium_potassium_pump__K_mKo = R(1.5);
    const R non_specific_calcium_activated_current__K_m_ns_Ca = R(0.0012);
    const R Na_Ca_exchanger__c1 = R(0.00025);
    const R Na_Ca_exchanger__c2 = R(0.0001);
    const R Na_Ca_exchanger__gamma = R(0.15);
    const R calcium_dynamics__Cao = R(1.8);
    const R ionic_concentrations__Nao = R(132.0);
    const R ionic_concentrations__Ko = R(4.5);
    const R i_stim = R(0);
    (void)i_stim;
    const R ATP_sensitive_potassium_current__g_K_ATP = ((ATP_sensitive_potassium_current__i_K_ATP_on * R(0.000193)) * (R(1) / ATP_sensitive_potassium_current__nicholsarea));
    const R ATP_sensitive_potassium_current__pATP = (R(1.0) * (R(1) / (R(1.0) + pow((ATP_sensitive_potassium_current__ATPi * (R(1) / ATP_sensitive_potassium_current__kATP)), ATP_sensitive_potassium_current__hATP))));
    const R ATP_sensitive_potassium_current__GKbaraATP = (ATP_sensitive_potassium_current__g_K_ATP * ATP_sensitive_potassium_current__pATP * pow((ionic_concentrations__Ko * (R(1) / R(4.0))), ATP_sensitive_potassium_current__nATP));
    const R time_independent_potassium_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Ko, y[24])));
    const R ATP_sensitive_potassium_current__i_K_ATP = (ATP_sensitive_potassium_current__GKbaraATP * (V - time_independent_potassium_current__E_K));
    const R L_type_Ca_channel__I_CaCa = cmx_remove_singularity(V, R(-0.0), R(1.3356169352749133e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Ca * cmx_ipow<2>(R(2.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Cai * y[13] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801712697600754, 0.0)) - (L_type_Ca_channel__gamma_Cao * calcium_dynamics__Cao))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel_f_Ca_gate__f_Ca = cmx_div(R(1.0), (R(1.0) + (y[13] * (R(1) / L_type_Ca_channel_f_Ca_gate__Km_Ca))));
    const R L_type_Ca_channel__i_CaCa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaCa);
    const R L_type_Ca_channel__I_CaK = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_K * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * y[24] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaK = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaK);
    const R L_type_Ca_channel__I_CaNa = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Na * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * y[23] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaNa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaNa);
    const R L_type_Ca_channel__i_Ca_L = (L_type_Ca_channel__i_CaCa + L_type_Ca_channel__i_CaK + L_type_Ca_channel__i_CaNa);
    const R Na_Ca_exchanger__i_NaCa = cmx_div((Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(y[23]) * calcium_dynamics__Cao) - (cmx_ipow<3>(ionic_concentrations__Nao) * y[13]))), (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(y[23]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * y[13])))));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(calcium_dynamics__Cao, y[13])));
    const R T_type_Ca_channel__i_Ca_T = (T_type_Ca_channel__g_CaT * y[6] * y[6] * y[7] * (V - calcium_background_current__E_Ca));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R fast_sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Nao, y[23])));
    const R fast_sodium_current__i_Na = (p[3] * cmx_ipow<3>(y[1]) * y[2] * y[3] * (V - fast_sodium_current__E_Na));
    const R non_specific_calcium_activated_current__P_ns_Ca = (R(0.0) * R(1.75e-07));
    const R non_specific_calcium_activated_current__I_ns_K = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * y[24] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_K = cmx_div((non_specific_calcium_activated_current__I_ns_K * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, y[13]))));
    const R non_specific_calcium_activated_current__I_ns_Na = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * y[23] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_Na = cmx_div((non_specific_calcium_activated_current__I_ns_Na * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, y[13]))));
    const R non_specific_calcium_activated_current__i_ns_Ca = (non_specific_calcium_activated_current__i_ns_Na + non_specific_calcium_activated_current__i_ns_K);
    const R plateau_potassium_current__Kp = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(7.488) - V) * (R(1) / R(5.98))), V, -0.2412533513192246, 1.806505094678354)));
    const R plateau_potassium_current__i_Kp = (plateau_potassium_current__g_Kp * plateau_potassium_current__Kp * (V - time_independent_potassium_current__E_K));
    const R rapid_delayed_rectifier_potassium_current__Rect = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003)));
    const R rapid_delayed_rectifier_potassium_current__g_Kr = (p[2] * R(0.02614) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (rapid_delayed_rectifier_potassium_current__g_Kr * y[8] * rapid_delayed_rectifier_potassium_current__Rect * (V - time_independent_potassium_current__E_K));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__I_pCa * y[13]), (sarcolemmal_calcium_pump__K_mpCa + y[13]));
    const R slow_delayed_rectifier_potassium_current__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)), (y[24] + (slow_delayed_rectifier_potassium_current__PNaK * y[23])))));
    const R slow_delayed_rectifier_potassium_current__g_Ks = (p[0] * R(0.433) * (R(1.0) + cmx_div(R(0.6), (R(1.0) + pow(cmx_div(R(3.8e-05), y[13]), R(1.4))))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (slow_delayed_rectifier_potassium_current__g_Ks * y[9] * y[10] * (V - slow_delayed_rectifier_potassium_current__E_Ks));
    const R sodium_activated_potassium_current__g_K_Na = (R(0.0) * R(0.12848));
    const R sodium_activated_potassium_current__pona = cmx_div(R(0.85), (R(1.0) + pow(cmx_div(sodium_activated_potassium_current__kdKNa, y[23]), sodium_activated_potassium_current__nKNa)));
    const R sodium_activated_potassium_current__pov = (R(0.8) - cmx_div(R(0.65), (R(1.0) + cmx_exp_v(((V + R(125.0)) * (R(1) / R(15.0))), V, 0.09617966939259756, 12.022458674074695))));
    const R sodium_activated_potassium_current__i_K_Na = (sodium_activated_potassium_current__g_K_Na * sodium_activated_potassium_current__pona * sodium_activated_potassium_current__pov * (V - time_independent_potassium_current__E_K));
    const R sodium_background_current__i_Na_b = (sodium_background_current__g_Nab * (V - fast_sodium_current__E_Na));
    const R sodium_potassium_pump__sigma = ((R(1.0) * (R(1) / R(7.0))) * cmx_expm1((ionic_concentrations__Nao * (R(1) / R(67.3)))));
    const R sodium_potassium_pump__f_NaK = cmx_div(R(1.0), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400856348800377, -0.0)) + (R(0.0365) * sodium_potassium_pump__sigma * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400856348800377, -0.0))));
    const R sodium_potassium_pump__i_NaK = ((cmx_div((sodium_potassium_pump__I_NaK * sodium_potassium_pump__f_NaK * R(1.0)), (R(1.0) + cmx_ipow<2>(cmx_div(sodium_potassium_pump__K_mNai, y[23])))) * ionic_concentrations__Ko) * (R(1) / (ionic_concentrations__Ko + sodium_potassium_pump__K_mKo)));
    const R time_independent_potassium_current__g_K1 = (R(0.75) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R time_independent_potassium_current_K1_gate__alpha_K1 = cmx_div(R(1020.0), (R(1.0) + cmx_exp_c((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215))), (((V - time_independent_potassium_current__E_K) - R(59.215))), 0.34408276725201775)));
    const R time_independent_potassium_current_K1_gate__beta_K1 = cmx_div((R(1000.0) * ((R(0.49124) * cmx_exp_c((R(0.08032) * ((V - time_independent_potassium_current__E_K) + R(5.476))), (((V - time_independent_potassium_current__E_K) + R(5.476))), 0.11587726568420155)) + cmx_exp_c((R(0.06175) * ((V - time_independent_potassium_current__E_K) - R(594.31))), (((V - time_independent_potassium_current__E_K) - R(594.31))), 0.08908641877489348))), (R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938)));
    const R time_independent_potassium_current_K1_gate__K1_infinity = cmx_div(time_independent_potassium_current_K1_gate__alpha_K1, (time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1));
    const R time_independent_potassium_current__i_K1 = (time_independent_potassium_current__g_K1 * time_independent_potassium_current_K1_gate__K1_infinity * (V - time_independent_potassium_current__E_K));
    const R transient_outward_current__g_to = (p[1] * R(0.5));
    const R transient_outward_current__rvdv = cmx_exp_v((V * (R(1) / R(100.0))), V, 0.014426950408889635, 0.0);
    const R transient_outward_current__i_to = (transient_outward_current__g_to * cmx_ipow<3>(y[11]) * y[12] * transient_outward_current__rvdv * (V - time_independent_potassium_current__E_K));
    const R dv = ((((-R(1.0)) * (R(1) / membrane__Cm)) * (fast_sodium_current__i_Na + L_type_Ca_channel__i_Ca_L + T_type_Ca_channel__i_Ca_T + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + sodium_activated_potassium_current__i_K_Na + ATP_sensitive_potassium_current__i_K_ATP + transient_outward_current__i_to + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + Na_Ca_exchanger__i_NaCa + sarcolemmal_calcium_pump__i_p_Ca + sodium_background_current__i_Na_b + calcium_background_current__i_Ca_b + sodium_potassium_pump__i_NaK + non_specific_calcium_activated_current__i_ns_Ca + i_stim)) * (R(1) / R(1000.0)));
    return -dv * cm;
}

// Backward Euler: the states solved by Newton (all but the linear gates).
constexpr int kNumNonlinear = 12;
static constexpr int kNonlinear[12] = {13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24};

// Linear states: y <- (y + a dt) / (1 - b dt) for dy/dt = a + b y, a and b
// functions of V (ComputeOneStepExceptVoltage, closed-form part).
template <class R>
CMX_HD inline void be_linear(R* y, R V, const R* p, R cm, R dt)
{
    CMX_MATH_USING
    (void)p; (void)cm; (void)V;
    const R fast_sodium_current_m_gate__delta_m = R(1e-05);
    const R L_type_Ca_channel_d_gate__E0_d = (V + R(10.0));
    const R L_type_Ca_channel_d_gate__d_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_c(((-L_type_Ca_channel_d_gate__E0_d) * (R(1) / R(6.24))), (-L_type_Ca_channel_d_gate__E0_d), 0.23120112834759027)));
    const R L_type_Ca_channel_d_gate__tau_d = (((fabs(L_type_Ca_channel_d_gate__E0_d) < R(1e-05))) ? ((R(0.001) * (R(1) / (R(0.035) * R(6.24))))) : (cmx_div((R(0.001) * L_type_Ca_channel_d_gate__d_infinity * (-cmx_expm1(((-L_type_Ca_channel_d_gate__E0_d) * (R(1) / R(6.24)))))), (R(0.035) * L_type_Ca_channel_d_gate__E0_d))));
    const R L_type_Ca_channel_d_gate__alpha_d = cmx_div(L_type_Ca_channel_d_gate__d_infinity, L_type_Ca_channel_d_gate__tau_d);
    const R L_type_Ca_channel_d_gate__beta_d = cmx_div((R(1.0) - L_type_Ca_channel_d_gate__d_infinity), L_type_Ca_channel_d_gate__tau_d);
    const R L_type_Ca_channel_f_gate__f_infinity = (cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(32.0)) * (R(1) / R(8.0))), V, 0.18033688011112042, 5.7707801635558535))) + cmx_div(R(0.6), (R(1.0) + cmx_exp_v(((R(50.0) - V) * (R(1) / R(20.0))), V, -0.07213475204444818, 3.6067376022224087))));
    const R L_type_Ca_channel_f_gate__tau_f = cmx_div(R(0.001), ((R(0.0197) * cmx_exp((-cmx_ipow<2>((R(0.0337) * (V + R(10.0))))))) + R(0.02)));
    const R L_type_Ca_channel_f_gate__alpha_f = cmx_div(L_type_Ca_channel_f_gate__f_infinity, L_type_Ca_channel_f_gate__tau_f);
    const R L_type_Ca_channel_f_gate__beta_f = cmx_div((R(1.0) - L_type_Ca_channel_f_gate__f_infinity), L_type_Ca_channel_f_gate__tau_f);
    const R T_type_Ca_channel_b_gate__b_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(14.0))) * (R(1) / R(10.8))), V, -0.13358287415638548, -1.870160238189397)));
    const R T_type_Ca_channel_b_gate__tau_b = (R(0.0037) + cmx_div(R(0.0061), (R(1.0) + cmx_exp_v(((V + R(25.0)) * (R(1) / R(4.5))), V, 0.3205988979753252, 8.01497244938313))));
    const R T_type_Ca_channel_g_gate__g_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(60.0)) * (R(1) / R(5.6))), V, 0.25762411444445776, 15.457446866667466)));
    const R T_type_Ca_channel_g_gate__tau_g = (((V <= R(0.0))) ? ((((-R(0.000875)) * V) + R(0.012))) : (R(0.012)));
    const R fast_sodium_current_h_gate__alpha_h = (((V < (-R(40.0)))) ? ((R(135.0) * cmx_exp_v(((R(80.0) + V) * (R(1) / (-R(6.8)))), V, -0.2121610354248476, -16.972882833987807))) : (R(0.0)));
    const R fast_sodium_current_h_gate__beta_h = (((V < (-R(40.0)))) ? (((R(3560.0) * cmx_exp_v((R(0.079) * V), V, 0.11397290823022813, 0.0)) + (R(310000000.0) * cmx_exp_v((R(0.35) * V), V, 0.5049432643111371, 0.0)))) : (cmx_div(R(1000.0), (R(0.13) * (R(1.0) + cmx_exp_v(((V + R(10.66)) * (R(1) / (-R(11.1)))), V, -0.12997252620621294, -1.3855071293582297))))));
    const R fast_sodium_current_j_gate__alpha_j = (((V < (-R(40.0)))) ? (cmx_div((R(1000.0) * (-((R(127140.0) * cmx_exp_v((R(0.2444) * V), V, 0.3525946679932627, 0.0)) + (R(3.474e-05) * cmx_exp_v(((-R(0.04391)) * V), V, -0.06334873924543438, -0.0)))) * (V + R(37.78))), (R(1.0) + cmx_exp_v((R(0.311) * (V + R(79.23))), V, 0.44867815771646763, 35.54877043587573)))) : (R(0.0)));
    const R fast_sodium_current_j_gate__beta_j = (((V < (-R(40.0)))) ? (cmx_div((R(121.2) * cmx_exp_v(((-R(0.01052)) * V), V, -0.015177151830151895, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.1378)) * (V + R(40.14))), V, -0.19880337663449915, -7.979967538108797)))) : (cmx_div((R(300.0) * cmx_exp_v(((-R(2.535e-07)) * V), V, -3.657231928653522e-07, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(32.0))), V, -0.14426950408889636, -4.616624130844683)))));
    const R fast_sodium_current_m_gate__E0_m = (V + R(47.13));
    const R fast_sodium_current_m_gate__alpha_m = (((fabs(fast_sodium_current_m_gate__E0_m) >= fast_sodium_current_m_gate__delta_m)) ? (cmx_div((R(320.0) * fast_sodium_current_m_gate__E0_m), (-cmx_expm1(((-R(0.1)) * fast_sodium_current_m_gate__E0_m))))) : (R(3200.0)));
    const R fast_sodium_current_m_gate__beta_m = (R(80.0) * cmx_exp_v(((-V) * (R(1) / R(11.0))), V, -0.1311540946262694, -0.0));
    const R rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr = cmx_remove_singularity(V, R(-38.90000000000011), R(6.896551724137951e-07), [&](const R V) { return cmx_remove_singularity(V, R(-14.2), R(8.130081300813008e-07), [&](const R V) { return cmx_div(R(0.001), (cmx_div((R(0.00138) * (V + R(14.2))), (-cmx_expm1(((-R(0.123)) * (V + R(14.2)))))) + cmx_div((R(0.00061) * (V + R(38.9))), cmx_expm1((R(0.145) * (V + R(38.9))))))); }); });
    const R rapid_delayed_rectifier_potassium_current_xr_gate__xr_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V + R(21.5))) * (R(1) / R(7.5))), V, -0.19235933878519512, -4.135725783881695)));
    const R slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1 = cmx_remove_singularity(V, R(-29.99999999999988), R(6.75675675675673e-07), [&](const R V) { return cmx_div(R(0.001), (cmx_div((R(7.19e-05) * (V + R(30.0))), (-cmx_expm1(((-R(0.148)) * (V + R(30.0)))))) + cmx_div((R(0.000131) * (V + R(30.0))), cmx_expm1((R(0.0687) * (V + R(30.0))))))); });
    const R slow_delayed_rectifier_potassium_current_xs1_gate__xs1_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V - R(1.5))) * (R(1) / R(16.7))), V, -0.08638892460412954, 0.1295833869061943)));
    const R slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2 = cmx_remove_singularity(V, R(-29.99999999999988), R(6.75675675675673e-07), [&](const R V) { return cmx_div((R(4.0) * R(0.001)), (cmx_div((R(7.19e-05) * (V + R(30.0))), (-cmx_expm1(((-R(0.148)) * (V + R(30.0)))))) + cmx_div((R(0.000131) * (V + R(30.0))), cmx_expm1((R(0.0687) * (V + R(30.0))))))); });
    const R slow_delayed_rectifier_potassium_current_xs2_gate__xs2_infinity = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((-(V - R(1.5))) * (R(1) / R(16.7))), V, -0.08638892460412954, 0.1295833869061943)));
    const R transient_outward_current_ydv_gate__alpha_ydv = cmx_div(R(15.0), (R(1.0) + cmx_exp_v(((V + R(60.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 17.31234049066756)));
    const R transient_outward_current_ydv_gate__beta_ydv = cmx_div((R(100.0) * cmx_exp_v(((V + R(25.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 7.213475204444817)), (R(1.0) + cmx_exp_v(((V + R(25.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 7.213475204444817)));
    const R transient_outward_current_ydv_gate__tau_ydv = cmx_div(R(1.0), (transient_outward_current_ydv_gate__alpha_ydv + transient_outward_current_ydv_gate__beta_ydv));
    const R transient_outward_current_ydv_gate__ydv_ss = cmx_div(transient_outward_current_ydv_gate__alpha_ydv, (transient_outward_current_ydv_gate__alpha_ydv + transient_outward_current_ydv_gate__beta_ydv));
    const R transient_outward_current_zdv_gate__alpha_zdv = cmx_div((R(10000.0) * cmx_exp_v(((V - R(40.0)) * (R(1) / R(25.0))), V, 0.05770780163555854, -2.3083120654223417)), (R(1.0) + cmx_exp_v(((V - R(40.0)) * (R(1) / R(25.0))), V, 0.05770780163555854, -2.3083120654223417)));
    const R transient_outward_current_zdv_gate__beta_zdv = cmx_div((R(10000.0) * cmx_exp_v(((-(V + R(90.0))) * (R(1) / R(25.0))), V, -0.05770780163555854, -5.193702147200268)), (R(1.0) + cmx_exp_v(((-(V + R(90.0))) * (R(1) / R(25.0))), V, -0.05770780163555854, -5.193702147200268)));
    const R transient_outward_current_zdv_gate__tau_zdv = cmx_div(R(1.0), (transient_outward_current_zdv_gate__alpha_zdv + transient_outward_current_zdv_gate__beta_zdv));
    const R transient_outward_current_zdv_gate__zdv_ss = cmx_div(transient_outward_current_zdv_gate__alpha_zdv, (transient_outward_current_zdv_gate__alpha_zdv + transient_outward_current_zdv_gate__beta_zdv));
    y[1] = (y[1] + (((fast_sodium_current_m_gate__alpha_m * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((fast_sodium_current_m_gate__alpha_m * (-R(1))) - (fast_sodium_current_m_gate__beta_m * R(1))) / R(1000.0))) * dt);
    y[2] = (y[2] + (((fast_sodium_current_h_gate__alpha_h * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((fast_sodium_current_h_gate__alpha_h * (-R(1))) - (fast_sodium_current_h_gate__beta_h * R(1))) / R(1000.0))) * dt);
    y[3] = (y[3] + (((fast_sodium_current_j_gate__alpha_j * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((fast_sodium_current_j_gate__alpha_j * (-R(1))) - (fast_sodium_current_j_gate__beta_j * R(1))) / R(1000.0))) * dt);
    y[4] = (y[4] + (((L_type_Ca_channel_d_gate__alpha_d * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((L_type_Ca_channel_d_gate__alpha_d * (-R(1))) - (L_type_Ca_channel_d_gate__beta_d * R(1))) / R(1000.0))) * dt);
    y[5] = (y[5] + (((L_type_Ca_channel_f_gate__alpha_f * R(1.0)) / R(1000.0))) * dt) / (R(1) - ((((L_type_Ca_channel_f_gate__alpha_f * (-R(1))) - (L_type_Ca_channel_f_gate__beta_f * R(1))) / R(1000.0))) * dt);
    y[6] = (y[6] + (((T_type_Ca_channel_b_gate__b_inf / T_type_Ca_channel_b_gate__tau_b) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / T_type_Ca_channel_b_gate__tau_b) / R(1000.0))) * dt);
    y[7] = (y[7] + (((T_type_Ca_channel_g_gate__g_inf / T_type_Ca_channel_g_gate__tau_g) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / T_type_Ca_channel_g_gate__tau_g) / R(1000.0))) * dt);
    y[8] = (y[8] + (((rapid_delayed_rectifier_potassium_current_xr_gate__xr_infinity / rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / rapid_delayed_rectifier_potassium_current_xr_gate__tau_xr) / R(1000.0))) * dt);
    y[9] = (y[9] + (((slow_delayed_rectifier_potassium_current_xs1_gate__xs1_infinity / slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / slow_delayed_rectifier_potassium_current_xs1_gate__tau_xs1) / R(1000.0))) * dt);
    y[10] = (y[10] + (((slow_delayed_rectifier_potassium_current_xs2_gate__xs2_infinity / slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / slow_delayed_rectifier_potassium_current_xs2_gate__tau_xs2) / R(1000.0))) * dt);
    y[11] = (y[11] + (((transient_outward_current_zdv_gate__zdv_ss / transient_outward_current_zdv_gate__tau_zdv) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / transient_outward_current_zdv_gate__tau_zdv) / R(1000.0))) * dt);
    y[12] = (y[12] + (((transient_outward_current_ydv_gate__ydv_ss / transient_outward_current_ydv_gate__tau_ydv) / R(1000.0))) * dt) / (R(1) - ((((-R(1)) / transient_outward_current_ydv_gate__tau_ydv) / R(1000.0))) * dt);
}

// Newton residual of the nonlinear states at guess g (y holds the other
// states, linear ones already advanced): g - y_old - dt f(g).
template <class R>
CMX_HD inline void be_residual(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* res)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.0);
    const R membrane__Cm = R(0.001);
    const R L_type_Ca_channel__gamma_Nai = R(0.75);
    const R L_type_Ca_channel__gamma_Nao = R(0.75);
    const R L_type_Ca_channel__gamma_Ki = R(0.75);
    const R L_type_Ca_channel__gamma_Ko = R(0.75);
    const R L_type_Ca_channel__P_Ca = R(0.00054);
    const R L_type_Ca_channel__P_Na = R(6.75e-07);
    const R L_type_Ca_channel__P_K = R(1.93e-07);
    const R L_type_Ca_channel__gamma_Cai = R(1.0);
    const R L_type_Ca_channel__gamma_Cao = R(0.341);
    const R L_type_Ca_channel_f_Ca_gate__Km_Ca = R(0.0006);
    const R T_type_Ca_channel__g_CaT = R(0.05);
    const R slow_delayed_rectifier_potassium_current__PNaK = R(0.01833);
    const R plateau_potassium_current__g_Kp = R(0.00552);
    const R sodium_activated_potassium_current__nKNa = R(2.8);
    const R sodium_activated_potassium_current__kdKNa = R(66.0);
    const R ATP_sensitive_potassium_current__i_K_ATP_on = R(1.0);
    const R ATP_sensitive_potassium_current__nATP = R(0.24);
    const R ATP_sensitive_potassium_current__nicholsarea = R(5e-05);
    const R ATP_sensitive_potassium_current__ATPi = R(3.0);
    const R ATP_sensitive_potassium_current__hATP = R(2.0);
    const R ATP_sensitive_potassium_current__kATP = R(0.00025);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.0005);
    const R sarcolemmal_calcium_pump__I_pCa = R(1.15);
    const R sodium_background_current__g_Nab = R(0.004);
    const R calcium_background_current__g_Cab = R(0.003016);
    const R sodium_potassium_pump__I_NaK = R(2.25);
    const R sodium_potassium_pump__K_mNai = R(10.0);
    const R sodium_potassium_pump__K_mKo = R(1.5);
    const R non_specific_calcium_activated_current__K_m_ns_Ca = R(0.0012);
    const R Na_Ca_exchanger__c1 = R(0.00025);
    const R Na_Ca_exchanger__c2 = R(0.0001);
    const R Na_Ca_exchanger__gamma = R(0.15);
    const R calcium_dynamics__G_rel_max = R(60000.0);
    const R calcium_dynamics__G_rel_overload = R(4000.0);
    const R calcium_dynamics__tau_tr = R(0.18);
    const R calcium_dynamics__K_mrel = R(0.0008);
    const R calcium_dynamics__delta_Ca_ith = R(0.00018);
    const R calcium_dynamics__CSQN_max = R(10.0);
    const R calcium_dynamics__K_mCSQN = R(0.8);
    const R calcium_dynamics__K_mup = R(0.00092);
    const R calcium_dynamics__I_up = R(8.75);
    const R calcium_dynamics__Ca_NSR_max = R(15.0);
    const R calcium_dynamics__Cao = R(1.8);
    const R calcium_dynamics__K_mTn = R(0.0005);
    const R calcium_dynamics__K_mCMDN = R(0.00238);
    const R calcium_dynamics__Tn_max = R(0.07);
    const R calcium_dynamics__CMDN_max = R(0.05);
    const R calcium_dynamics__CSQNthresh = R(0.7);
    const R calcium_dynamics__Logicthresh = R(0.98);
    const R ionic_concentrations__Nao = R(132.0);
    const R ionic_concentrations__Ko = R(4.5);
    const R ionic_concentrations__A_cap = R(1.434e-07);
    const R ionic_concentrations__preplength = R(0.001);
    const R ionic_concentrations__radius = R(0.00011);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R ATP_sensitive_potassium_current__g_K_ATP = ((ATP_sensitive_potassium_current__i_K_ATP_on * R(0.000193)) * (R(1) / ATP_sensitive_potassium_current__nicholsarea));
    const R ATP_sensitive_potassium_current__pATP = (R(1.0) * (R(1) / (R(1.0) + pow((ATP_sensitive_potassium_current__ATPi * (R(1) / ATP_sensitive_potassium_current__kATP)), ATP_sensitive_potassium_current__hATP))));
    const R ATP_sensitive_potassium_current__GKbaraATP = (ATP_sensitive_potassium_current__g_K_ATP * ATP_sensitive_potassium_current__pATP * pow((ionic_concentrations__Ko * (R(1) / R(4.0))), ATP_sensitive_potassium_current__nATP));
    const R time_independent_potassium_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Ko, g[11])));
    const R ATP_sensitive_potassium_current__i_K_ATP = (ATP_sensitive_potassium_current__GKbaraATP * (V - time_independent_potassium_current__E_K));
    const R L_type_Ca_channel__I_CaCa = cmx_remove_singularity(V, R(-0.0), R(1.3356169352749133e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Ca * cmx_ipow<2>(R(2.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Cai * g[0] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801712697600754, 0.0)) - (L_type_Ca_channel__gamma_Cao * calcium_dynamics__Cao))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel_f_Ca_gate__f_Ca = cmx_div(R(1.0), (R(1.0) + (g[0] * (R(1) / L_type_Ca_channel_f_Ca_gate__Km_Ca))));
    const R L_type_Ca_channel__i_CaCa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaCa);
    const R L_type_Ca_channel__I_CaK = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_K * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * g[11] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaK = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaK);
    const R L_type_Ca_channel__I_CaNa = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Na * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * g[10] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaNa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaNa);
    const R Na_Ca_exchanger__i_NaCa = cmx_div((Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) - (cmx_ipow<3>(ionic_concentrations__Nao) * g[0]))), (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0])))));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(calcium_dynamics__Cao, g[0])));
    const R T_type_Ca_channel__i_Ca_T = (T_type_Ca_channel__g_CaT * y[6] * y[6] * y[7] * (V - calcium_background_current__E_Ca));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R ionic_concentrations__volume = (R(3.14159265358979323846) * ionic_concentrations__preplength * cmx_ipow<2>(ionic_concentrations__radius));
    const R ionic_concentrations__V_myo = (R(0.68) * ionic_concentrations__volume);
    const R calcium_dynamics__V_JSR = ((R(0.0048) * (R(1) / R(0.68))) * ionic_concentrations__V_myo);
    const R calcium_dynamics__V_NSR = ((R(0.0552) * (R(1) / R(0.68))) * ionic_concentrations__V_myo);
    const R calcium_dynamics__K_leak = (calcium_dynamics__I_up * (R(1) / calcium_dynamics__Ca_NSR_max));
    const R calcium_dynamics__i_leak = (calcium_dynamics__K_leak * g[2]);
    const R calcium_dynamics__G_rel = (((g[6] > calcium_dynamics__delta_Ca_ith)) ? ((cmx_div((calcium_dynamics__G_rel_max * (g[6] - calcium_dynamics__delta_Ca_ith)), ((calcium_dynamics__K_mrel + g[6]) - calcium_dynamics__delta_Ca_ith)) * (R(1.0) - g[4]) * g[4])) : (((((g[6] <= calcium_dynamics__delta_Ca_ith) && (g[8] > R(0.0)))) ? ((calcium_dynamics__G_rel_overload * (R(1.0) - g[8]) * g[8])) : (R(0.0)))));
    const R calcium_dynamics__i_rel = (calcium_dynamics__G_rel * (g[1] - g[0]));
    const R calcium_dynamics__i_tr = ((g[2] - g[1]) * (R(1) / calcium_dynamics__tau_tr));
    const R calcium_dynamics__i_up = cmx_div((calcium_dynamics__I_up * g[0]), (g[0] + calcium_dynamics__K_mup));
    const R fast_sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Nao, g[10])));
    const R fast_sodium_current__i_Na = (p[3] * cmx_ipow<3>(y[1]) * y[2] * y[3] * (V - fast_sodium_current__E_Na));
    const R L_type_Ca_channel__i_Ca_L = (L_type_Ca_channel__i_CaCa + L_type_Ca_channel__i_CaK + L_type_Ca_channel__i_CaNa);
    const R non_specific_calcium_activated_current__P_ns_Ca = (R(0.0) * R(1.75e-07));
    const R non_specific_calcium_activated_current__I_ns_K = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * g[11] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_K = cmx_div((non_specific_calcium_activated_current__I_ns_K * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))));
    const R non_specific_calcium_activated_current__I_ns_Na = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * g[10] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_Na = cmx_div((non_specific_calcium_activated_current__I_ns_Na * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))));
    const R non_specific_calcium_activated_current__i_ns_Ca = (non_specific_calcium_activated_current__i_ns_Na + non_specific_calcium_activated_current__i_ns_K);
    const R plateau_potassium_current__Kp = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(7.488) - V) * (R(1) / R(5.98))), V, -0.2412533513192246, 1.806505094678354)));
    const R plateau_potassium_current__i_Kp = (plateau_potassium_current__g_Kp * plateau_potassium_current__Kp * (V - time_independent_potassium_current__E_K));
    const R rapid_delayed_rectifier_potassium_current__Rect = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003)));
    const R rapid_delayed_rectifier_potassium_current__g_Kr = (p[2] * R(0.02614) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (rapid_delayed_rectifier_potassium_current__g_Kr * y[8] * rapid_delayed_rectifier_potassium_current__Rect * (V - time_independent_potassium_current__E_K));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__I_pCa * g[0]), (sarcolemmal_calcium_pump__K_mpCa + g[0]));
    const R slow_delayed_rectifier_potassium_current__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)), (g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])))));
    const R slow_delayed_rectifier_potassium_current__g_Ks = (p[0] * R(0.433) * (R(1.0) + cmx_div(R(0.6), (R(1.0) + pow(cmx_div(R(3.8e-05), g[0]), R(1.4))))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (slow_delayed_rectifier_potassium_current__g_Ks * y[9] * y[10] * (V - slow_delayed_rectifier_potassium_current__E_Ks));
    const R sodium_activated_potassium_current__g_K_Na = (R(0.0) * R(0.12848));
    const R sodium_activated_potassium_current__pona = cmx_div(R(0.85), (R(1.0) + pow(cmx_div(sodium_activated_potassium_current__kdKNa, g[10]), sodium_activated_potassium_current__nKNa)));
    const R sodium_activated_potassium_current__pov = (R(0.8) - cmx_div(R(0.65), (R(1.0) + cmx_exp_v(((V + R(125.0)) * (R(1) / R(15.0))), V, 0.09617966939259756, 12.022458674074695))));
    const R sodium_activated_potassium_current__i_K_Na = (sodium_activated_potassium_current__g_K_Na * sodium_activated_potassium_current__pona * sodium_activated_potassium_current__pov * (V - time_independent_potassium_current__E_K));
    const R sodium_background_current__i_Na_b = (sodium_background_current__g_Nab * (V - fast_sodium_current__E_Na));
    const R sodium_potassium_pump__sigma = ((R(1.0) * (R(1) / R(7.0))) * cmx_expm1((ionic_concentrations__Nao * (R(1) / R(67.3)))));
    const R sodium_potassium_pump__f_NaK = cmx_div(R(1.0), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400856348800377, -0.0)) + (R(0.0365) * sodium_potassium_pump__sigma * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400856348800377, -0.0))));
    const R sodium_potassium_pump__i_NaK = ((cmx_div((sodium_potassium_pump__I_NaK * sodium_potassium_pump__f_NaK * R(1.0)), (R(1.0) + cmx_ipow<2>(cmx_div(sodium_potassium_pump__K_mNai, g[10])))) * ionic_concentrations__Ko) * (R(1) / (ionic_concentrations__Ko + sodium_potassium_pump__K_mKo)));
    const R time_independent_potassium_current__g_K1 = (R(0.75) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R time_independent_potassium_current_K1_gate__alpha_K1 = cmx_div(R(1020.0), (R(1.0) + cmx_exp_c((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215))), (((V - time_independent_potassium_current__E_K) - R(59.215))), 0.34408276725201775)));
    const R time_independent_potassium_current_K1_gate__beta_K1 = cmx_div((R(1000.0) * ((R(0.49124) * cmx_exp_c((R(0.08032) * ((V - time_independent_potassium_current__E_K) + R(5.476))), (((V - time_independent_potassium_current__E_K) + R(5.476))), 0.11587726568420155)) + cmx_exp_c((R(0.06175) * ((V - time_independent_potassium_current__E_K) - R(594.31))), (((V - time_independent_potassium_current__E_K) - R(594.31))), 0.08908641877489348))), (R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938)));
    const R time_independent_potassium_current_K1_gate__K1_infinity = cmx_div(time_independent_potassium_current_K1_gate__alpha_K1, (time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1));
    const R time_independent_potassium_current__i_K1 = (time_independent_potassium_current__g_K1 * time_independent_potassium_current_K1_gate__K1_infinity * (V - time_independent_potassium_current__E_K));
    const R transient_outward_current__g_to = (p[1] * R(0.5));
    const R transient_outward_current__rvdv = cmx_exp_v((V * (R(1) / R(100.0))), V, 0.014426950408889635, 0.0);
    const R transient_outward_current__i_to = (transient_outward_current__g_to * cmx_ipow<3>(y[11]) * y[12] * transient_outward_current__rvdv * (V - time_independent_potassium_current__E_K));
    const R membrane__dVdt = (((-R(1.0)) * (R(1) / membrane__Cm)) * (fast_sodium_current__i_Na + L_type_Ca_channel__i_Ca_L + T_type_Ca_channel__i_Ca_T + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + sodium_activated_potassium_current__i_K_Na + ATP_sensitive_potassium_current__i_K_ATP + transient_outward_current__i_to + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + Na_Ca_exchanger__i_NaCa + sarcolemmal_calcium_pump__i_p_Ca + sodium_background_current__i_Na_b + calcium_background_current__i_Ca_b + sodium_potassium_pump__i_NaK + non_specific_calcium_activated_current__i_ns_Ca + i_stim));
    res[0] = g[0] - y[13] - dt * ((cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))) * ((((-ionic_concentrations__A_cap) * (((L_type_Ca_channel__i_CaCa + T_type_Ca_channel__i_Ca_T) - (R(2.0) * Na_Ca_exchanger__i_NaCa)) + sarcolemmal_calcium_pump__i_p_Ca + calcium_background_current__i_Ca_b)) * (R(1) / (R(2.0) * ionic_concentrations__V_myo * membrane__F))) + ((calcium_dynamics__i_rel * calcium_dynamics__V_JSR) * (R(1) / ionic_concentrations__V_myo)) + (((calcium_dynamics__i_leak - calcium_dynamics__i_up) * calcium_dynamics__V_NSR) * (R(1) / ionic_concentrations__V_myo)))) * (R(1) / R(1000.0)));
    res[1] = g[1] - y[14] - dt * ((cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))) * (calcium_dynamics__i_tr - calcium_dynamics__i_rel)) * (R(1) / R(1000.0)));
    res[2] = g[2] - y[15] - dt * ((((((-calcium_dynamics__i_tr) * calcium_dynamics__V_JSR) * (R(1) / calcium_dynamics__V_NSR)) - calcium_dynamics__i_leak) + calcium_dynamics__i_up) * (R(1) / R(1000.0)));
    res[3] = g[3] - y[16] - dt * ((((membrane__dVdt > R(150000.0))) ? (((R(100000.0) * (R(1.0) - g[3])) - (R(500.0) * g[3]))) : (((-R(500.0)) * g[3]))) * (R(1) / R(1000.0)));
    res[4] = g[4] - y[17] - dt * (((((g[3] < R(0.2)) && (g[3] > R(0.18)))) ? (((R(100000.0) * (R(1.0) - g[4])) - (R(500.0) * g[4]))) : (((-R(500.0)) * g[4]))) * (R(1) / R(1000.0)));
    res[5] = g[5] - y[18] - dt * (((((g[3] < R(0.2)) && (g[3] > R(0.18)))) ? (((R(100000.0) * (R(1.0) - g[5])) - (R(500.0) * g[5]))) : (((-R(10.0)) * g[5]))) * (R(1) / R(1000.0)));
    res[6] = g[6] - y[19] - dt * ((((g[3] > R(0.2))) ? ((((-ionic_concentrations__A_cap) * (((L_type_Ca_channel__i_CaCa + T_type_Ca_channel__i_Ca_T) - Na_Ca_exchanger__i_NaCa) + sarcolemmal_calcium_pump__i_p_Ca + calcium_background_current__i_Ca_b)) * (R(1) / (R(2.0) * ionic_concentrations__V_myo * membrane__F)))) : (((((g[4] > R(0.01)) && (g[3] <= R(0.2)))) ? (R(0.0)) : (((-R(500.0)) * g[6]))))) * (R(1) / R(1000.0)));
    res[7] = g[7] - y[20] - dt * (((((cmx_div(R(1.0), (R(1.0) + cmx_div(calcium_dynamics__K_mCSQN, g[1]))) > calcium_dynamics__CSQNthresh) && (g[9] < R(0.37)) && (g[5] < R(0.37)))) ? ((R(50000.0) * (R(1.0) - g[7]))) : (((-R(500.0)) * g[7]))) * (R(1) / R(1000.0)));
    res[8] = g[8] - y[21] - dt * (((((g[7] > calcium_dynamics__Logicthresh) && (g[8] < calcium_dynamics__Logicthresh))) ? ((R(50000.0) * (R(1.0) - g[8]))) : (((-R(500.0)) * g[8]))) * (R(1) / R(1000.0)));
    res[9] = g[9] - y[22] - dt * (((((g[7] > calcium_dynamics__Logicthresh) && (g[9] < calcium_dynamics__Logicthresh))) ? ((R(50000.0) * (R(1.0) - g[9]))) : (((-R(10.0)) * g[9]))) * (R(1) / R(1000.0)));
    res[10] = g[10] - y[23] - dt * ((((-(fast_sodium_current__i_Na + L_type_Ca_channel__i_CaNa + sodium_background_current__i_Na_b + non_specific_calcium_activated_current__i_ns_Na + (Na_Ca_exchanger__i_NaCa * R(3.0)) + (sodium_potassium_pump__i_NaK * R(3.0)))) * ionic_concentrations__A_cap) * (R(1) / (ionic_concentrations__V_myo * membrane__F))) * (R(1) / R(1000.0)));
    res[11] = g[11] - y[24] - dt * ((((-(L_type_Ca_channel__i_CaK + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + sodium_activated_potassium_current__i_K_Na + ATP_sensitive_potassium_current__i_K_ATP + transient_outward_current__i_to + non_specific_calcium_activated_current__i_ns_K + ((-sodium_potassium_pump__i_NaK) * R(2.0)))) * ionic_concentrations__A_cap) * (R(1) / (ionic_concentrations__V_myo * membrane__F))) * (R(1) / R(1000.0)));
}

// Jacobian of be_residual in g: J[i*N + j] = delta_ij - dt d f_i / d g_j.
template <class R>
CMX_HD inline void be_jacobian(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* J)
{
    CMX_MATH_USING
    (void)y; (void)p; (void)cm; (void)i_stim_area;
    const R membrane__R = R(8314.0);
    const R membrane__T = R(310.0);
    const R membrane__F = R(96485.0);
    const R membrane__Cm = R(0.001);
    const R L_type_Ca_channel__gamma_Nai = R(0.75);
    const R L_type_Ca_channel__gamma_Nao = R(0.75);
    const R L_type_Ca_channel__gamma_Ki = R(0.75);
    const R L_type_Ca_channel__gamma_Ko = R(0.75);
    const R L_type_Ca_channel__P_Ca = R(0.00054);
    const R L_type_Ca_channel__P_Na = R(6.75e-07);
    const R L_type_Ca_channel__P_K = R(1.93e-07);
    const R L_type_Ca_channel__gamma_Cai = R(1.0);
    const R L_type_Ca_channel__gamma_Cao = R(0.341);
    const R L_type_Ca_channel_f_Ca_gate__Km_Ca = R(0.0006);
    const R T_type_Ca_channel__g_CaT = R(0.05);
    const R slow_delayed_rectifier_potassium_current__PNaK = R(0.01833);
    const R plateau_potassium_current__g_Kp = R(0.00552);
    const R sodium_activated_potassium_current__nKNa = R(2.8);
    const R sodium_activated_potassium_current__kdKNa = R(66.0);
    const R ATP_sensitive_potassium_current__i_K_ATP_on = R(1.0);
    const R ATP_sensitive_potassium_current__nATP = R(0.24);
    const R ATP_sensitive_potassium_current__nicholsarea = R(5e-05);
    const R ATP_sensitive_potassium_current__ATPi = R(3.0);
    const R ATP_sensitive_potassium_current__hATP = R(2.0);
    const R ATP_sensitive_potassium_current__kATP = R(0.00025);
    const R sarcolemmal_calcium_pump__K_mpCa = R(0.0005);
    const R sarcolemmal_calcium_pump__I_pCa = R(1.15);
    const R sodium_background_current__g_Nab = R(0.004);
    const R calcium_background_current__g_Cab = R(0.003016);
    const R sodium_potassium_pump__I_NaK = R(2.25);
    const R sodium_potassium_pump__K_mNai = R(10.0);
    const R sodium_potassium_pump__K_mKo = R(1.5);
    const R non_specific_calcium_activated_current__K_m_ns_Ca = R(0.0012);
    const R Na_Ca_exchanger__c1 = R(0.00025);
    const R Na_Ca_exchanger__c2 = R(0.0001);
    const R Na_Ca_exchanger__gamma = R(0.15);
    const R calcium_dynamics__G_rel_max = R(60000.0);
    const R calcium_dynamics__G_rel_overload = R(4000.0);
    const R calcium_dynamics__tau_tr = R(0.18);
    const R calcium_dynamics__K_mrel = R(0.0008);
    const R calcium_dynamics__delta_Ca_ith = R(0.00018);
    const R calcium_dynamics__CSQN_max = R(10.0);
    const R calcium_dynamics__K_mCSQN = R(0.8);
    const R calcium_dynamics__K_mup = R(0.00092);
    const R calcium_dynamics__I_up = R(8.75);
    const R calcium_dynamics__Ca_NSR_max = R(15.0);
    const R calcium_dynamics__Cao = R(1.8);
    const R calcium_dynamics__K_mTn = R(0.0005);
    const R calcium_dynamics__K_mCMDN = R(0.00238);
    const R calcium_dynamics__Tn_max = R(0.07);
    const R calcium_dynamics__CMDN_max = R(0.05);
    const R calcium_dynamics__CSQNthresh = R(0.7);
    const R calcium_dynamics__Logicthresh = R(0.98);
    const R ionic_concentrations__Nao = R(132.0);
    const R ionic_concentrations__Ko = R(4.5);
    const R ionic_concentrations__A_cap = R(1.434e-07);
    const R ionic_concentrations__preplength = R(0.001);
    const R ionic_concentrations__radius = R(0.00011);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R ATP_sensitive_potassium_current__g_K_ATP = ((ATP_sensitive_potassium_current__i_K_ATP_on * R(0.000193)) * (R(1) / ATP_sensitive_potassium_current__nicholsarea));
    const R ATP_sensitive_potassium_current__pATP = (R(1.0) * (R(1) / (R(1.0) + pow((ATP_sensitive_potassium_current__ATPi * (R(1) / ATP_sensitive_potassium_current__kATP)), ATP_sensitive_potassium_current__hATP))));
    const R ATP_sensitive_potassium_current__GKbaraATP = (ATP_sensitive_potassium_current__g_K_ATP * ATP_sensitive_potassium_current__pATP * pow((ionic_concentrations__Ko * (R(1) / R(4.0))), ATP_sensitive_potassium_current__nATP));
    const R time_independent_potassium_current__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Ko, g[11])));
    const R ATP_sensitive_potassium_current__i_K_ATP = (ATP_sensitive_potassium_current__GKbaraATP * (V - time_independent_potassium_current__E_K));
    const R L_type_Ca_channel__I_CaCa = cmx_remove_singularity(V, R(-0.0), R(1.3356169352749133e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Ca * cmx_ipow<2>(R(2.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Cai * g[0] * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801712697600754, 0.0)) - (L_type_Ca_channel__gamma_Cao * calcium_dynamics__Cao))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel_f_Ca_gate__f_Ca = cmx_div(R(1.0), (R(1.0) + (g[0] * (R(1) / L_type_Ca_channel_f_Ca_gate__Km_Ca))));
    const R L_type_Ca_channel__i_CaCa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaCa);
    const R L_type_Ca_channel__I_CaK = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_K * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * g[11] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaK = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaK);
    const R L_type_Ca_channel__I_CaNa = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((L_type_Ca_channel__P_Na * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * g[10] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_channel__i_CaNa = (y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca * L_type_Ca_channel__I_CaNa);
    const R Na_Ca_exchanger__i_NaCa = cmx_div((Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) - (cmx_ipow<3>(ionic_concentrations__Nao) * g[0]))), (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0])))));
    const R calcium_background_current__E_Ca = (((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F))) * cmx_log(cmx_div(calcium_dynamics__Cao, g[0])));
    const R T_type_Ca_channel__i_Ca_T = (T_type_Ca_channel__g_CaT * y[6] * y[6] * y[7] * (V - calcium_background_current__E_Ca));
    const R calcium_background_current__i_Ca_b = (calcium_background_current__g_Cab * (V - calcium_background_current__E_Ca));
    const R ionic_concentrations__volume = (R(3.14159265358979323846) * ionic_concentrations__preplength * cmx_ipow<2>(ionic_concentrations__radius));
    const R ionic_concentrations__V_myo = (R(0.68) * ionic_concentrations__volume);
    const R calcium_dynamics__V_JSR = ((R(0.0048) * (R(1) / R(0.68))) * ionic_concentrations__V_myo);
    const R calcium_dynamics__V_NSR = ((R(0.0552) * (R(1) / R(0.68))) * ionic_concentrations__V_myo);
    const R calcium_dynamics__K_leak = (calcium_dynamics__I_up * (R(1) / calcium_dynamics__Ca_NSR_max));
    const R calcium_dynamics__i_leak = (calcium_dynamics__K_leak * g[2]);
    const R calcium_dynamics__G_rel = (((g[6] > calcium_dynamics__delta_Ca_ith)) ? ((cmx_div((calcium_dynamics__G_rel_max * (g[6] - calcium_dynamics__delta_Ca_ith)), ((calcium_dynamics__K_mrel + g[6]) - calcium_dynamics__delta_Ca_ith)) * (R(1.0) - g[4]) * g[4])) : (((((g[6] <= calcium_dynamics__delta_Ca_ith) && (g[8] > R(0.0)))) ? ((calcium_dynamics__G_rel_overload * (R(1.0) - g[8]) * g[8])) : (R(0.0)))));
    const R calcium_dynamics__i_rel = (calcium_dynamics__G_rel * (g[1] - g[0]));
    const R calcium_dynamics__i_tr = ((g[2] - g[1]) * (R(1) / calcium_dynamics__tau_tr));
    const R calcium_dynamics__i_up = cmx_div((calcium_dynamics__I_up * g[0]), (g[0] + calcium_dynamics__K_mup));
    const R fast_sodium_current__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(ionic_concentrations__Nao, g[10])));
    const R fast_sodium_current__i_Na = (p[3] * cmx_ipow<3>(y[1]) * y[2] * y[3] * (V - fast_sodium_current__E_Na));
    const R L_type_Ca_channel__i_Ca_L = (L_type_Ca_channel__i_CaCa + L_type_Ca_channel__i_CaK + L_type_Ca_channel__i_CaNa);
    const R non_specific_calcium_activated_current__P_ns_Ca = (R(0.0) * R(1.75e-07));
    const R non_specific_calcium_activated_current__I_ns_K = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Ki * g[11] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Ko * ionic_concentrations__Ko))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_K = cmx_div((non_specific_calcium_activated_current__I_ns_K * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))));
    const R non_specific_calcium_activated_current__I_ns_Na = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div((((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((L_type_Ca_channel__gamma_Nai * g[10] * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0)) - (L_type_Ca_channel__gamma_Nao * ionic_concentrations__Nao))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R non_specific_calcium_activated_current__i_ns_Na = cmx_div((non_specific_calcium_activated_current__I_ns_Na * R(1.0)), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))));
    const R non_specific_calcium_activated_current__i_ns_Ca = (non_specific_calcium_activated_current__i_ns_Na + non_specific_calcium_activated_current__i_ns_K);
    const R plateau_potassium_current__Kp = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(7.488) - V) * (R(1) / R(5.98))), V, -0.2412533513192246, 1.806505094678354)));
    const R plateau_potassium_current__i_Kp = (plateau_potassium_current__g_Kp * plateau_potassium_current__Kp * (V - time_independent_potassium_current__E_K));
    const R rapid_delayed_rectifier_potassium_current__Rect = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(9.0)) * (R(1) / R(22.4))), V, 0.06440602861111444, 0.57965425750003)));
    const R rapid_delayed_rectifier_potassium_current__g_Kr = (p[2] * R(0.02614) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R rapid_delayed_rectifier_potassium_current__i_Kr = (rapid_delayed_rectifier_potassium_current__g_Kr * y[8] * rapid_delayed_rectifier_potassium_current__Rect * (V - time_independent_potassium_current__E_K));
    const R sarcolemmal_calcium_pump__i_p_Ca = cmx_div((sarcolemmal_calcium_pump__I_pCa * g[0]), (sarcolemmal_calcium_pump__K_mpCa + g[0]));
    const R slow_delayed_rectifier_potassium_current__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)), (g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])))));
    const R slow_delayed_rectifier_potassium_current__g_Ks = (p[0] * R(0.433) * (R(1.0) + cmx_div(R(0.6), (R(1.0) + pow(cmx_div(R(3.8e-05), g[0]), R(1.4))))));
    const R slow_delayed_rectifier_potassium_current__i_Ks = (slow_delayed_rectifier_potassium_current__g_Ks * y[9] * y[10] * (V - slow_delayed_rectifier_potassium_current__E_Ks));
    const R sodium_activated_potassium_current__g_K_Na = (R(0.0) * R(0.12848));
    const R sodium_activated_potassium_current__pona = cmx_div(R(0.85), (R(1.0) + pow(cmx_div(sodium_activated_potassium_current__kdKNa, g[10]), sodium_activated_potassium_current__nKNa)));
    const R sodium_activated_potassium_current__pov = (R(0.8) - cmx_div(R(0.65), (R(1.0) + cmx_exp_v(((V + R(125.0)) * (R(1) / R(15.0))), V, 0.09617966939259756, 12.022458674074695))));
    const R sodium_activated_potassium_current__i_K_Na = (sodium_activated_potassium_current__g_K_Na * sodium_activated_potassium_current__pona * sodium_activated_potassium_current__pov * (V - time_independent_potassium_current__E_K));
    const R sodium_background_current__i_Na_b = (sodium_background_current__g_Nab * (V - fast_sodium_current__E_Na));
    const R sodium_potassium_pump__sigma = ((R(1.0) * (R(1) / R(7.0))) * cmx_expm1((ionic_concentrations__Nao * (R(1) / R(67.3)))));
    const R sodium_potassium_pump__f_NaK = cmx_div(R(1.0), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400856348800377, -0.0)) + (R(0.0365) * sodium_potassium_pump__sigma * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400856348800377, -0.0))));
    const R sodium_potassium_pump__i_NaK = ((cmx_div((sodium_potassium_pump__I_NaK * sodium_potassium_pump__f_NaK * R(1.0)), (R(1.0) + cmx_ipow<2>(cmx_div(sodium_potassium_pump__K_mNai, g[10])))) * ionic_concentrations__Ko) * (R(1) / (ionic_concentrations__Ko + sodium_potassium_pump__K_mKo)));
    const R time_independent_potassium_current__g_K1 = (R(0.75) * sqrt((ionic_concentrations__Ko * (R(1) / R(5.4)))));
    const R time_independent_potassium_current_K1_gate__alpha_K1 = cmx_div(R(1020.0), (R(1.0) + cmx_exp_c((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215))), (((V - time_independent_potassium_current__E_K) - R(59.215))), 0.34408276725201775)));
    const R time_independent_potassium_current_K1_gate__beta_K1 = cmx_div((R(1000.0) * ((R(0.49124) * cmx_exp_c((R(0.08032) * ((V - time_independent_potassium_current__E_K) + R(5.476))), (((V - time_independent_potassium_current__E_K) + R(5.476))), 0.11587726568420155)) + cmx_exp_c((R(0.06175) * ((V - time_independent_potassium_current__E_K) - R(594.31))), (((V - time_independent_potassium_current__E_K) - R(594.31))), 0.08908641877489348))), (R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938)));
    const R time_independent_potassium_current_K1_gate__K1_infinity = cmx_div(time_independent_potassium_current_K1_gate__alpha_K1, (time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1));
    const R time_independent_potassium_current__i_K1 = (time_independent_potassium_current__g_K1 * time_independent_potassium_current_K1_gate__K1_infinity * (V - time_independent_potassium_current__E_K));
    const R transient_outward_current__g_to = (p[1] * R(0.5));
    const R transient_outward_current__rvdv = cmx_exp_v((V * (R(1) / R(100.0))), V, 0.014426950408889635, 0.0);
    const R transient_outward_current__i_to = (transient_outward_current__g_to * cmx_ipow<3>(y[11]) * y[12] * transient_outward_current__rvdv * (V - time_independent_potassium_current__E_K));
    const R membrane__dVdt = (((-R(1.0)) * (R(1) / membrane__Cm)) * (fast_sodium_current__i_Na + L_type_Ca_channel__i_Ca_L + T_type_Ca_channel__i_Ca_T + rapid_delayed_rectifier_potassium_current__i_Kr + slow_delayed_rectifier_potassium_current__i_Ks + sodium_activated_potassium_current__i_K_Na + ATP_sensitive_potassium_current__i_K_ATP + transient_outward_current__i_to + time_independent_potassium_current__i_K1 + plateau_potassium_current__i_Kp + Na_Ca_exchanger__i_NaCa + sarcolemmal_calcium_pump__i_p_Ca + sodium_background_current__i_Na_b + calcium_background_current__i_Ca_b + sodium_potassium_pump__i_NaK + non_specific_calcium_activated_current__i_ns_Ca + i_stim));
    const R d_time_independent_potassium_current__E_K__d11 = ((cmx_div(cmx_div((-(ionic_concentrations__Ko * R(1))), (g[11] * g[11])), cmx_div(ionic_concentrations__Ko, g[11])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_ATP_sensitive_potassium_current__i_K_ATP__d11 = (((-d_time_independent_potassium_current__E_K__d11) * ATP_sensitive_potassium_current__GKbaraATP));
    const R d_L_type_Ca_channel__I_CaCa__d0 = cmx_remove_singularity(V, R(-0.0), R(1.3356169352749133e-06), [&](const R V) { return cmx_div(((((R(1) * L_type_Ca_channel__gamma_Cai * cmx_exp_v(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801712697600754, 0.0))) * ((L_type_Ca_channel__P_Ca * cmx_ipow<2>(R(2.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((R(2.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_L_type_Ca_channel_f_Ca_gate__f_Ca__d0 = cmx_div((-(R(1.0) * (cmx_div(R(1), L_type_Ca_channel_f_Ca_gate__Km_Ca)))), ((R(1.0) + (g[0] * (R(1) / L_type_Ca_channel_f_Ca_gate__Km_Ca))) * (R(1.0) + (g[0] * (R(1) / L_type_Ca_channel_f_Ca_gate__Km_Ca)))));
    const R d_L_type_Ca_channel__i_CaCa__d0 = ((d_L_type_Ca_channel_f_Ca_gate__f_Ca__d0 * y[4] * y[5] * L_type_Ca_channel__I_CaCa) + (d_L_type_Ca_channel__I_CaCa__d0 * y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca));
    const R d_L_type_Ca_channel__I_CaK__d11 = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div(((((R(1) * L_type_Ca_channel__gamma_Ki * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0))) * ((L_type_Ca_channel__P_K * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_L_type_Ca_channel__i_CaK__d0 = ((d_L_type_Ca_channel_f_Ca_gate__f_Ca__d0 * y[4] * y[5] * L_type_Ca_channel__I_CaK));
    const R d_L_type_Ca_channel__i_CaK__d11 = ((d_L_type_Ca_channel__I_CaK__d11 * y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca));
    const R d_L_type_Ca_channel__I_CaNa__d10 = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div(((((R(1) * L_type_Ca_channel__gamma_Nai * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0))) * ((L_type_Ca_channel__P_Na * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_L_type_Ca_channel__i_CaNa__d0 = ((d_L_type_Ca_channel_f_Ca_gate__f_Ca__d0 * y[4] * y[5] * L_type_Ca_channel__I_CaNa));
    const R d_L_type_Ca_channel__i_CaNa__d10 = ((d_L_type_Ca_channel__I_CaNa__d10 * y[4] * y[5] * L_type_Ca_channel_f_Ca_gate__f_Ca));
    const R d_Na_Ca_exchanger__i_NaCa__d0 = cmx_div(((((-((R(1) * cmx_ipow<3>(ionic_concentrations__Nao)))) * Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0))) * (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0])))) - (Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) - (cmx_ipow<3>(ionic_concentrations__Nao) * g[0]))) * ((((((R(1) * cmx_ipow<3>(ionic_concentrations__Nao)))) * Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0))))), ((R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0])))) * (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0]))))));
    const R d_Na_Ca_exchanger__i_NaCa__d10 = cmx_div(((((((R(3.0) * cmx_ipow<2>(g[10]) * R(1)) * cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * calcium_dynamics__Cao)) * Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0))) * (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0])))) - (Na_Ca_exchanger__c1 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) - (cmx_ipow<3>(ionic_concentrations__Nao) * g[0]))) * (((((((R(3.0) * cmx_ipow<2>(g[10]) * R(1)) * cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * calcium_dynamics__Cao))) * Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0))))), ((R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0])))) * (R(1.0) + (Na_Ca_exchanger__c2 * cmx_exp_v((((Na_Ca_exchanger__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.0459072789648032, -0.0) * ((cmx_exp_v(((V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0) * cmx_ipow<3>(g[10]) * calcium_dynamics__Cao) + (cmx_ipow<3>(ionic_concentrations__Nao) * g[0]))))));
    const R d_calcium_background_current__E_Ca__d0 = ((cmx_div(cmx_div((-(calcium_dynamics__Cao * R(1))), (g[0] * g[0])), cmx_div(calcium_dynamics__Cao, g[0])) * ((membrane__R * membrane__T) * (R(1) / (R(2.0) * membrane__F)))));
    const R d_T_type_Ca_channel__i_Ca_T__d0 = (((-d_calcium_background_current__E_Ca__d0) * T_type_Ca_channel__g_CaT * y[6] * y[6] * y[7]));
    const R d_calcium_background_current__i_Ca_b__d0 = (((-d_calcium_background_current__E_Ca__d0) * calcium_background_current__g_Cab));
    const R d_calcium_dynamics__i_leak__d2 = ((R(1) * calcium_dynamics__K_leak));
    const R d_calcium_dynamics__G_rel__d4 = (((g[6] > calcium_dynamics__delta_Ca_ith)) ? ((((-R(1)) * cmx_div((calcium_dynamics__G_rel_max * (g[6] - calcium_dynamics__delta_Ca_ith)), ((calcium_dynamics__K_mrel + g[6]) - calcium_dynamics__delta_Ca_ith)) * g[4]) + (R(1) * cmx_div((calcium_dynamics__G_rel_max * (g[6] - calcium_dynamics__delta_Ca_ith)), ((calcium_dynamics__K_mrel + g[6]) - calcium_dynamics__delta_Ca_ith)) * (R(1.0) - g[4])))) : (((((g[6] <= calcium_dynamics__delta_Ca_ith) && (g[8] > R(0.0)))) ? (R(0)) : (R(0)))));
    const R d_calcium_dynamics__G_rel__d6 = (((g[6] > calcium_dynamics__delta_Ca_ith)) ? (((cmx_div((((R(1) * calcium_dynamics__G_rel_max)) * ((calcium_dynamics__K_mrel + g[6]) - calcium_dynamics__delta_Ca_ith) - (calcium_dynamics__G_rel_max * (g[6] - calcium_dynamics__delta_Ca_ith)) * (R(1))), (((calcium_dynamics__K_mrel + g[6]) - calcium_dynamics__delta_Ca_ith) * ((calcium_dynamics__K_mrel + g[6]) - calcium_dynamics__delta_Ca_ith))) * (R(1.0) - g[4]) * g[4]))) : (((((g[6] <= calcium_dynamics__delta_Ca_ith) && (g[8] > R(0.0)))) ? (R(0)) : (R(0)))));
    const R d_calcium_dynamics__G_rel__d8 = (((g[6] > calcium_dynamics__delta_Ca_ith)) ? (R(0)) : (((((g[6] <= calcium_dynamics__delta_Ca_ith) && (g[8] > R(0.0)))) ? ((((-R(1)) * calcium_dynamics__G_rel_overload * g[8]) + (R(1) * calcium_dynamics__G_rel_overload * (R(1.0) - g[8])))) : (R(0)))));
    const R d_calcium_dynamics__i_rel__d0 = (((-R(1)) * calcium_dynamics__G_rel));
    const R d_calcium_dynamics__i_rel__d1 = ((R(1) * calcium_dynamics__G_rel));
    const R d_calcium_dynamics__i_rel__d4 = ((d_calcium_dynamics__G_rel__d4 * (g[1] - g[0])));
    const R d_calcium_dynamics__i_rel__d6 = ((d_calcium_dynamics__G_rel__d6 * (g[1] - g[0])));
    const R d_calcium_dynamics__i_rel__d8 = ((d_calcium_dynamics__G_rel__d8 * (g[1] - g[0])));
    const R d_calcium_dynamics__i_tr__d1 = cmx_div((-R(1)), calcium_dynamics__tau_tr);
    const R d_calcium_dynamics__i_tr__d2 = cmx_div(R(1), calcium_dynamics__tau_tr);
    const R d_calcium_dynamics__i_up__d0 = cmx_div((((R(1) * calcium_dynamics__I_up)) * (g[0] + calcium_dynamics__K_mup) - (calcium_dynamics__I_up * g[0]) * (R(1))), ((g[0] + calcium_dynamics__K_mup) * (g[0] + calcium_dynamics__K_mup)));
    const R d_fast_sodium_current__E_Na__d10 = ((cmx_div(cmx_div((-(ionic_concentrations__Nao * R(1))), (g[10] * g[10])), cmx_div(ionic_concentrations__Nao, g[10])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_fast_sodium_current__i_Na__d10 = (((-d_fast_sodium_current__E_Na__d10) * p[3] * cmx_ipow<3>(y[1]) * y[2] * y[3]));
    const R d_L_type_Ca_channel__i_Ca_L__d0 = (d_L_type_Ca_channel__i_CaCa__d0 + d_L_type_Ca_channel__i_CaK__d0 + d_L_type_Ca_channel__i_CaNa__d0);
    const R d_L_type_Ca_channel__i_Ca_L__d10 = (d_L_type_Ca_channel__i_CaNa__d10);
    const R d_L_type_Ca_channel__i_Ca_L__d11 = (d_L_type_Ca_channel__i_CaK__d11);
    const R d_non_specific_calcium_activated_current__I_ns_K__d11 = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div(((((R(1) * L_type_Ca_channel__gamma_Ki * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0))) * ((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_non_specific_calcium_activated_current__i_ns_K__d0 = cmx_div((-((non_specific_calcium_activated_current__I_ns_K * R(1.0)) * ((R(3.0) * cmx_ipow<2>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0])) * cmx_div((-(non_specific_calcium_activated_current__K_m_ns_Ca * R(1))), (g[0] * g[0])))))), ((R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))) * (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0])))));
    const R d_non_specific_calcium_activated_current__i_ns_K__d11 = cmx_div(((d_non_specific_calcium_activated_current__I_ns_K__d11 * R(1.0))), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))));
    const R d_non_specific_calcium_activated_current__I_ns_Na__d10 = cmx_remove_singularity(V, R(-0.0), R(2.6712338705498266e-06), [&](const R V) { return cmx_div(((((R(1) * L_type_Ca_channel__gamma_Nai * cmx_exp_v(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.05400856348800377, 0.0))) * ((non_specific_calcium_activated_current__P_ns_Ca * cmx_ipow<2>(R(1.0)) * V * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((R(1.0) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_non_specific_calcium_activated_current__i_ns_Na__d0 = cmx_div((-((non_specific_calcium_activated_current__I_ns_Na * R(1.0)) * ((R(3.0) * cmx_ipow<2>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0])) * cmx_div((-(non_specific_calcium_activated_current__K_m_ns_Ca * R(1))), (g[0] * g[0])))))), ((R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))) * (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0])))));
    const R d_non_specific_calcium_activated_current__i_ns_Na__d10 = cmx_div(((d_non_specific_calcium_activated_current__I_ns_Na__d10 * R(1.0))), (R(1.0) + cmx_ipow<3>(cmx_div(non_specific_calcium_activated_current__K_m_ns_Ca, g[0]))));
    const R d_non_specific_calcium_activated_current__i_ns_Ca__d0 = (d_non_specific_calcium_activated_current__i_ns_Na__d0 + d_non_specific_calcium_activated_current__i_ns_K__d0);
    const R d_non_specific_calcium_activated_current__i_ns_Ca__d10 = (d_non_specific_calcium_activated_current__i_ns_Na__d10);
    const R d_non_specific_calcium_activated_current__i_ns_Ca__d11 = (d_non_specific_calcium_activated_current__i_ns_K__d11);
    const R d_plateau_potassium_current__i_Kp__d11 = (((-d_time_independent_potassium_current__E_K__d11) * plateau_potassium_current__g_Kp * plateau_potassium_current__Kp));
    const R d_rapid_delayed_rectifier_potassium_current__i_Kr__d11 = (((-d_time_independent_potassium_current__E_K__d11) * rapid_delayed_rectifier_potassium_current__g_Kr * y[8] * rapid_delayed_rectifier_potassium_current__Rect));
    const R d_sarcolemmal_calcium_pump__i_p_Ca__d0 = cmx_div((((R(1) * sarcolemmal_calcium_pump__I_pCa)) * (sarcolemmal_calcium_pump__K_mpCa + g[0]) - (sarcolemmal_calcium_pump__I_pCa * g[0]) * (R(1))), ((sarcolemmal_calcium_pump__K_mpCa + g[0]) * (sarcolemmal_calcium_pump__K_mpCa + g[0])));
    const R d_slow_delayed_rectifier_potassium_current__E_Ks__d10 = ((cmx_div(cmx_div((-((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)) * (((R(1) * slow_delayed_rectifier_potassium_current__PNaK))))), ((g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])) * (g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])))), cmx_div((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)), (g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_slow_delayed_rectifier_potassium_current__E_Ks__d11 = ((cmx_div(cmx_div((-((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)) * (R(1)))), ((g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])) * (g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])))), cmx_div((ionic_concentrations__Ko + (slow_delayed_rectifier_potassium_current__PNaK * ionic_concentrations__Nao)), (g[11] + (slow_delayed_rectifier_potassium_current__PNaK * g[10])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_slow_delayed_rectifier_potassium_current__g_Ks__d0 = (((cmx_div((-(R(0.6) * ((R(1.4) * pow(cmx_div(R(3.8e-05), g[0]), R(1.4) - R(1)) * cmx_div((-(R(3.8e-05) * R(1))), (g[0] * g[0])))))), ((R(1.0) + pow(cmx_div(R(3.8e-05), g[0]), R(1.4))) * (R(1.0) + pow(cmx_div(R(3.8e-05), g[0]), R(1.4)))))) * p[0] * R(0.433)));
    const R d_slow_delayed_rectifier_potassium_current__i_Ks__d0 = ((d_slow_delayed_rectifier_potassium_current__g_Ks__d0 * y[9] * y[10] * (V - slow_delayed_rectifier_potassium_current__E_Ks)));
    const R d_slow_delayed_rectifier_potassium_current__i_Ks__d10 = (((-d_slow_delayed_rectifier_potassium_current__E_Ks__d10) * slow_delayed_rectifier_potassium_current__g_Ks * y[9] * y[10]));
    const R d_slow_delayed_rectifier_potassium_current__i_Ks__d11 = (((-d_slow_delayed_rectifier_potassium_current__E_Ks__d11) * slow_delayed_rectifier_potassium_current__g_Ks * y[9] * y[10]));
    const R d_sodium_activated_potassium_current__pona__d10 = cmx_div((-(R(0.85) * ((sodium_activated_potassium_current__nKNa * pow(cmx_div(sodium_activated_potassium_current__kdKNa, g[10]), sodium_activated_potassium_current__nKNa - R(1)) * cmx_div((-(sodium_activated_potassium_current__kdKNa * R(1))), (g[10] * g[10])))))), ((R(1.0) + pow(cmx_div(sodium_activated_potassium_current__kdKNa, g[10]), sodium_activated_potassium_current__nKNa)) * (R(1.0) + pow(cmx_div(sodium_activated_potassium_current__kdKNa, g[10]), sodium_activated_potassium_current__nKNa))));
    const R d_sodium_activated_potassium_current__i_K_Na__d10 = ((d_sodium_activated_potassium_current__pona__d10 * sodium_activated_potassium_current__g_K_Na * sodium_activated_potassium_current__pov * (V - time_independent_potassium_current__E_K)));
    const R d_sodium_activated_potassium_current__i_K_Na__d11 = (((-d_time_independent_potassium_current__E_K__d11) * sodium_activated_potassium_current__g_K_Na * sodium_activated_potassium_current__pona * sodium_activated_potassium_current__pov));
    const R d_sodium_background_current__i_Na_b__d10 = (((-d_fast_sodium_current__E_Na__d10) * sodium_background_current__g_Nab));
    const R d_sodium_potassium_pump__i_NaK__d10 = cmx_div(((cmx_div((-((sodium_potassium_pump__I_NaK * sodium_potassium_pump__f_NaK * R(1.0)) * ((R(2.0) * cmx_ipow<1>(cmx_div(sodium_potassium_pump__K_mNai, g[10])) * cmx_div((-(sodium_potassium_pump__K_mNai * R(1))), (g[10] * g[10])))))), ((R(1.0) + cmx_ipow<2>(cmx_div(sodium_potassium_pump__K_mNai, g[10]))) * (R(1.0) + cmx_ipow<2>(cmx_div(sodium_potassium_pump__K_mNai, g[10]))))) * ionic_concentrations__Ko)), (ionic_concentrations__Ko + sodium_potassium_pump__K_mKo));
    const R d_time_independent_potassium_current_K1_gate__alpha_K1__d11 = cmx_div((-(R(1020.0) * ((cmx_exp((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215)))) * (((-d_time_independent_potassium_current__E_K__d11) * R(0.2385))))))), ((R(1.0) + cmx_exp_c((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215))), (((V - time_independent_potassium_current__E_K) - R(59.215))), 0.34408276725201775)) * (R(1.0) + cmx_exp_c((R(0.2385) * ((V - time_independent_potassium_current__E_K) - R(59.215))), (((V - time_independent_potassium_current__E_K) - R(59.215))), 0.34408276725201775))));
    const R d_time_independent_potassium_current_K1_gate__beta_K1__d11 = cmx_div((((((((cmx_exp((R(0.08032) * ((V - time_independent_potassium_current__E_K) + R(5.476)))) * ((((-d_time_independent_potassium_current__E_K__d11)) * R(0.08032)))) * R(0.49124))) + (cmx_exp((R(0.06175) * ((V - time_independent_potassium_current__E_K) - R(594.31)))) * (((-d_time_independent_potassium_current__E_K__d11) * R(0.06175))))) * R(1000.0))) * (R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938)) - (R(1000.0) * ((R(0.49124) * cmx_exp_c((R(0.08032) * ((V - time_independent_potassium_current__E_K) + R(5.476))), (((V - time_independent_potassium_current__E_K) + R(5.476))), 0.11587726568420155)) + cmx_exp_c((R(0.06175) * ((V - time_independent_potassium_current__E_K) - R(594.31))), (((V - time_independent_potassium_current__E_K) - R(594.31))), 0.08908641877489348))) * ((cmx_exp(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753)))) * ((((-d_time_independent_potassium_current__E_K__d11)) * (-R(0.5143))))))), ((R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938)) * (R(1.0) + cmx_exp_c(((-R(0.5143)) * ((V - time_independent_potassium_current__E_K) + R(4.753))), (((V - time_independent_potassium_current__E_K) + R(4.753))), -0.7419780595291938))));
    const R d_time_independent_potassium_current_K1_gate__K1_infinity__d11 = cmx_div((d_time_independent_potassium_current_K1_gate__alpha_K1__d11 * (time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1) - time_independent_potassium_current_K1_gate__alpha_K1 * (d_time_independent_potassium_current_K1_gate__alpha_K1__d11 + d_time_independent_potassium_current_K1_gate__beta_K1__d11)), ((time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1) * (time_independent_potassium_current_K1_gate__alpha_K1 + time_independent_potassium_current_K1_gate__beta_K1)));
    const R d_time_independent_potassium_current__i_K1__d11 = ((d_time_independent_potassium_current_K1_gate__K1_infinity__d11 * time_independent_potassium_current__g_K1 * (V - time_independent_potassium_current__E_K)) + ((-d_time_independent_potassium_current__E_K__d11) * time_independent_potassium_current__g_K1 * time_independent_potassium_current_K1_gate__K1_infinity));
    const R d_transient_outward_current__i_to__d11 = (((-d_time_independent_potassium_current__E_K__d11) * transient_outward_current__g_to * cmx_ipow<3>(y[11]) * y[12] * transient_outward_current__rvdv));
    const R d_membrane__dVdt__d0 = (((d_L_type_Ca_channel__i_Ca_L__d0 + d_T_type_Ca_channel__i_Ca_T__d0 + d_slow_delayed_rectifier_potassium_current__i_Ks__d0 + d_Na_Ca_exchanger__i_NaCa__d0 + d_sarcolemmal_calcium_pump__i_p_Ca__d0 + d_calcium_background_current__i_Ca_b__d0 + d_non_specific_calcium_activated_current__i_ns_Ca__d0) * ((-R(1.0)) * (R(1) / membrane__Cm))));
    const R d_membrane__dVdt__d10 = (((d_fast_sodium_current__i_Na__d10 + d_L_type_Ca_channel__i_Ca_L__d10 + d_slow_delayed_rectifier_potassium_current__i_Ks__d10 + d_sodium_activated_potassium_current__i_K_Na__d10 + d_Na_Ca_exchanger__i_NaCa__d10 + d_sodium_background_current__i_Na_b__d10 + d_sodium_potassium_pump__i_NaK__d10 + d_non_specific_calcium_activated_current__i_ns_Ca__d10) * ((-R(1.0)) * (R(1) / membrane__Cm))));
    const R d_membrane__dVdt__d11 = (((d_L_type_Ca_channel__i_Ca_L__d11 + d_rapid_delayed_rectifier_potassium_current__i_Kr__d11 + d_slow_delayed_rectifier_potassium_current__i_Ks__d11 + d_sodium_activated_potassium_current__i_K_Na__d11 + d_ATP_sensitive_potassium_current__i_K_ATP__d11 + d_transient_outward_current__i_to__d11 + d_time_independent_potassium_current__i_K1__d11 + d_plateau_potassium_current__i_Kp__d11 + d_non_specific_calcium_activated_current__i_ns_Ca__d11) * ((-R(1.0)) * (R(1) / membrane__Cm))));
    J[0] = R(1) - dt * cmx_div(((cmx_div((-(R(1.0) * (cmx_div((-((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN) * (R(2.0) * cmx_ipow<1>((calcium_dynamics__K_mCMDN + g[0])) * (R(1))))), (cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0])) * cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0])))) + cmx_div((-((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn) * (R(2.0) * cmx_ipow<1>((calcium_dynamics__K_mTn + g[0])) * (R(1))))), (cmx_ipow<2>((calcium_dynamics__K_mTn + g[0])) * cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), ((R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0])))) * (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0])))))) * ((((-ionic_concentrations__A_cap) * (((L_type_Ca_channel__i_CaCa + T_type_Ca_channel__i_Ca_T) - (R(2.0) * Na_Ca_exchanger__i_NaCa)) + sarcolemmal_calcium_pump__i_p_Ca + calcium_background_current__i_Ca_b)) * (R(1) / (R(2.0) * ionic_concentrations__V_myo * membrane__F))) + ((calcium_dynamics__i_rel * calcium_dynamics__V_JSR) * (R(1) / ionic_concentrations__V_myo)) + (((calcium_dynamics__i_leak - calcium_dynamics__i_up) * calcium_dynamics__V_NSR) * (R(1) / ionic_concentrations__V_myo)))) + ((cmx_div((((((d_L_type_Ca_channel__i_CaCa__d0 + d_T_type_Ca_channel__i_Ca_T__d0) - ((d_Na_Ca_exchanger__i_NaCa__d0 * R(2.0)))) + d_sarcolemmal_calcium_pump__i_p_Ca__d0 + d_calcium_background_current__i_Ca_b__d0) * (-ionic_concentrations__A_cap))), (R(2.0) * ionic_concentrations__V_myo * membrane__F)) + cmx_div(((d_calcium_dynamics__i_rel__d0 * calcium_dynamics__V_JSR)), ionic_concentrations__V_myo) + cmx_div((((-d_calcium_dynamics__i_up__d0) * calcium_dynamics__V_NSR)), ionic_concentrations__V_myo)) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), R(1000.0));
    J[1] = R(0) - dt * cmx_div((((cmx_div(((d_calcium_dynamics__i_rel__d1 * calcium_dynamics__V_JSR)), ionic_concentrations__V_myo)) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), R(1000.0));
    J[2] = R(0) - dt * cmx_div((((cmx_div(((d_calcium_dynamics__i_leak__d2 * calcium_dynamics__V_NSR)), ionic_concentrations__V_myo)) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), R(1000.0));
    J[3] = R(0);
    J[4] = R(0) - dt * cmx_div((((cmx_div(((d_calcium_dynamics__i_rel__d4 * calcium_dynamics__V_JSR)), ionic_concentrations__V_myo)) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), R(1000.0));
    J[5] = R(0);
    J[6] = R(0) - dt * cmx_div((((cmx_div(((d_calcium_dynamics__i_rel__d6 * calcium_dynamics__V_JSR)), ionic_concentrations__V_myo)) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), R(1000.0));
    J[7] = R(0);
    J[8] = R(0) - dt * cmx_div((((cmx_div(((d_calcium_dynamics__i_rel__d8 * calcium_dynamics__V_JSR)), ionic_concentrations__V_myo)) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), R(1000.0));
    J[9] = R(0);
    J[10] = R(0) - dt * cmx_div((((cmx_div(((((-((d_Na_Ca_exchanger__i_NaCa__d10 * R(2.0))))) * (-ionic_concentrations__A_cap))), (R(2.0) * ionic_concentrations__V_myo * membrane__F))) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CMDN_max * calcium_dynamics__K_mCMDN), cmx_ipow<2>((calcium_dynamics__K_mCMDN + g[0]))) + cmx_div((calcium_dynamics__Tn_max * calcium_dynamics__K_mTn), cmx_ipow<2>((calcium_dynamics__K_mTn + g[0]))))))), R(1000.0));
    J[11] = R(0);
    J[12] = R(0) - dt * cmx_div((((-d_calcium_dynamics__i_rel__d0) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))))), R(1000.0));
    J[13] = R(1) - dt * cmx_div(((cmx_div((-(R(1.0) * (cmx_div((-((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN) * (R(2.0) * cmx_ipow<1>((calcium_dynamics__K_mCSQN + g[1])) * (R(1))))), (cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1])) * cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))))), ((R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1])))) * (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1])))))) * (calcium_dynamics__i_tr - calcium_dynamics__i_rel)) + ((d_calcium_dynamics__i_tr__d1 - d_calcium_dynamics__i_rel__d1) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))))), R(1000.0));
    J[14] = R(0) - dt * cmx_div(((d_calcium_dynamics__i_tr__d2 * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))))), R(1000.0));
    J[15] = R(0);
    J[16] = R(0) - dt * cmx_div((((-d_calcium_dynamics__i_rel__d4) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))))), R(1000.0));
    J[17] = R(0);
    J[18] = R(0) - dt * cmx_div((((-d_calcium_dynamics__i_rel__d6) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))))), R(1000.0));
    J[19] = R(0);
    J[20] = R(0) - dt * cmx_div((((-d_calcium_dynamics__i_rel__d8) * cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__CSQN_max * calcium_dynamics__K_mCSQN), cmx_ipow<2>((calcium_dynamics__K_mCSQN + g[1]))))))), R(1000.0));
    J[21] = R(0);
    J[22] = R(0);
    J[23] = R(0);
    J[24] = R(0) - dt * cmx_div((d_calcium_dynamics__i_up__d0), R(1000.0));
    J[25] = R(0) - dt * cmx_div((cmx_div((((-d_calcium_dynamics__i_tr__d1) * calcium_dynamics__V_JSR)), calcium_dynamics__V_NSR)), R(1000.0));
    J[26] = R(1) - dt * cmx_div(((cmx_div((((-d_calcium_dynamics__i_tr__d2) * calcium_dynamics__V_JSR)), calcium_dynamics__V_NSR) - d_calcium_dynamics__i_leak__d2)), R(1000.0));
    J[27] = R(0);
    J[28] = R(0);
    J[29] = R(0);
    J[30] = R(0);
    J[31] = R(0);
    J[32] = R(0);
    J[33] = R(0);
    J[34] = R(0);
    J[35] = R(0);
    J[36] = R(0);
    J[37] = R(0);
    J[38] = R(0);
    J[39] = R(1) - dt * cmx_div((((membrane__dVdt > R(150000.0))) ? (((((-R(1)) * R(100000.0))) - ((R(1) * R(500.0))))) : (((R(1) * (-R(500.0)))))), R(1000.0));
    J[40] = R(0);
    J[41] = R(0);
    J[42] = R(0);
    J[43] = R(0);
    J[44] = R(0);
    J[45] = R(0);
    J[46] = R(0);
    J[47] = R(0);
    J[48] = R(0);
    J[49] = R(0);
    J[50] = R(0);
    J[51] = R(0);
    J[52] = R(1) - dt * cmx_div(((((g[3] < R(0.2)) && (g[3] > R(0.18)))) ? (((((-R(1)) * R(100000.0))) - ((R(1) * R(500.0))))) : (((R(1) * (-R(500.0)))))), R(1000.0));
    J[53] = R(0);
    J[54] = R(0);
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
    J[65] = R(1) - dt * cmx_div(((((g[3] < R(0.2)) && (g[3] > R(0.18)))) ? (((((-R(1)) * R(100000.0))) - ((R(1) * R(500.0))))) : (((R(1) * (-R(10.0)))))), R(1000.0));
    J[66] = R(0);
    J[67] = R(0);
    J[68] = R(0);
    J[69] = R(0);
    J[70] = R(0);
    J[71] = R(0);
    J[72] = R(0) - dt * cmx_div((((g[3] > R(0.2))) ? (cmx_div((((((d_L_type_Ca_channel__i_CaCa__d0 + d_T_type_Ca_channel__i_Ca_T__d0) - d_Na_Ca_exchanger__i_NaCa__d0) + d_sarcolemmal_calcium_pump__i_p_Ca__d0 + d_calcium_background_current__i_Ca_b__d0) * (-ionic_concentrations__A_cap))), (R(2.0) * ionic_concentrations__V_myo * membrane__F))) : (((((g[4] > R(0.01)) && (g[3] <= R(0.2)))) ? (R(0)) : (R(0))))), R(1000.0));
    J[73] = R(0);
    J[74] = R(0);
    J[75] = R(0);
    J[76] = R(0);
    J[77] = R(0);
    J[78] = R(1) - dt * cmx_div((((g[3] > R(0.2))) ? (R(0)) : (((((g[4] > R(0.01)) && (g[3] <= R(0.2)))) ? (R(0)) : (((R(1) * (-R(500.0)))))))), R(1000.0));
    J[79] = R(0);
    J[80] = R(0);
    J[81] = R(0);
    J[82] = R(0) - dt * cmx_div((((g[3] > R(0.2))) ? (cmx_div(((((-d_Na_Ca_exchanger__i_NaCa__d10)) * (-ionic_concentrations__A_cap))), (R(2.0) * ionic_concentrations__V_myo * membrane__F))) : (((((g[4] > R(0.01)) && (g[3] <= R(0.2)))) ? (R(0)) : (R(0))))), R(1000.0));
    J[83] = R(0);
    J[84] = R(0);
    J[85] = R(0);
    J[86] = R(0);
    J[87] = R(0);
    J[88] = R(0);
    J[89] = R(0);
    J[90] = R(0);
    J[91] = R(1) - dt * cmx_div(((((cmx_div(R(1.0), (R(1.0) + cmx_div(calcium_dynamics__K_mCSQN, g[1]))) > calcium_dynamics__CSQNthresh) && (g[9] < R(0.37)) && (g[5] < R(0.37)))) ? ((((-R(1)) * R(50000.0)))) : (((R(1) * (-R(500.0)))))), R(1000.0));
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
    J[103] = R(0);
    J[104] = R(1) - dt * cmx_div(((((g[7] > calcium_dynamics__Logicthresh) && (g[8] < calcium_dynamics__Logicthresh))) ? ((((-R(1)) * R(50000.0)))) : (((R(1) * (-R(500.0)))))), R(1000.0));
    J[105] = R(0);
    J[106] = R(0);
    J[107] = R(0);
    J[108] = R(0);
    J[109] = R(0);
    J[110] = R(0);
    J[111] = R(0);
    J[112] = R(0);
    J[113] = R(0);
    J[114] = R(0);
    J[115] = R(0);
    J[116] = R(0);
    J[117] = R(1) - dt * cmx_div(((((g[7] > calcium_dynamics__Logicthresh) && (g[9] < calcium_dynamics__Logicthresh))) ? ((((-R(1)) * R(50000.0)))) : (((R(1) * (-R(10.0)))))), R(1000.0));
    J[118] = R(0);
    J[119] = R(0);
    J[120] = R(0) - dt * cmx_div(cmx_div((((-(d_L_type_Ca_channel__i_CaNa__d0 + d_non_specific_calcium_activated_current__i_ns_Na__d0 + ((d_Na_Ca_exchanger__i_NaCa__d0 * R(3.0))))) * ionic_concentrations__A_cap)), (ionic_concentrations__V_myo * membrane__F)), R(1000.0));
    J[121] = R(0);
    J[122] = R(0);
    J[123] = R(0);
    J[124] = R(0);
    J[125] = R(0);
    J[126] = R(0);
    J[127] = R(0);
    J[128] = R(0);
    J[129] = R(0);
    J[130] = R(1) - dt * cmx_div(cmx_div((((-(d_fast_sodium_current__i_Na__d10 + d_L_type_Ca_channel__i_CaNa__d10 + d_sodium_background_current__i_Na_b__d10 + d_non_specific_calcium_activated_current__i_ns_Na__d10 + ((d_Na_Ca_exchanger__i_NaCa__d10 * R(3.0))) + ((d_sodium_potassium_pump__i_NaK__d10 * R(3.0))))) * ionic_concentrations__A_cap)), (ionic_concentrations__V_myo * membrane__F)), R(1000.0));
    J[131] = R(0);
    J[132] = R(0) - dt * cmx_div(cmx_div((((-(d_L_type_Ca_channel__i_CaK__d0 + d_slow_delayed_rectifier_potassium_current__i_Ks__d0 + d_non_specific_calcium_activated_current__i_ns_K__d0)) * ionic_concentrations__A_cap)), (ionic_concentrations__V_myo * membrane__F)), R(1000.0));
    J[133] = R(0);
    J[134] = R(0);
    J[135] = R(0);
    J[136] = R(0);
    J[137] = R(0);
    J[138] = R(0);
    J[139] = R(0);
    J[140] = R(0);
    J[141] = R(0);
    J[142] = R(0) - dt * cmx_div(cmx_div((((-(d_slow_delayed_rectifier_potassium_current__i_Ks__d10 + d_sodium_activated_potassium_current__i_K_Na__d10 + (((-d_sodium_potassium_pump__i_NaK__d10) * R(2.0))))) * ionic_concentrations__A_cap)), (ionic_concentrations__V_myo * membrane__F)), R(1000.0));
    J[143] = R(1) - dt * cmx_div(cmx_div((((-(d_L_type_Ca_channel__i_CaK__d11 + d_rapid_delayed_rectifier_potassium_current__i_Kr__d11 + d_slow_delayed_rectifier_potassium_current__i_Ks__d11 + d_time_independent_potassium_current__i_K1__d11 + d_plateau_potassium_current__i_Kp__d11 + d_sodium_activated_potassium_current__i_K_Na__d11 + d_ATP_sensitive_potassium_current__i_K_ATP__d11 + d_transient_outward_current__i_to__d11 + d_non_specific_calcium_activated_current__i_ns_K__d11)) * ionic_concentrations__A_cap)), (ionic_concentrations__V_myo * membrane__F)), R(1000.0));
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

}}  // namespace cmx::fr2000
