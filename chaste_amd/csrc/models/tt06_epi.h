// Generated by chaste_amd/codegen/cellml2cpp.py from TenTusscher2006Epi.cellml -- do not edit.
// Model: tentusscher_model_2006_epi
#pragma once
#include "model_common.h"

namespace cmx { namespace tt06_epi {

constexpr int kNumStates = 19;
constexpr int kVIndex = 0;
constexpr int kNumParams = 6;
constexpr bool kPerCapacitanceUnits = true;

// State variables (CellML component__variable), V first then document order.
static const char* const kStateNames[kNumStates] = {
    "membrane__V",
    "rapid_time_dependent_potassium_current_Xr1_gate__Xr1",
    "rapid_time_dependent_potassium_current_Xr2_gate__Xr2",
    "slow_time_dependent_potassium_current_Xs_gate__Xs",
    "fast_sodium_current_m_gate__m",
    "fast_sodium_current_h_gate__h",
    "fast_sodium_current_j_gate__j",
    "L_type_Ca_current_d_gate__d",
    "L_type_Ca_current_f_gate__f",
    "L_type_Ca_current_f2_gate__f2",
    "L_type_Ca_current_fCass_gate__fCass",
    "transient_outward_current_s_gate__s",
    "transient_outward_current_r_gate__r",
    "calcium_dynamics__Ca_i",
    "calcium_dynamics__Ca_SR",
    "calcium_dynamics__Ca_ss",
    "calcium_dynamics__R_prime",
    "sodium_dynamics__Na_i",
    "potassium_dynamics__K_i",
};
static constexpr double kInitialState[kNumStates] = {
    -85.23,  // membrane__V
    0.00621,  // rapid_time_dependent_potassium_current_Xr1_gate__Xr1
    0.4712,  // rapid_time_dependent_potassium_current_Xr2_gate__Xr2
    0.0095,  // slow_time_dependent_potassium_current_Xs_gate__Xs
    0.00172,  // fast_sodium_current_m_gate__m
    0.7444,  // fast_sodium_current_h_gate__h
    0.7045,  // fast_sodium_current_j_gate__j
    3.373e-05,  // L_type_Ca_current_d_gate__d
    0.7888,  // L_type_Ca_current_f_gate__f
    0.9755,  // L_type_Ca_current_f2_gate__f2
    0.9953,  // L_type_Ca_current_fCass_gate__fCass
    0.999998,  // transient_outward_current_s_gate__s
    2.42e-08,  // transient_outward_current_r_gate__r
    0.000126,  // calcium_dynamics__Ca_i
    3.64,  // calcium_dynamics__Ca_SR
    0.00036,  // calcium_dynamics__Ca_ss
    0.9073,  // calcium_dynamics__R_prime
    8.604,  // sodium_dynamics__Na_i
    136.89,  // potassium_dynamics__K_i
};
static const char* const kParamNames[6] = {
    "fast_sodium_current__g_Na",
    "L_type_Ca_current__g_CaL",
    "rapid_time_dependent_potassium_current__g_Kr",
    "slow_time_dependent_potassium_current__g_Ks",
    "inward_rectifier_potassium_current__g_K1",
    "transient_outward_current__g_to",
};
// The names the reference gives the same parameters (the CellML oxmeta term,
// else the variable's cmeta:id; empty when the file annotates neither).
static const char* const kParamAliases[6] = {
    "",
    "",
    "",
    "",
    "",
    "",
};
static constexpr double kParamDefaults[6] = {
    14.838,
    3.98e-05,
    0.153,
    0.392,
    5.405,
    0.294,
};
// 1 = integrated by Rush-Larsen (affine gate, V-dependent coefficients only)
static constexpr int kIsGate[kNumStates] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 0, 0, 0, 0, 0, 0};

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
    const R membrane__Cm = R(0.185);
    const R membrane__V_c = R(0.016404);
    const R reversal_potentials__P_kna = R(0.03);
    const R rapid_time_dependent_potassium_current__ScaleFactorGkr = R(1.0);
    const R slow_time_dependent_potassium_current__ScaleFactorGks = R(1.0);
    const R sodium_background_current__g_bna = R(0.00029);
    const R calcium_background_current__g_bca = R(0.000592);
    const R transient_outward_current__ScaleFactorIto = R(1.0);
    const R sodium_potassium_pump_current__P_NaK = R(2.724);
    const R sodium_potassium_pump_current__K_mk = R(1.0);
    const R sodium_potassium_pump_current__K_mNa = R(40.0);
    const R sodium_calcium_exchanger_current__K_NaCa = R(1000.0);
    const R sodium_calcium_exchanger_current__K_sat = R(0.1);
    const R sodium_calcium_exchanger_current__alpha = R(2.5);
    const R sodium_calcium_exchanger_current__gamma = R(0.35);
    const R sodium_calcium_exchanger_current__Km_Ca = R(1.38);
    const R sodium_calcium_exchanger_current__Km_Nai = R(87.5);
    const R calcium_pump_current__g_pCa = R(0.1238);
    const R calcium_pump_current__K_pCa = R(0.0005);
    const R potassium_pump_current__g_pK = R(0.0146);
    const R calcium_dynamics__Ca_o = R(2.0);
    const R calcium_dynamics__k1_prime = R(0.15);
    const R calcium_dynamics__k2_prime = R(0.045);
    const R calcium_dynamics__k3 = R(0.06);
    const R calcium_dynamics__k4 = R(0.005);
    const R calcium_dynamics__EC = R(1.5);
    const R calcium_dynamics__max_sr = R(2.5);
    const R calcium_dynamics__min_sr = R(1.0);
    const R calcium_dynamics__V_rel = R(0.102);
    const R calcium_dynamics__V_xfer = R(0.0038);
    const R calcium_dynamics__K_up = R(0.00025);
    const R calcium_dynamics__V_leak = R(0.00036);
    const R calcium_dynamics__Vmax_up = R(0.006375);
    const R calcium_dynamics__Buf_c = R(0.2);
    const R calcium_dynamics__K_buf_c = R(0.001);
    const R calcium_dynamics__Buf_sr = R(10.0);
    const R calcium_dynamics__K_buf_sr = R(0.3);
    const R calcium_dynamics__Buf_ss = R(0.4);
    const R calcium_dynamics__K_buf_ss = R(0.00025);
    const R calcium_dynamics__V_sr = R(0.001094);
    const R calcium_dynamics__V_ss = R(5.468e-05);
    const R sodium_dynamics__Na_o = R(140.0);
    const R potassium_dynamics__K_o = R(5.4);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_CaL = cmx_remove_singularity(V, R(15.00000000000003), R(1.335688032984785e-06), [&](const R V) { return cmx_div((((p[1] * y[7] * y[8] * y[9] * y[10] * R(4.0) * (V - R(15.0)) * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((R(0.25) * y[15] * cmx_exp_v(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801137730230755, -1.620170659534613)) - calcium_dynamics__Ca_o)), cmx_expm1(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current_d_gate__d_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(8.0)) - V) * (R(1) / R(7.5))), V, -0.19235933878519512, -1.538874710281561)));
    const R L_type_Ca_current_d_gate__alpha_d = (cmx_div(R(1.4), (R(1.0) + cmx_exp_v((((-R(35.0)) - V) * (R(1) / R(13.0))), V, -0.11097654160684334, -3.8841789562395173))) + R(0.25));
    const R L_type_Ca_current_d_gate__beta_d = cmx_div(R(1.4), (R(1.0) + cmx_exp_v(((V + R(5.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 1.4426950408889634)));
    const R L_type_Ca_current_d_gate__gamma_d = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(50.0) - V) * (R(1) / R(20.0))), V, -0.07213475204444818, 3.6067376022224087)));
    const R L_type_Ca_current_d_gate__tau_d = ((L_type_Ca_current_d_gate__alpha_d * L_type_Ca_current_d_gate__beta_d) + L_type_Ca_current_d_gate__gamma_d);
    const R L_type_Ca_current_f2_gate__f2_inf = (cmx_div(R(0.67), (R(1.0) + cmx_exp_v(((V + R(35.0)) * (R(1) / R(7.0))), V, 0.2060992915555662, 7.213475204444817))) + R(0.33));
    const R L_type_Ca_current_f2_gate__tau_f2 = ((R(562.0) * cmx_exp_c(((-cmx_ipow<2>((V + R(27.0)))) * (R(1) / R(240.0))), (-cmx_ipow<2>((V + R(27.0)))), 0.0060112293370373476)) + cmx_div(R(31.0), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(10.0))), V, -0.14426950408889636, 3.6067376022224087))) + cmx_div(R(80.0), (R(1.0) + cmx_exp_v(((V + R(30.0)) * (R(1) / R(10.0))), V, 0.14426950408889636, 4.32808512266689))));
    const R L_type_Ca_current_fCass_gate__fCass_inf = (cmx_div(R(0.6), (R(1.0) + cmx_ipow<2>((y[15] * (R(1) / R(0.05)))))) + R(0.4));
    const R L_type_Ca_current_fCass_gate__tau_fCass = (cmx_div(R(80.0), (R(1.0) + cmx_ipow<2>((y[15] * (R(1) / R(0.05)))))) + R(2.0));
    const R L_type_Ca_current_f_gate__f_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(20.0)) * (R(1) / R(7.0))), V, 0.20609929155556622, 4.121985831111324)));
    const R L_type_Ca_current_f_gate__tau_f = ((R(1102.5) * cmx_exp_c(((-cmx_ipow<2>((V + R(27.0)))) * (R(1) / R(225.0))), (-cmx_ipow<2>((V + R(27.0)))), 0.006411977959506504)) + cmx_div(R(200.0), (R(1.0) + cmx_exp_v(((R(13.0) - V) * (R(1) / R(10.0))), V, -0.14426950408889636, 1.8755035531556525))) + cmx_div(R(180.0), (R(1.0) + cmx_exp_v(((V + R(30.0)) * (R(1) / R(10.0))), V, 0.14426950408889636, 4.32808512266689))) + R(20.0));
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(calcium_dynamics__Ca_o, y[13])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_dynamics__Ca_i_bufc = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_c * calcium_dynamics__K_buf_c), cmx_ipow<2>((y[13] + calcium_dynamics__K_buf_c)))));
    const R calcium_dynamics__Ca_sr_bufsr = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_sr * calcium_dynamics__K_buf_sr), cmx_ipow<2>((y[14] + calcium_dynamics__K_buf_sr)))));
    const R calcium_dynamics__Ca_ss_bufss = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_ss * calcium_dynamics__K_buf_ss), cmx_ipow<2>((y[15] + calcium_dynamics__K_buf_ss)))));
    const R calcium_dynamics__i_leak = (calcium_dynamics__V_leak * (y[14] - y[13]));
    const R calcium_dynamics__kcasr = (calcium_dynamics__max_sr - cmx_div((calcium_dynamics__max_sr - calcium_dynamics__min_sr), (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__EC, y[14])))));
    const R calcium_dynamics__k1 = cmx_div(calcium_dynamics__k1_prime, calcium_dynamics__kcasr);
    const R calcium_dynamics__O = cmx_div((calcium_dynamics__k1 * cmx_ipow<2>(y[15]) * y[16]), (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(y[15]))));
    const R calcium_dynamics__i_rel = (calcium_dynamics__V_rel * calcium_dynamics__O * (y[14] - y[15]));
    const R calcium_dynamics__i_up = cmx_div(calcium_dynamics__Vmax_up, (R(1.0) + cmx_div(cmx_ipow<2>(calcium_dynamics__K_up), cmx_ipow<2>(y[13]))));
    const R calcium_dynamics__i_xfer = (calcium_dynamics__V_xfer * (y[15] - y[13]));
    const R calcium_dynamics__k2 = (calcium_dynamics__k2_prime * calcium_dynamics__kcasr);
    const R calcium_pump_current__i_p_Ca = cmx_div((calcium_pump_current__g_pCa * y[13]), (y[13] + calcium_pump_current__K_pCa));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(sodium_dynamics__Na_o, y[17])));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * y[6] * (V - reversal_potentials__E_Na));
    const R fast_sodium_current_h_gate__h_inf = cmx_div(R(1.0), cmx_ipow<2>((R(1.0) + cmx_exp_v(((V + R(71.55)) * (R(1) / R(7.43))), V, 0.19417160711829928, 13.89297848931431))));
    const R fast_sodium_current_h_gate__alpha_h = (((V < (-R(40.0)))) ? ((R(0.057) * cmx_exp_v(((-(V + R(80.0))) * (R(1) / R(6.8))), V, -0.2121610354248476, -16.972882833987807))) : (R(0.0)));
    const R fast_sodium_current_h_gate__beta_h = (((V < (-R(40.0)))) ? (((R(2.7) * cmx_exp_v((R(0.079) * V), V, 0.11397290823022813, 0.0)) + (R(310000.0) * cmx_exp_v((R(0.3485) * V), V, 0.5027792217498037, 0.0)))) : (cmx_div(R(0.77), (R(0.13) * (R(1.0) + cmx_exp_v(((V + R(10.66)) * (R(1) / (-R(11.1)))), V, -0.12997252620621294, -1.3855071293582297))))));
    const R fast_sodium_current_h_gate__tau_h = cmx_div(R(1.0), (fast_sodium_current_h_gate__alpha_h + fast_sodium_current_h_gate__beta_h));
    const R fast_sodium_current_j_gate__j_inf = cmx_div(R(1.0), cmx_ipow<2>((R(1.0) + cmx_exp_v(((V + R(71.55)) * (R(1) / R(7.43))), V, 0.19417160711829928, 13.89297848931431))));
    const R fast_sodium_current_j_gate__alpha_j = (((V < (-R(40.0)))) ? (cmx_div(((((-R(25428.0)) * cmx_exp_v((R(0.2444) * V), V, 0.3525946679932627, 0.0)) - (R(6.948e-06) * cmx_exp_v(((-R(0.04391)) * V), V, -0.06334873924543438, -0.0))) * (V + R(37.78))), (R(1.0) + cmx_exp_v((R(0.311) * (V + R(79.23))), V, 0.44867815771646763, 35.54877043587573)))) : (R(0.0)));
    const R fast_sodium_current_j_gate__beta_j = (((V < (-R(40.0)))) ? (cmx_div((R(0.02424) * cmx_exp_v(((-R(0.01052)) * V), V, -0.015177151830151895, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.1378)) * (V + R(40.14))), V, -0.19880337663449915, -7.979967538108797)))) : (cmx_div((R(0.6) * cmx_exp_v((R(0.057) * V), V, 0.08223361733067092, 0.0)), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(32.0))), V, -0.14426950408889636, -4.616624130844683)))));
    const R fast_sodium_current_j_gate__tau_j = cmx_div(R(1.0), (fast_sodium_current_j_gate__alpha_j + fast_sodium_current_j_gate__beta_j));
    const R fast_sodium_current_m_gate__m_inf = cmx_div(R(1.0), cmx_ipow<2>((R(1.0) + cmx_exp_v((((-R(56.86)) - V) * (R(1) / R(9.03))), V, -0.1597668926787335, -9.084345517712785))));
    const R fast_sodium_current_m_gate__alpha_m = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(60.0)) - V) * (R(1) / R(5.0))), V, -0.2885390081777927, -17.31234049066756)));
    const R fast_sodium_current_m_gate__beta_m = (cmx_div(R(0.1), (R(1.0) + cmx_exp_v(((V + R(35.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 10.098865286222743))) + cmx_div(R(0.1), (R(1.0) + cmx_exp_v(((V - R(50.0)) * (R(1) / R(200.0))), V, 0.007213475204444817, -0.36067376022224085))));
    const R fast_sodium_current_m_gate__tau_m = (fast_sodium_current_m_gate__alpha_m * fast_sodium_current_m_gate__beta_m);
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(potassium_dynamics__K_o, y[18])));
    const R inward_rectifier_potassium_current__alpha_K1 = cmx_div(R(0.1), (R(1.0) + cmx_exp_c((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0))), (((V - reversal_potentials__E_K) - R(200.0))), 0.0865617024533378)));
    const R inward_rectifier_potassium_current__beta_K1 = cmx_div(((R(3.0) * cmx_exp_c((R(0.0002) * ((V - reversal_potentials__E_K) + R(100.0))), (((V - reversal_potentials__E_K) + R(100.0))), 0.0002885390081777927)) + cmx_exp_c((R(0.1) * ((V - reversal_potentials__E_K) - R(10.0))), (((V - reversal_potentials__E_K) - R(10.0))), 0.14426950408889636)), (R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817)));
    const R inward_rectifier_potassium_current__xK1_inf = cmx_div(inward_rectifier_potassium_current__alpha_K1, (inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1));
    const R inward_rectifier_potassium_current__i_K1 = (p[4] * inward_rectifier_potassium_current__xK1_inf * (V - reversal_potentials__E_K));
    const R potassium_pump_current__i_p_K = cmx_div((potassium_pump_current__g_pK * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(5.98))), V, -0.24125335131922465, 6.0313337829806155)));
    const R rapid_time_dependent_potassium_current__i_Kr = (rapid_time_dependent_potassium_current__ScaleFactorGkr * p[2] * sqrt((potassium_dynamics__K_o * (R(1) / R(5.4)))) * y[1] * y[2] * (V - reversal_potentials__E_K));
    const R rapid_time_dependent_potassium_current_Xr1_gate__alpha_xr1 = cmx_div(R(450.0), (R(1.0) + cmx_exp_v((((-R(45.0)) - V) * (R(1) / R(10.0))), V, -0.14426950408889636, -6.4921276840003355)));
    const R rapid_time_dependent_potassium_current_Xr1_gate__beta_xr1 = cmx_div(R(6.0), (R(1.0) + cmx_exp_v(((V + R(30.0)) * (R(1) / R(11.5))), V, 0.12545174268599682, 3.7635522805799044)));
    const R rapid_time_dependent_potassium_current_Xr1_gate__tau_xr1 = (rapid_time_dependent_potassium_current_Xr1_gate__alpha_xr1 * rapid_time_dependent_potassium_current_Xr1_gate__beta_xr1);
    const R rapid_time_dependent_potassium_current_Xr1_gate__xr1_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(26.0)) - V) * (R(1) / R(7.0))), V, -0.20609929155556622, -5.358581580444722)));
    const R rapid_time_dependent_potassium_current_Xr2_gate__alpha_xr2 = cmx_div(R(3.0), (R(1.0) + cmx_exp_v((((-R(60.0)) - V) * (R(1) / R(20.0))), V, -0.07213475204444818, -4.32808512266689)));
    const R rapid_time_dependent_potassium_current_Xr2_gate__beta_xr2 = cmx_div(R(1.12), (R(1.0) + cmx_exp_v(((V - R(60.0)) * (R(1) / R(20.0))), V, 0.07213475204444818, -4.32808512266689)));
    const R rapid_time_dependent_potassium_current_Xr2_gate__tau_xr2 = (rapid_time_dependent_potassium_current_Xr2_gate__alpha_xr2 * rapid_time_dependent_potassium_current_Xr2_gate__beta_xr2);
    const R rapid_time_dependent_potassium_current_Xr2_gate__xr2_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(88.0)) * (R(1) / R(24.0))), V, 0.06011229337037347, 5.289881816592866)));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)), (y[18] + (reversal_potentials__P_kna * y[17])))));
    const R slow_time_dependent_potassium_current__i_Ks = (slow_time_dependent_potassium_current__ScaleFactorGks * p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R slow_time_dependent_potassium_current_Xs_gate__alpha_xs = cmx_div(R(1400.0), sqrt((R(1.0) + cmx_exp_v(((R(5.0) - V) * (R(1) / R(6.0))), V, -0.2404491734814939, 1.2022458674074696))));
    const R slow_time_dependent_potassium_current_Xs_gate__beta_xs = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(35.0)) * (R(1) / R(15.0))), V, 0.09617966939259756, -3.3662884287409147)));
    const R slow_time_dependent_potassium_current_Xs_gate__tau_xs = ((slow_time_dependent_potassium_current_Xs_gate__alpha_xs * slow_time_dependent_potassium_current_Xs_gate__beta_xs) + R(80.0));
    const R slow_time_dependent_potassium_current_Xs_gate__xs_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(5.0)) - V) * (R(1) / R(14.0))), V, -0.10304964577778311, -0.5152482288889155)));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger_current__i_NaCa = cmx_div((sodium_calcium_exchanger_current__K_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger_current__gamma * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.01890199102790382, 0.0) * cmx_ipow<3>(y[17]) * calcium_dynamics__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0) * cmx_ipow<3>(sodium_dynamics__Na_o) * y[13] * sodium_calcium_exchanger_current__alpha))), ((cmx_ipow<3>(sodium_calcium_exchanger_current__Km_Nai) + cmx_ipow<3>(sodium_dynamics__Na_o)) * (sodium_calcium_exchanger_current__Km_Ca + calcium_dynamics__Ca_o) * (R(1.0) + (sodium_calcium_exchanger_current__K_sat * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0)))));
    const R sodium_potassium_pump_current__i_NaK = cmx_div(cmx_div((((sodium_potassium_pump_current__P_NaK * potassium_dynamics__K_o) * (R(1) / (potassium_dynamics__K_o + sodium_potassium_pump_current__K_mk))) * y[17]), (y[17] + sodium_potassium_pump_current__K_mNa)), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400568865115377, -0.0)) + (R(0.0353) * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115376, -0.0))));
    const R transient_outward_current__i_to = (transient_outward_current__ScaleFactorIto * p[5] * y[12] * y[11] * (V - reversal_potentials__E_K));
    const R transient_outward_current_r_gate__r_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(20.0) - V) * (R(1) / R(6.0))), V, -0.2404491734814939, 4.8089834696298785)));
    const R transient_outward_current_r_gate__tau_r = ((R(9.5) * cmx_exp_c(((-cmx_ipow<2>((V + R(40.0)))) * (R(1) / R(1800.0))), (-cmx_ipow<2>((V + R(40.0)))), 0.000801497244938313)) + R(0.8));
    const R transient_outward_current_s_gate__s_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(20.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 5.7707801635558535)));
    const R transient_outward_current_s_gate__tau_s = ((R(85.0) * cmx_exp_c(((-cmx_ipow<2>((V + R(45.0)))) * (R(1) / R(320.0))), (-cmx_ipow<2>((V + R(45.0)))), 0.004508422002778011)) + cmx_div(R(5.0), (R(1.0) + cmx_exp_v(((V - R(20.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, -5.7707801635558535))) + R(3.0));
    if (RL) {
        g_k[1] = cmx_div(R(-1), rapid_time_dependent_potassium_current_Xr1_gate__tau_xr1);
        g_inf[1] = rapid_time_dependent_potassium_current_Xr1_gate__xr1_inf;
    } else {
        dy[1] = cmx_div((rapid_time_dependent_potassium_current_Xr1_gate__xr1_inf - y[1]), rapid_time_dependent_potassium_current_Xr1_gate__tau_xr1);
    }
    if (RL) {
        g_k[2] = cmx_div(R(-1), rapid_time_dependent_potassium_current_Xr2_gate__tau_xr2);
        g_inf[2] = rapid_time_dependent_potassium_current_Xr2_gate__xr2_inf;
    } else {
        dy[2] = cmx_div((rapid_time_dependent_potassium_current_Xr2_gate__xr2_inf - y[2]), rapid_time_dependent_potassium_current_Xr2_gate__tau_xr2);
    }
    if (RL) {
        g_k[3] = cmx_div(R(-1), slow_time_dependent_potassium_current_Xs_gate__tau_xs);
        g_inf[3] = slow_time_dependent_potassium_current_Xs_gate__xs_inf;
    } else {
        dy[3] = cmx_div((slow_time_dependent_potassium_current_Xs_gate__xs_inf - y[3]), slow_time_dependent_potassium_current_Xs_gate__tau_xs);
    }
    if (RL) {
        g_k[4] = cmx_div(R(-1), fast_sodium_current_m_gate__tau_m);
        g_inf[4] = fast_sodium_current_m_gate__m_inf;
    } else {
        dy[4] = cmx_div((fast_sodium_current_m_gate__m_inf - y[4]), fast_sodium_current_m_gate__tau_m);
    }
    if (RL) {
        g_k[5] = cmx_div(R(-1), fast_sodium_current_h_gate__tau_h);
        g_inf[5] = fast_sodium_current_h_gate__h_inf;
    } else {
        dy[5] = cmx_div((fast_sodium_current_h_gate__h_inf - y[5]), fast_sodium_current_h_gate__tau_h);
    }
    if (RL) {
        g_k[6] = cmx_div(R(-1), fast_sodium_current_j_gate__tau_j);
        g_inf[6] = fast_sodium_current_j_gate__j_inf;
    } else {
        dy[6] = cmx_div((fast_sodium_current_j_gate__j_inf - y[6]), fast_sodium_current_j_gate__tau_j);
    }
    if (RL) {
        g_k[7] = cmx_div(R(-1), L_type_Ca_current_d_gate__tau_d);
        g_inf[7] = L_type_Ca_current_d_gate__d_inf;
    } else {
        dy[7] = cmx_div((L_type_Ca_current_d_gate__d_inf - y[7]), L_type_Ca_current_d_gate__tau_d);
    }
    if (RL) {
        g_k[8] = cmx_div(R(-1), L_type_Ca_current_f_gate__tau_f);
        g_inf[8] = L_type_Ca_current_f_gate__f_inf;
    } else {
        dy[8] = cmx_div((L_type_Ca_current_f_gate__f_inf - y[8]), L_type_Ca_current_f_gate__tau_f);
    }
    if (RL) {
        g_k[9] = cmx_div(R(-1), L_type_Ca_current_f2_gate__tau_f2);
        g_inf[9] = L_type_Ca_current_f2_gate__f2_inf;
    } else {
        dy[9] = cmx_div((L_type_Ca_current_f2_gate__f2_inf - y[9]), L_type_Ca_current_f2_gate__tau_f2);
    }
    dy[10] = cmx_div((L_type_Ca_current_fCass_gate__fCass_inf - y[10]), L_type_Ca_current_fCass_gate__tau_fCass);
    if (RL) {
        g_k[11] = cmx_div(R(-1), transient_outward_current_s_gate__tau_s);
        g_inf[11] = transient_outward_current_s_gate__s_inf;
    } else {
        dy[11] = cmx_div((transient_outward_current_s_gate__s_inf - y[11]), transient_outward_current_s_gate__tau_s);
    }
    if (RL) {
        g_k[12] = cmx_div(R(-1), transient_outward_current_r_gate__tau_r);
        g_inf[12] = transient_outward_current_r_gate__r_inf;
    } else {
        dy[12] = cmx_div((transient_outward_current_r_gate__r_inf - y[12]), transient_outward_current_r_gate__tau_r);
    }
    dy[13] = (calcium_dynamics__Ca_i_bufc * (((((calcium_dynamics__i_leak - calcium_dynamics__i_up) * calcium_dynamics__V_sr) * (R(1) / membrane__V_c)) + calcium_dynamics__i_xfer) - ((((calcium_background_current__i_b_Ca + calcium_pump_current__i_p_Ca) - (R(2.0) * sodium_calcium_exchanger_current__i_NaCa)) * membrane__Cm) * (R(1) / (R(2.0) * membrane__V_c * membrane__F)))));
    dy[14] = (calcium_dynamics__Ca_sr_bufsr * (calcium_dynamics__i_up - (calcium_dynamics__i_rel + calcium_dynamics__i_leak)));
    dy[15] = (calcium_dynamics__Ca_ss_bufss * (((((-L_type_Ca_current__i_CaL) * membrane__Cm) * (R(1) / (R(2.0) * calcium_dynamics__V_ss * membrane__F))) + ((calcium_dynamics__i_rel * calcium_dynamics__V_sr) * (R(1) / calcium_dynamics__V_ss))) - ((calcium_dynamics__i_xfer * membrane__V_c) * (R(1) / calcium_dynamics__V_ss))));
    dy[16] = (((-calcium_dynamics__k2) * y[15] * y[16]) + (calcium_dynamics__k4 * (R(1.0) - y[16])));
    dy[17] = (((-(fast_sodium_current__i_Na + sodium_background_current__i_b_Na + (R(3.0) * sodium_potassium_pump_current__i_NaK) + (R(3.0) * sodium_calcium_exchanger_current__i_NaCa))) * (R(1) / (membrane__V_c * membrane__F))) * membrane__Cm);
    dy[18] = (((-((inward_rectifier_potassium_current__i_K1 + transient_outward_current__i_to + rapid_time_dependent_potassium_current__i_Kr + slow_time_dependent_potassium_current__i_Ks + potassium_pump_current__i_p_K + i_stim) - (R(2.0) * sodium_potassium_pump_current__i_NaK))) * (R(1) / (membrane__V_c * membrane__F))) * membrane__Cm);
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
    const R reversal_potentials__P_kna = R(0.03);
    const R rapid_time_dependent_potassium_current__ScaleFactorGkr = R(1.0);
    const R slow_time_dependent_potassium_current__ScaleFactorGks = R(1.0);
    const R sodium_background_current__g_bna = R(0.00029);
    const R calcium_background_current__g_bca = R(0.000592);
    const R transient_outward_current__ScaleFactorIto = R(1.0);
    const R sodium_potassium_pump_current__P_NaK = R(2.724);
    const R sodium_potassium_pump_current__K_mk = R(1.0);
    const R sodium_potassium_pump_current__K_mNa = R(40.0);
    const R sodium_calcium_exchanger_current__K_NaCa = R(1000.0);
    const R sodium_calcium_exchanger_current__K_sat = R(0.1);
    const R sodium_calcium_exchanger_current__alpha = R(2.5);
    const R sodium_calcium_exchanger_current__gamma = R(0.35);
    const R sodium_calcium_exchanger_current__Km_Ca = R(1.38);
    const R sodium_calcium_exchanger_current__Km_Nai = R(87.5);
    const R calcium_pump_current__g_pCa = R(0.1238);
    const R calcium_pump_current__K_pCa = R(0.0005);
    const R potassium_pump_current__g_pK = R(0.0146);
    const R calcium_dynamics__Ca_o = R(2.0);
    const R sodium_dynamics__Na_o = R(140.0);
    const R potassium_dynamics__K_o = R(5.4);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_CaL = cmx_remove_singularity(V, R(15.00000000000003), R(1.335688032984785e-06), [&](const R V) { return cmx_div((((p[1] * y[7] * y[8] * y[9] * y[10] * R(4.0) * (V - R(15.0)) * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((R(0.25) * y[15] * cmx_exp_v(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801137730230755, -1.620170659534613)) - calcium_dynamics__Ca_o)), cmx_expm1(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(calcium_dynamics__Ca_o, y[13])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_pump_current__i_p_Ca = cmx_div((calcium_pump_current__g_pCa * y[13]), (y[13] + calcium_pump_current__K_pCa));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(sodium_dynamics__Na_o, y[17])));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * y[6] * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(potassium_dynamics__K_o, y[18])));
    const R inward_rectifier_potassium_current__alpha_K1 = cmx_div(R(0.1), (R(1.0) + cmx_exp_c((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0))), (((V - reversal_potentials__E_K) - R(200.0))), 0.0865617024533378)));
    const R inward_rectifier_potassium_current__beta_K1 = cmx_div(((R(3.0) * cmx_exp_c((R(0.0002) * ((V - reversal_potentials__E_K) + R(100.0))), (((V - reversal_potentials__E_K) + R(100.0))), 0.0002885390081777927)) + cmx_exp_c((R(0.1) * ((V - reversal_potentials__E_K) - R(10.0))), (((V - reversal_potentials__E_K) - R(10.0))), 0.14426950408889636)), (R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817)));
    const R inward_rectifier_potassium_current__xK1_inf = cmx_div(inward_rectifier_potassium_current__alpha_K1, (inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1));
    const R inward_rectifier_potassium_current__i_K1 = (p[4] * inward_rectifier_potassium_current__xK1_inf * (V - reversal_potentials__E_K));
    const R potassium_pump_current__i_p_K = cmx_div((potassium_pump_current__g_pK * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(5.98))), V, -0.24125335131922465, 6.0313337829806155)));
    const R rapid_time_dependent_potassium_current__i_Kr = (rapid_time_dependent_potassium_current__ScaleFactorGkr * p[2] * sqrt((potassium_dynamics__K_o * (R(1) / R(5.4)))) * y[1] * y[2] * (V - reversal_potentials__E_K));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)), (y[18] + (reversal_potentials__P_kna * y[17])))));
    const R slow_time_dependent_potassium_current__i_Ks = (slow_time_dependent_potassium_current__ScaleFactorGks * p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger_current__i_NaCa = cmx_div((sodium_calcium_exchanger_current__K_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger_current__gamma * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.01890199102790382, 0.0) * cmx_ipow<3>(y[17]) * calcium_dynamics__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0) * cmx_ipow<3>(sodium_dynamics__Na_o) * y[13] * sodium_calcium_exchanger_current__alpha))), ((cmx_ipow<3>(sodium_calcium_exchanger_current__Km_Nai) + cmx_ipow<3>(sodium_dynamics__Na_o)) * (sodium_calcium_exchanger_current__Km_Ca + calcium_dynamics__Ca_o) * (R(1.0) + (sodium_calcium_exchanger_current__K_sat * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0)))));
    const R sodium_potassium_pump_current__i_NaK = cmx_div(cmx_div((((sodium_potassium_pump_current__P_NaK * potassium_dynamics__K_o) * (R(1) / (potassium_dynamics__K_o + sodium_potassium_pump_current__K_mk))) * y[17]), (y[17] + sodium_potassium_pump_current__K_mNa)), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400568865115377, -0.0)) + (R(0.0353) * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115376, -0.0))));
    const R transient_outward_current__i_to = (transient_outward_current__ScaleFactorIto * p[5] * y[12] * y[11] * (V - reversal_potentials__E_K));
    return (((-R(1.0)) * (R(1) / R(1.0))) * (inward_rectifier_potassium_current__i_K1 + transient_outward_current__i_to + rapid_time_dependent_potassium_current__i_Kr + slow_time_dependent_potassium_current__i_Ks + L_type_Ca_current__i_CaL + sodium_potassium_pump_current__i_NaK + fast_sodium_current__i_Na + sodium_background_current__i_b_Na + sodium_calcium_exchanger_current__i_NaCa + calcium_background_current__i_b_Ca + potassium_pump_current__i_p_K + calcium_pump_current__i_p_Ca + i_stim));
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
    const R reversal_potentials__P_kna = R(0.03);
    const R rapid_time_dependent_potassium_current__ScaleFactorGkr = R(1.0);
    const R slow_time_dependent_potassium_current__ScaleFactorGks = R(1.0);
    const R sodium_background_current__g_bna = R(0.00029);
    const R calcium_background_current__g_bca = R(0.000592);
    const R transient_outward_current__ScaleFactorIto = R(1.0);
    const R sodium_potassium_pump_current__P_NaK = R(2.724);
    const R sodium_potassium_pump_current__K_mk = R(1.0);
    const R sodium_potassium_pump_current__K_mNa = R(40.0);
    const R sodium_calcium_exchanger_current__K_NaCa = R(1000.0);
    const R sodium_calcium_exchanger_current__K_sat = R(0.1);
    const R sodium_calcium_exchanger_current__alpha = R(2.5);
    const R sodium_calcium_exchanger_current__gamma = R(0.35);
    const R sodium_calcium_exchanger_current__Km_Ca = R(1.38);
    const R sodium_calcium_exchanger_current__Km_Nai = R(87.5);
    const R calcium_pump_current__g_pCa = R(0.1238);
    const R calcium_pump_current__K_pCa = R(0.0005);
    const R potassium_pump_current__g_pK = R(0.0146);
    const R calcium_dynamics__Ca_o = R(2.0);
    const R sodium_dynamics__Na_o = R(140.0);
    const R potassium_dynamics__K_o = R(5.4);
    const R i_stim = R(0);
    (void)i_stim;
    const R L_type_Ca_current__i_CaL = cmx_remove_singularity(V, R(15.00000000000003), R(1.335688032984785e-06), [&](const R V) { return cmx_div((((p[1] * y[7] * y[8] * y[9] * y[10] * R(4.0) * (V - R(15.0)) * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((R(0.25) * y[15] * cmx_exp_v(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801137730230755, -1.620170659534613)) - calcium_dynamics__Ca_o)), cmx_expm1(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(calcium_dynamics__Ca_o, y[13])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_pump_current__i_p_Ca = cmx_div((calcium_pump_current__g_pCa * y[13]), (y[13] + calcium_pump_current__K_pCa));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(sodium_dynamics__Na_o, y[17])));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * y[6] * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(potassium_dynamics__K_o, y[18])));
    const R inward_rectifier_potassium_current__alpha_K1 = cmx_div(R(0.1), (R(1.0) + cmx_exp_c((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0))), (((V - reversal_potentials__E_K) - R(200.0))), 0.0865617024533378)));
    const R inward_rectifier_potassium_current__beta_K1 = cmx_div(((R(3.0) * cmx_exp_c((R(0.0002) * ((V - reversal_potentials__E_K) + R(100.0))), (((V - reversal_potentials__E_K) + R(100.0))), 0.0002885390081777927)) + cmx_exp_c((R(0.1) * ((V - reversal_potentials__E_K) - R(10.0))), (((V - reversal_potentials__E_K) - R(10.0))), 0.14426950408889636)), (R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817)));
    const R inward_rectifier_potassium_current__xK1_inf = cmx_div(inward_rectifier_potassium_current__alpha_K1, (inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1));
    const R inward_rectifier_potassium_current__i_K1 = (p[4] * inward_rectifier_potassium_current__xK1_inf * (V - reversal_potentials__E_K));
    const R potassium_pump_current__i_p_K = cmx_div((potassium_pump_current__g_pK * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(5.98))), V, -0.24125335131922465, 6.0313337829806155)));
    const R rapid_time_dependent_potassium_current__i_Kr = (rapid_time_dependent_potassium_current__ScaleFactorGkr * p[2] * sqrt((potassium_dynamics__K_o * (R(1) / R(5.4)))) * y[1] * y[2] * (V - reversal_potentials__E_K));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)), (y[18] + (reversal_potentials__P_kna * y[17])))));
    const R slow_time_dependent_potassium_current__i_Ks = (slow_time_dependent_potassium_current__ScaleFactorGks * p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger_current__i_NaCa = cmx_div((sodium_calcium_exchanger_current__K_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger_current__gamma * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.01890199102790382, 0.0) * cmx_ipow<3>(y[17]) * calcium_dynamics__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0) * cmx_ipow<3>(sodium_dynamics__Na_o) * y[13] * sodium_calcium_exchanger_current__alpha))), ((cmx_ipow<3>(sodium_calcium_exchanger_current__Km_Nai) + cmx_ipow<3>(sodium_dynamics__Na_o)) * (sodium_calcium_exchanger_current__Km_Ca + calcium_dynamics__Ca_o) * (R(1.0) + (sodium_calcium_exchanger_current__K_sat * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0)))));
    const R sodium_potassium_pump_current__i_NaK = cmx_div(cmx_div((((sodium_potassium_pump_current__P_NaK * potassium_dynamics__K_o) * (R(1) / (potassium_dynamics__K_o + sodium_potassium_pump_current__K_mk))) * y[17]), (y[17] + sodium_potassium_pump_current__K_mNa)), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400568865115377, -0.0)) + (R(0.0353) * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115376, -0.0))));
    const R transient_outward_current__i_to = (transient_outward_current__ScaleFactorIto * p[5] * y[12] * y[11] * (V - reversal_potentials__E_K));
    const R dv = (((-R(1.0)) * (R(1) / R(1.0))) * (inward_rectifier_potassium_current__i_K1 + transient_outward_current__i_to + rapid_time_dependent_potassium_current__i_Kr + slow_time_dependent_potassium_current__i_Ks + L_type_Ca_current__i_CaL + sodium_potassium_pump_current__i_NaK + fast_sodium_current__i_Na + sodium_background_current__i_b_Na + sodium_calcium_exchanger_current__i_NaCa + calcium_background_current__i_b_Ca + potassium_pump_current__i_p_K + calcium_pump_current__i_p_Ca + i_stim));
    return -dv * cm;
}

// Backward Euler: the states solved by Newton (all but the linear gates).
constexpr int kNumNonlinear = 7;
static constexpr int kNonlinear[7] = {10, 13, 14, 15, 16, 17, 18};

// Linear states: y <- (y + a dt) / (1 - b dt) for dy/dt = a + b y, a and b
// functions of V (ComputeOneStepExceptVoltage, closed-form part).
template <class R>
CMX_HD inline void be_linear(R* y, R V, const R* p, R cm, R dt)
{
    CMX_MATH_USING
    (void)p; (void)cm; (void)V;
    const R L_type_Ca_current_d_gate__d_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(8.0)) - V) * (R(1) / R(7.5))), V, -0.19235933878519512, -1.538874710281561)));
    const R L_type_Ca_current_d_gate__alpha_d = (cmx_div(R(1.4), (R(1.0) + cmx_exp_v((((-R(35.0)) - V) * (R(1) / R(13.0))), V, -0.11097654160684334, -3.8841789562395173))) + R(0.25));
    const R L_type_Ca_current_d_gate__beta_d = cmx_div(R(1.4), (R(1.0) + cmx_exp_v(((V + R(5.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 1.4426950408889634)));
    const R L_type_Ca_current_d_gate__gamma_d = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(50.0) - V) * (R(1) / R(20.0))), V, -0.07213475204444818, 3.6067376022224087)));
    const R L_type_Ca_current_d_gate__tau_d = ((L_type_Ca_current_d_gate__alpha_d * L_type_Ca_current_d_gate__beta_d) + L_type_Ca_current_d_gate__gamma_d);
    const R L_type_Ca_current_f2_gate__f2_inf = (cmx_div(R(0.67), (R(1.0) + cmx_exp_v(((V + R(35.0)) * (R(1) / R(7.0))), V, 0.2060992915555662, 7.213475204444817))) + R(0.33));
    const R L_type_Ca_current_f2_gate__tau_f2 = ((R(562.0) * cmx_exp_c(((-cmx_ipow<2>((V + R(27.0)))) * (R(1) / R(240.0))), (-cmx_ipow<2>((V + R(27.0)))), 0.0060112293370373476)) + cmx_div(R(31.0), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(10.0))), V, -0.14426950408889636, 3.6067376022224087))) + cmx_div(R(80.0), (R(1.0) + cmx_exp_v(((V + R(30.0)) * (R(1) / R(10.0))), V, 0.14426950408889636, 4.32808512266689))));
    const R L_type_Ca_current_f_gate__f_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(20.0)) * (R(1) / R(7.0))), V, 0.20609929155556622, 4.121985831111324)));
    const R L_type_Ca_current_f_gate__tau_f = ((R(1102.5) * cmx_exp_c(((-cmx_ipow<2>((V + R(27.0)))) * (R(1) / R(225.0))), (-cmx_ipow<2>((V + R(27.0)))), 0.006411977959506504)) + cmx_div(R(200.0), (R(1.0) + cmx_exp_v(((R(13.0) - V) * (R(1) / R(10.0))), V, -0.14426950408889636, 1.8755035531556525))) + cmx_div(R(180.0), (R(1.0) + cmx_exp_v(((V + R(30.0)) * (R(1) / R(10.0))), V, 0.14426950408889636, 4.32808512266689))) + R(20.0));
    const R fast_sodium_current_h_gate__h_inf = cmx_div(R(1.0), cmx_ipow<2>((R(1.0) + cmx_exp_v(((V + R(71.55)) * (R(1) / R(7.43))), V, 0.19417160711829928, 13.89297848931431))));
    const R fast_sodium_current_h_gate__alpha_h = (((V < (-R(40.0)))) ? ((R(0.057) * cmx_exp_v(((-(V + R(80.0))) * (R(1) / R(6.8))), V, -0.2121610354248476, -16.972882833987807))) : (R(0.0)));
    const R fast_sodium_current_h_gate__beta_h = (((V < (-R(40.0)))) ? (((R(2.7) * cmx_exp_v((R(0.079) * V), V, 0.11397290823022813, 0.0)) + (R(310000.0) * cmx_exp_v((R(0.3485) * V), V, 0.5027792217498037, 0.0)))) : (cmx_div(R(0.77), (R(0.13) * (R(1.0) + cmx_exp_v(((V + R(10.66)) * (R(1) / (-R(11.1)))), V, -0.12997252620621294, -1.3855071293582297))))));
    const R fast_sodium_current_h_gate__tau_h = cmx_div(R(1.0), (fast_sodium_current_h_gate__alpha_h + fast_sodium_current_h_gate__beta_h));
    const R fast_sodium_current_j_gate__j_inf = cmx_div(R(1.0), cmx_ipow<2>((R(1.0) + cmx_exp_v(((V + R(71.55)) * (R(1) / R(7.43))), V, 0.19417160711829928, 13.89297848931431))));
    const R fast_sodium_current_j_gate__alpha_j = (((V < (-R(40.0)))) ? (cmx_div(((((-R(25428.0)) * cmx_exp_v((R(0.2444) * V), V, 0.3525946679932627, 0.0)) - (R(6.948e-06) * cmx_exp_v(((-R(0.04391)) * V), V, -0.06334873924543438, -0.0))) * (V + R(37.78))), (R(1.0) + cmx_exp_v((R(0.311) * (V + R(79.23))), V, 0.44867815771646763, 35.54877043587573)))) : (R(0.0)));
    const R fast_sodium_current_j_gate__beta_j = (((V < (-R(40.0)))) ? (cmx_div((R(0.02424) * cmx_exp_v(((-R(0.01052)) * V), V, -0.015177151830151895, -0.0)), (R(1.0) + cmx_exp_v(((-R(0.1378)) * (V + R(40.14))), V, -0.19880337663449915, -7.979967538108797)))) : (cmx_div((R(0.6) * cmx_exp_v((R(0.057) * V), V, 0.08223361733067092, 0.0)), (R(1.0) + cmx_exp_v(((-R(0.1)) * (V + R(32.0))), V, -0.14426950408889636, -4.616624130844683)))));
    const R fast_sodium_current_j_gate__tau_j = cmx_div(R(1.0), (fast_sodium_current_j_gate__alpha_j + fast_sodium_current_j_gate__beta_j));
    const R fast_sodium_current_m_gate__m_inf = cmx_div(R(1.0), cmx_ipow<2>((R(1.0) + cmx_exp_v((((-R(56.86)) - V) * (R(1) / R(9.03))), V, -0.1597668926787335, -9.084345517712785))));
    const R fast_sodium_current_m_gate__alpha_m = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(60.0)) - V) * (R(1) / R(5.0))), V, -0.2885390081777927, -17.31234049066756)));
    const R fast_sodium_current_m_gate__beta_m = (cmx_div(R(0.1), (R(1.0) + cmx_exp_v(((V + R(35.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 10.098865286222743))) + cmx_div(R(0.1), (R(1.0) + cmx_exp_v(((V - R(50.0)) * (R(1) / R(200.0))), V, 0.007213475204444817, -0.36067376022224085))));
    const R fast_sodium_current_m_gate__tau_m = (fast_sodium_current_m_gate__alpha_m * fast_sodium_current_m_gate__beta_m);
    const R rapid_time_dependent_potassium_current_Xr1_gate__alpha_xr1 = cmx_div(R(450.0), (R(1.0) + cmx_exp_v((((-R(45.0)) - V) * (R(1) / R(10.0))), V, -0.14426950408889636, -6.4921276840003355)));
    const R rapid_time_dependent_potassium_current_Xr1_gate__beta_xr1 = cmx_div(R(6.0), (R(1.0) + cmx_exp_v(((V + R(30.0)) * (R(1) / R(11.5))), V, 0.12545174268599682, 3.7635522805799044)));
    const R rapid_time_dependent_potassium_current_Xr1_gate__tau_xr1 = (rapid_time_dependent_potassium_current_Xr1_gate__alpha_xr1 * rapid_time_dependent_potassium_current_Xr1_gate__beta_xr1);
    const R rapid_time_dependent_potassium_current_Xr1_gate__xr1_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(26.0)) - V) * (R(1) / R(7.0))), V, -0.20609929155556622, -5.358581580444722)));
    const R rapid_time_dependent_potassium_current_Xr2_gate__alpha_xr2 = cmx_div(R(3.0), (R(1.0) + cmx_exp_v((((-R(60.0)) - V) * (R(1) / R(20.0))), V, -0.07213475204444818, -4.32808512266689)));
    const R rapid_time_dependent_potassium_current_Xr2_gate__beta_xr2 = cmx_div(R(1.12), (R(1.0) + cmx_exp_v(((V - R(60.0)) * (R(1) / R(20.0))), V, 0.07213475204444818, -4.32808512266689)));
    const R rapid_time_dependent_potassium_current_Xr2_gate__tau_xr2 = (rapid_time_dependent_potassium_current_Xr2_gate__alpha_xr2 * rapid_time_dependent_potassium_current_Xr2_gate__beta_xr2);
    const R rapid_time_dependent_potassium_current_Xr2_gate__xr2_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(88.0)) * (R(1) / R(24.0))), V, 0.06011229337037347, 5.289881816592866)));
    const R slow_time_dependent_potassium_current_Xs_gate__alpha_xs = cmx_div(R(1400.0), sqrt((R(1.0) + cmx_exp_v(((R(5.0) - V) * (R(1) / R(6.0))), V, -0.2404491734814939, 1.2022458674074696))));
    const R slow_time_dependent_potassium_current_Xs_gate__beta_xs = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V - R(35.0)) * (R(1) / R(15.0))), V, 0.09617966939259756, -3.3662884287409147)));
    const R slow_time_dependent_potassium_current_Xs_gate__tau_xs = ((slow_time_dependent_potassium_current_Xs_gate__alpha_xs * slow_time_dependent_potassium_current_Xs_gate__beta_xs) + R(80.0));
    const R slow_time_dependent_potassium_current_Xs_gate__xs_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v((((-R(5.0)) - V) * (R(1) / R(14.0))), V, -0.10304964577778311, -0.5152482288889155)));
    const R transient_outward_current_r_gate__r_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((R(20.0) - V) * (R(1) / R(6.0))), V, -0.2404491734814939, 4.8089834696298785)));
    const R transient_outward_current_r_gate__tau_r = ((R(9.5) * cmx_exp_c(((-cmx_ipow<2>((V + R(40.0)))) * (R(1) / R(1800.0))), (-cmx_ipow<2>((V + R(40.0)))), 0.000801497244938313)) + R(0.8));
    const R transient_outward_current_s_gate__s_inf = cmx_div(R(1.0), (R(1.0) + cmx_exp_v(((V + R(20.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, 5.7707801635558535)));
    const R transient_outward_current_s_gate__tau_s = ((R(85.0) * cmx_exp_c(((-cmx_ipow<2>((V + R(45.0)))) * (R(1) / R(320.0))), (-cmx_ipow<2>((V + R(45.0)))), 0.004508422002778011)) + cmx_div(R(5.0), (R(1.0) + cmx_exp_v(((V - R(20.0)) * (R(1) / R(5.0))), V, 0.2885390081777927, -5.7707801635558535))) + R(3.0));
    y[1] = (y[1] + ((rapid_time_dependent_potassium_current_Xr1_gate__xr1_inf / rapid_time_dependent_potassium_current_Xr1_gate__tau_xr1)) * dt) / (R(1) - (((-R(1)) / rapid_time_dependent_potassium_current_Xr1_gate__tau_xr1)) * dt);
    y[2] = (y[2] + ((rapid_time_dependent_potassium_current_Xr2_gate__xr2_inf / rapid_time_dependent_potassium_current_Xr2_gate__tau_xr2)) * dt) / (R(1) - (((-R(1)) / rapid_time_dependent_potassium_current_Xr2_gate__tau_xr2)) * dt);
    y[3] = (y[3] + ((slow_time_dependent_potassium_current_Xs_gate__xs_inf / slow_time_dependent_potassium_current_Xs_gate__tau_xs)) * dt) / (R(1) - (((-R(1)) / slow_time_dependent_potassium_current_Xs_gate__tau_xs)) * dt);
    y[4] = (y[4] + ((fast_sodium_current_m_gate__m_inf / fast_sodium_current_m_gate__tau_m)) * dt) / (R(1) - (((-R(1)) / fast_sodium_current_m_gate__tau_m)) * dt);
    y[5] = (y[5] + ((fast_sodium_current_h_gate__h_inf / fast_sodium_current_h_gate__tau_h)) * dt) / (R(1) - (((-R(1)) / fast_sodium_current_h_gate__tau_h)) * dt);
    y[6] = (y[6] + ((fast_sodium_current_j_gate__j_inf / fast_sodium_current_j_gate__tau_j)) * dt) / (R(1) - (((-R(1)) / fast_sodium_current_j_gate__tau_j)) * dt);
    y[7] = (y[7] + ((L_type_Ca_current_d_gate__d_inf / L_type_Ca_current_d_gate__tau_d)) * dt) / (R(1) - (((-R(1)) / L_type_Ca_current_d_gate__tau_d)) * dt);
    y[8] = (y[8] + ((L_type_Ca_current_f_gate__f_inf / L_type_Ca_current_f_gate__tau_f)) * dt) / (R(1) - (((-R(1)) / L_type_Ca_current_f_gate__tau_f)) * dt);
    y[9] = (y[9] + ((L_type_Ca_current_f2_gate__f2_inf / L_type_Ca_current_f2_gate__tau_f2)) * dt) / (R(1) - (((-R(1)) / L_type_Ca_current_f2_gate__tau_f2)) * dt);
    y[11] = (y[11] + ((transient_outward_current_s_gate__s_inf / transient_outward_current_s_gate__tau_s)) * dt) / (R(1) - (((-R(1)) / transient_outward_current_s_gate__tau_s)) * dt);
    y[12] = (y[12] + ((transient_outward_current_r_gate__r_inf / transient_outward_current_r_gate__tau_r)) * dt) / (R(1) - (((-R(1)) / transient_outward_current_r_gate__tau_r)) * dt);
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
    const R membrane__Cm = R(0.185);
    const R membrane__V_c = R(0.016404);
    const R reversal_potentials__P_kna = R(0.03);
    const R rapid_time_dependent_potassium_current__ScaleFactorGkr = R(1.0);
    const R slow_time_dependent_potassium_current__ScaleFactorGks = R(1.0);
    const R sodium_background_current__g_bna = R(0.00029);
    const R calcium_background_current__g_bca = R(0.000592);
    const R transient_outward_current__ScaleFactorIto = R(1.0);
    const R sodium_potassium_pump_current__P_NaK = R(2.724);
    const R sodium_potassium_pump_current__K_mk = R(1.0);
    const R sodium_potassium_pump_current__K_mNa = R(40.0);
    const R sodium_calcium_exchanger_current__K_NaCa = R(1000.0);
    const R sodium_calcium_exchanger_current__K_sat = R(0.1);
    const R sodium_calcium_exchanger_current__alpha = R(2.5);
    const R sodium_calcium_exchanger_current__gamma = R(0.35);
    const R sodium_calcium_exchanger_current__Km_Ca = R(1.38);
    const R sodium_calcium_exchanger_current__Km_Nai = R(87.5);
    const R calcium_pump_current__g_pCa = R(0.1238);
    const R calcium_pump_current__K_pCa = R(0.0005);
    const R potassium_pump_current__g_pK = R(0.0146);
    const R calcium_dynamics__Ca_o = R(2.0);
    const R calcium_dynamics__k1_prime = R(0.15);
    const R calcium_dynamics__k2_prime = R(0.045);
    const R calcium_dynamics__k3 = R(0.06);
    const R calcium_dynamics__k4 = R(0.005);
    const R calcium_dynamics__EC = R(1.5);
    const R calcium_dynamics__max_sr = R(2.5);
    const R calcium_dynamics__min_sr = R(1.0);
    const R calcium_dynamics__V_rel = R(0.102);
    const R calcium_dynamics__V_xfer = R(0.0038);
    const R calcium_dynamics__K_up = R(0.00025);
    const R calcium_dynamics__V_leak = R(0.00036);
    const R calcium_dynamics__Vmax_up = R(0.006375);
    const R calcium_dynamics__Buf_c = R(0.2);
    const R calcium_dynamics__K_buf_c = R(0.001);
    const R calcium_dynamics__Buf_sr = R(10.0);
    const R calcium_dynamics__K_buf_sr = R(0.3);
    const R calcium_dynamics__Buf_ss = R(0.4);
    const R calcium_dynamics__K_buf_ss = R(0.00025);
    const R calcium_dynamics__V_sr = R(0.001094);
    const R calcium_dynamics__V_ss = R(5.468e-05);
    const R sodium_dynamics__Na_o = R(140.0);
    const R potassium_dynamics__K_o = R(5.4);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_CaL = cmx_remove_singularity(V, R(15.00000000000003), R(1.335688032984785e-06), [&](const R V) { return cmx_div((((p[1] * y[7] * y[8] * y[9] * g[0] * R(4.0) * (V - R(15.0)) * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((R(0.25) * g[3] * cmx_exp_v(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801137730230755, -1.620170659534613)) - calcium_dynamics__Ca_o)), cmx_expm1(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current_fCass_gate__fCass_inf = (cmx_div(R(0.6), (R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05)))))) + R(0.4));
    const R L_type_Ca_current_fCass_gate__tau_fCass = (cmx_div(R(80.0), (R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05)))))) + R(2.0));
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(calcium_dynamics__Ca_o, g[1])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_dynamics__Ca_i_bufc = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_c * calcium_dynamics__K_buf_c), cmx_ipow<2>((g[1] + calcium_dynamics__K_buf_c)))));
    const R calcium_dynamics__Ca_sr_bufsr = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_sr * calcium_dynamics__K_buf_sr), cmx_ipow<2>((g[2] + calcium_dynamics__K_buf_sr)))));
    const R calcium_dynamics__Ca_ss_bufss = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_ss * calcium_dynamics__K_buf_ss), cmx_ipow<2>((g[3] + calcium_dynamics__K_buf_ss)))));
    const R calcium_dynamics__i_leak = (calcium_dynamics__V_leak * (g[2] - g[1]));
    const R calcium_dynamics__kcasr = (calcium_dynamics__max_sr - cmx_div((calcium_dynamics__max_sr - calcium_dynamics__min_sr), (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__EC, g[2])))));
    const R calcium_dynamics__k1 = cmx_div(calcium_dynamics__k1_prime, calcium_dynamics__kcasr);
    const R calcium_dynamics__O = cmx_div((calcium_dynamics__k1 * cmx_ipow<2>(g[3]) * g[4]), (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3]))));
    const R calcium_dynamics__i_rel = (calcium_dynamics__V_rel * calcium_dynamics__O * (g[2] - g[3]));
    const R calcium_dynamics__i_up = cmx_div(calcium_dynamics__Vmax_up, (R(1.0) + cmx_div(cmx_ipow<2>(calcium_dynamics__K_up), cmx_ipow<2>(g[1]))));
    const R calcium_dynamics__i_xfer = (calcium_dynamics__V_xfer * (g[3] - g[1]));
    const R calcium_dynamics__k2 = (calcium_dynamics__k2_prime * calcium_dynamics__kcasr);
    const R calcium_pump_current__i_p_Ca = cmx_div((calcium_pump_current__g_pCa * g[1]), (g[1] + calcium_pump_current__K_pCa));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(sodium_dynamics__Na_o, g[5])));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * y[6] * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(potassium_dynamics__K_o, g[6])));
    const R inward_rectifier_potassium_current__alpha_K1 = cmx_div(R(0.1), (R(1.0) + cmx_exp_c((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0))), (((V - reversal_potentials__E_K) - R(200.0))), 0.0865617024533378)));
    const R inward_rectifier_potassium_current__beta_K1 = cmx_div(((R(3.0) * cmx_exp_c((R(0.0002) * ((V - reversal_potentials__E_K) + R(100.0))), (((V - reversal_potentials__E_K) + R(100.0))), 0.0002885390081777927)) + cmx_exp_c((R(0.1) * ((V - reversal_potentials__E_K) - R(10.0))), (((V - reversal_potentials__E_K) - R(10.0))), 0.14426950408889636)), (R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817)));
    const R inward_rectifier_potassium_current__xK1_inf = cmx_div(inward_rectifier_potassium_current__alpha_K1, (inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1));
    const R inward_rectifier_potassium_current__i_K1 = (p[4] * inward_rectifier_potassium_current__xK1_inf * (V - reversal_potentials__E_K));
    const R potassium_pump_current__i_p_K = cmx_div((potassium_pump_current__g_pK * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(5.98))), V, -0.24125335131922465, 6.0313337829806155)));
    const R rapid_time_dependent_potassium_current__i_Kr = (rapid_time_dependent_potassium_current__ScaleFactorGkr * p[2] * sqrt((potassium_dynamics__K_o * (R(1) / R(5.4)))) * y[1] * y[2] * (V - reversal_potentials__E_K));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)), (g[6] + (reversal_potentials__P_kna * g[5])))));
    const R slow_time_dependent_potassium_current__i_Ks = (slow_time_dependent_potassium_current__ScaleFactorGks * p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger_current__i_NaCa = cmx_div((sodium_calcium_exchanger_current__K_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger_current__gamma * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.01890199102790382, 0.0) * cmx_ipow<3>(g[5]) * calcium_dynamics__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0) * cmx_ipow<3>(sodium_dynamics__Na_o) * g[1] * sodium_calcium_exchanger_current__alpha))), ((cmx_ipow<3>(sodium_calcium_exchanger_current__Km_Nai) + cmx_ipow<3>(sodium_dynamics__Na_o)) * (sodium_calcium_exchanger_current__Km_Ca + calcium_dynamics__Ca_o) * (R(1.0) + (sodium_calcium_exchanger_current__K_sat * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0)))));
    const R sodium_potassium_pump_current__i_NaK = cmx_div(cmx_div((((sodium_potassium_pump_current__P_NaK * potassium_dynamics__K_o) * (R(1) / (potassium_dynamics__K_o + sodium_potassium_pump_current__K_mk))) * g[5]), (g[5] + sodium_potassium_pump_current__K_mNa)), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400568865115377, -0.0)) + (R(0.0353) * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115376, -0.0))));
    const R transient_outward_current__i_to = (transient_outward_current__ScaleFactorIto * p[5] * y[12] * y[11] * (V - reversal_potentials__E_K));
    res[0] = g[0] - y[10] - dt * cmx_div((L_type_Ca_current_fCass_gate__fCass_inf - g[0]), L_type_Ca_current_fCass_gate__tau_fCass);
    res[1] = g[1] - y[13] - dt * (calcium_dynamics__Ca_i_bufc * (((((calcium_dynamics__i_leak - calcium_dynamics__i_up) * calcium_dynamics__V_sr) * (R(1) / membrane__V_c)) + calcium_dynamics__i_xfer) - ((((calcium_background_current__i_b_Ca + calcium_pump_current__i_p_Ca) - (R(2.0) * sodium_calcium_exchanger_current__i_NaCa)) * membrane__Cm) * (R(1) / (R(2.0) * membrane__V_c * membrane__F)))));
    res[2] = g[2] - y[14] - dt * (calcium_dynamics__Ca_sr_bufsr * (calcium_dynamics__i_up - (calcium_dynamics__i_rel + calcium_dynamics__i_leak)));
    res[3] = g[3] - y[15] - dt * (calcium_dynamics__Ca_ss_bufss * (((((-L_type_Ca_current__i_CaL) * membrane__Cm) * (R(1) / (R(2.0) * calcium_dynamics__V_ss * membrane__F))) + ((calcium_dynamics__i_rel * calcium_dynamics__V_sr) * (R(1) / calcium_dynamics__V_ss))) - ((calcium_dynamics__i_xfer * membrane__V_c) * (R(1) / calcium_dynamics__V_ss))));
    res[4] = g[4] - y[16] - dt * (((-calcium_dynamics__k2) * g[3] * g[4]) + (calcium_dynamics__k4 * (R(1.0) - g[4])));
    res[5] = g[5] - y[17] - dt * (((-(fast_sodium_current__i_Na + sodium_background_current__i_b_Na + (R(3.0) * sodium_potassium_pump_current__i_NaK) + (R(3.0) * sodium_calcium_exchanger_current__i_NaCa))) * (R(1) / (membrane__V_c * membrane__F))) * membrane__Cm);
    res[6] = g[6] - y[18] - dt * (((-((inward_rectifier_potassium_current__i_K1 + transient_outward_current__i_to + rapid_time_dependent_potassium_current__i_Kr + slow_time_dependent_potassium_current__i_Ks + potassium_pump_current__i_p_K + i_stim) - (R(2.0) * sodium_potassium_pump_current__i_NaK))) * (R(1) / (membrane__V_c * membrane__F))) * membrane__Cm);
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
    const R membrane__Cm = R(0.185);
    const R membrane__V_c = R(0.016404);
    const R reversal_potentials__P_kna = R(0.03);
    const R rapid_time_dependent_potassium_current__ScaleFactorGkr = R(1.0);
    const R slow_time_dependent_potassium_current__ScaleFactorGks = R(1.0);
    const R sodium_background_current__g_bna = R(0.00029);
    const R calcium_background_current__g_bca = R(0.000592);
    const R transient_outward_current__ScaleFactorIto = R(1.0);
    const R sodium_potassium_pump_current__P_NaK = R(2.724);
    const R sodium_potassium_pump_current__K_mk = R(1.0);
    const R sodium_potassium_pump_current__K_mNa = R(40.0);
    const R sodium_calcium_exchanger_current__K_NaCa = R(1000.0);
    const R sodium_calcium_exchanger_current__K_sat = R(0.1);
    const R sodium_calcium_exchanger_current__alpha = R(2.5);
    const R sodium_calcium_exchanger_current__gamma = R(0.35);
    const R sodium_calcium_exchanger_current__Km_Ca = R(1.38);
    const R sodium_calcium_exchanger_current__Km_Nai = R(87.5);
    const R calcium_pump_current__g_pCa = R(0.1238);
    const R calcium_pump_current__K_pCa = R(0.0005);
    const R potassium_pump_current__g_pK = R(0.0146);
    const R calcium_dynamics__Ca_o = R(2.0);
    const R calcium_dynamics__k1_prime = R(0.15);
    const R calcium_dynamics__k2_prime = R(0.045);
    const R calcium_dynamics__k3 = R(0.06);
    const R calcium_dynamics__k4 = R(0.005);
    const R calcium_dynamics__EC = R(1.5);
    const R calcium_dynamics__max_sr = R(2.5);
    const R calcium_dynamics__min_sr = R(1.0);
    const R calcium_dynamics__V_rel = R(0.102);
    const R calcium_dynamics__V_xfer = R(0.0038);
    const R calcium_dynamics__K_up = R(0.00025);
    const R calcium_dynamics__V_leak = R(0.00036);
    const R calcium_dynamics__Vmax_up = R(0.006375);
    const R calcium_dynamics__Buf_c = R(0.2);
    const R calcium_dynamics__K_buf_c = R(0.001);
    const R calcium_dynamics__Buf_sr = R(10.0);
    const R calcium_dynamics__K_buf_sr = R(0.3);
    const R calcium_dynamics__Buf_ss = R(0.4);
    const R calcium_dynamics__K_buf_ss = R(0.00025);
    const R calcium_dynamics__V_sr = R(0.001094);
    const R calcium_dynamics__V_ss = R(5.468e-05);
    const R sodium_dynamics__Na_o = R(140.0);
    const R potassium_dynamics__K_o = R(5.4);
    const R i_stim = i_stim_area / cm;
    (void)i_stim;
    const R L_type_Ca_current__i_CaL = cmx_remove_singularity(V, R(15.00000000000003), R(1.335688032984785e-06), [&](const R V) { return cmx_div((((p[1] * y[7] * y[8] * y[9] * g[0] * R(4.0) * (V - R(15.0)) * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))) * ((R(0.25) * g[3] * cmx_exp_v(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801137730230755, -1.620170659534613)) - calcium_dynamics__Ca_o)), cmx_expm1(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R L_type_Ca_current_fCass_gate__fCass_inf = (cmx_div(R(0.6), (R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05)))))) + R(0.4));
    const R L_type_Ca_current_fCass_gate__tau_fCass = (cmx_div(R(80.0), (R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05)))))) + R(2.0));
    const R reversal_potentials__E_Ca = (((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(calcium_dynamics__Ca_o, g[1])));
    const R calcium_background_current__i_b_Ca = (calcium_background_current__g_bca * (V - reversal_potentials__E_Ca));
    const R calcium_dynamics__Ca_i_bufc = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_c * calcium_dynamics__K_buf_c), cmx_ipow<2>((g[1] + calcium_dynamics__K_buf_c)))));
    const R calcium_dynamics__Ca_sr_bufsr = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_sr * calcium_dynamics__K_buf_sr), cmx_ipow<2>((g[2] + calcium_dynamics__K_buf_sr)))));
    const R calcium_dynamics__Ca_ss_bufss = cmx_div(R(1.0), (R(1.0) + cmx_div((calcium_dynamics__Buf_ss * calcium_dynamics__K_buf_ss), cmx_ipow<2>((g[3] + calcium_dynamics__K_buf_ss)))));
    const R calcium_dynamics__i_leak = (calcium_dynamics__V_leak * (g[2] - g[1]));
    const R calcium_dynamics__kcasr = (calcium_dynamics__max_sr - cmx_div((calcium_dynamics__max_sr - calcium_dynamics__min_sr), (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__EC, g[2])))));
    const R calcium_dynamics__k1 = cmx_div(calcium_dynamics__k1_prime, calcium_dynamics__kcasr);
    const R calcium_dynamics__O = cmx_div((calcium_dynamics__k1 * cmx_ipow<2>(g[3]) * g[4]), (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3]))));
    const R calcium_dynamics__i_rel = (calcium_dynamics__V_rel * calcium_dynamics__O * (g[2] - g[3]));
    const R calcium_dynamics__i_up = cmx_div(calcium_dynamics__Vmax_up, (R(1.0) + cmx_div(cmx_ipow<2>(calcium_dynamics__K_up), cmx_ipow<2>(g[1]))));
    const R calcium_dynamics__i_xfer = (calcium_dynamics__V_xfer * (g[3] - g[1]));
    const R calcium_dynamics__k2 = (calcium_dynamics__k2_prime * calcium_dynamics__kcasr);
    const R calcium_pump_current__i_p_Ca = cmx_div((calcium_pump_current__g_pCa * g[1]), (g[1] + calcium_pump_current__K_pCa));
    const R reversal_potentials__E_Na = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(sodium_dynamics__Na_o, g[5])));
    const R fast_sodium_current__i_Na = (p[0] * cmx_ipow<3>(y[4]) * y[5] * y[6] * (V - reversal_potentials__E_Na));
    const R reversal_potentials__E_K = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div(potassium_dynamics__K_o, g[6])));
    const R inward_rectifier_potassium_current__alpha_K1 = cmx_div(R(0.1), (R(1.0) + cmx_exp_c((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0))), (((V - reversal_potentials__E_K) - R(200.0))), 0.0865617024533378)));
    const R inward_rectifier_potassium_current__beta_K1 = cmx_div(((R(3.0) * cmx_exp_c((R(0.0002) * ((V - reversal_potentials__E_K) + R(100.0))), (((V - reversal_potentials__E_K) + R(100.0))), 0.0002885390081777927)) + cmx_exp_c((R(0.1) * ((V - reversal_potentials__E_K) - R(10.0))), (((V - reversal_potentials__E_K) - R(10.0))), 0.14426950408889636)), (R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817)));
    const R inward_rectifier_potassium_current__xK1_inf = cmx_div(inward_rectifier_potassium_current__alpha_K1, (inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1));
    const R inward_rectifier_potassium_current__i_K1 = (p[4] * inward_rectifier_potassium_current__xK1_inf * (V - reversal_potentials__E_K));
    const R potassium_pump_current__i_p_K = cmx_div((potassium_pump_current__g_pK * (V - reversal_potentials__E_K)), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(5.98))), V, -0.24125335131922465, 6.0313337829806155)));
    const R rapid_time_dependent_potassium_current__i_Kr = (rapid_time_dependent_potassium_current__ScaleFactorGkr * p[2] * sqrt((potassium_dynamics__K_o * (R(1) / R(5.4)))) * y[1] * y[2] * (V - reversal_potentials__E_K));
    const R reversal_potentials__E_Ks = (((membrane__R * membrane__T) * (R(1) / membrane__F)) * cmx_log(cmx_div((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)), (g[6] + (reversal_potentials__P_kna * g[5])))));
    const R slow_time_dependent_potassium_current__i_Ks = (slow_time_dependent_potassium_current__ScaleFactorGks * p[3] * cmx_ipow<2>(y[3]) * (V - reversal_potentials__E_Ks));
    const R sodium_background_current__i_b_Na = (sodium_background_current__g_bna * (V - reversal_potentials__E_Na));
    const R sodium_calcium_exchanger_current__i_NaCa = cmx_div((sodium_calcium_exchanger_current__K_NaCa * ((cmx_exp_v(((sodium_calcium_exchanger_current__gamma * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.01890199102790382, 0.0) * cmx_ipow<3>(g[5]) * calcium_dynamics__Ca_o) - (cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0) * cmx_ipow<3>(sodium_dynamics__Na_o) * g[1] * sodium_calcium_exchanger_current__alpha))), ((cmx_ipow<3>(sodium_calcium_exchanger_current__Km_Nai) + cmx_ipow<3>(sodium_dynamics__Na_o)) * (sodium_calcium_exchanger_current__Km_Ca + calcium_dynamics__Ca_o) * (R(1.0) + (sodium_calcium_exchanger_current__K_sat * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0)))));
    const R sodium_potassium_pump_current__i_NaK = cmx_div(cmx_div((((sodium_potassium_pump_current__P_NaK * potassium_dynamics__K_o) * (R(1) / (potassium_dynamics__K_o + sodium_potassium_pump_current__K_mk))) * g[5]), (g[5] + sodium_potassium_pump_current__K_mNa)), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400568865115377, -0.0)) + (R(0.0353) * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115376, -0.0))));
    const R transient_outward_current__i_to = (transient_outward_current__ScaleFactorIto * p[5] * y[12] * y[11] * (V - reversal_potentials__E_K));
    const R d_L_type_Ca_current__i_CaL__d0 = cmx_remove_singularity(V, R(15.00000000000003), R(1.335688032984785e-06), [&](const R V) { return cmx_div(((cmx_div(((R(1) * p[1] * y[7] * y[8] * y[9] * R(4.0) * (V - R(15.0)) * cmx_ipow<2>(membrane__F))), (membrane__R * membrane__T)) * ((R(0.25) * g[3] * cmx_exp_v(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801137730230755, -1.620170659534613)) - calcium_dynamics__Ca_o))), cmx_expm1(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_L_type_Ca_current__i_CaL__d3 = cmx_remove_singularity(V, R(15.00000000000003), R(1.335688032984785e-06), [&](const R V) { return cmx_div(((((R(1) * R(0.25) * cmx_exp_v(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.10801137730230755, -1.620170659534613))) * ((p[1] * y[7] * y[8] * y[9] * g[0] * R(4.0) * (V - R(15.0)) * cmx_ipow<2>(membrane__F)) * (R(1) / (membrane__R * membrane__T))))), cmx_expm1(((R(2.0) * (V - R(15.0)) * membrane__F) * (R(1) / (membrane__R * membrane__T))))); });
    const R d_L_type_Ca_current_fCass_gate__fCass_inf__d3 = (cmx_div((-(R(0.6) * ((R(2.0) * cmx_ipow<1>((g[3] * (R(1) / R(0.05)))) * cmx_div(R(1), R(0.05)))))), ((R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05))))) * (R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05))))))));
    const R d_L_type_Ca_current_fCass_gate__tau_fCass__d3 = (cmx_div((-(R(80.0) * ((R(2.0) * cmx_ipow<1>((g[3] * (R(1) / R(0.05)))) * cmx_div(R(1), R(0.05)))))), ((R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05))))) * (R(1.0) + cmx_ipow<2>((g[3] * (R(1) / R(0.05))))))));
    const R d_reversal_potentials__E_Ca__d1 = ((cmx_div(cmx_div((-(calcium_dynamics__Ca_o * R(1))), (g[1] * g[1])), cmx_div(calcium_dynamics__Ca_o, g[1])) * ((R(0.5) * membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_calcium_background_current__i_b_Ca__d1 = (((-d_reversal_potentials__E_Ca__d1) * calcium_background_current__g_bca));
    const R d_calcium_dynamics__Ca_i_bufc__d1 = cmx_div((-(R(1.0) * (cmx_div((-((calcium_dynamics__Buf_c * calcium_dynamics__K_buf_c) * (R(2.0) * cmx_ipow<1>((g[1] + calcium_dynamics__K_buf_c)) * (R(1))))), (cmx_ipow<2>((g[1] + calcium_dynamics__K_buf_c)) * cmx_ipow<2>((g[1] + calcium_dynamics__K_buf_c))))))), ((R(1.0) + cmx_div((calcium_dynamics__Buf_c * calcium_dynamics__K_buf_c), cmx_ipow<2>((g[1] + calcium_dynamics__K_buf_c)))) * (R(1.0) + cmx_div((calcium_dynamics__Buf_c * calcium_dynamics__K_buf_c), cmx_ipow<2>((g[1] + calcium_dynamics__K_buf_c))))));
    const R d_calcium_dynamics__Ca_sr_bufsr__d2 = cmx_div((-(R(1.0) * (cmx_div((-((calcium_dynamics__Buf_sr * calcium_dynamics__K_buf_sr) * (R(2.0) * cmx_ipow<1>((g[2] + calcium_dynamics__K_buf_sr)) * (R(1))))), (cmx_ipow<2>((g[2] + calcium_dynamics__K_buf_sr)) * cmx_ipow<2>((g[2] + calcium_dynamics__K_buf_sr))))))), ((R(1.0) + cmx_div((calcium_dynamics__Buf_sr * calcium_dynamics__K_buf_sr), cmx_ipow<2>((g[2] + calcium_dynamics__K_buf_sr)))) * (R(1.0) + cmx_div((calcium_dynamics__Buf_sr * calcium_dynamics__K_buf_sr), cmx_ipow<2>((g[2] + calcium_dynamics__K_buf_sr))))));
    const R d_calcium_dynamics__Ca_ss_bufss__d3 = cmx_div((-(R(1.0) * (cmx_div((-((calcium_dynamics__Buf_ss * calcium_dynamics__K_buf_ss) * (R(2.0) * cmx_ipow<1>((g[3] + calcium_dynamics__K_buf_ss)) * (R(1))))), (cmx_ipow<2>((g[3] + calcium_dynamics__K_buf_ss)) * cmx_ipow<2>((g[3] + calcium_dynamics__K_buf_ss))))))), ((R(1.0) + cmx_div((calcium_dynamics__Buf_ss * calcium_dynamics__K_buf_ss), cmx_ipow<2>((g[3] + calcium_dynamics__K_buf_ss)))) * (R(1.0) + cmx_div((calcium_dynamics__Buf_ss * calcium_dynamics__K_buf_ss), cmx_ipow<2>((g[3] + calcium_dynamics__K_buf_ss))))));
    const R d_calcium_dynamics__i_leak__d1 = (((-R(1)) * calcium_dynamics__V_leak));
    const R d_calcium_dynamics__i_leak__d2 = ((R(1) * calcium_dynamics__V_leak));
    const R d_calcium_dynamics__kcasr__d2 = (-cmx_div((-((calcium_dynamics__max_sr - calcium_dynamics__min_sr) * ((R(2.0) * cmx_ipow<1>(cmx_div(calcium_dynamics__EC, g[2])) * cmx_div((-(calcium_dynamics__EC * R(1))), (g[2] * g[2])))))), ((R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__EC, g[2]))) * (R(1.0) + cmx_ipow<2>(cmx_div(calcium_dynamics__EC, g[2]))))));
    const R d_calcium_dynamics__k1__d2 = cmx_div((-(calcium_dynamics__k1_prime * d_calcium_dynamics__kcasr__d2)), (calcium_dynamics__kcasr * calcium_dynamics__kcasr));
    const R d_calcium_dynamics__O__d2 = cmx_div((((d_calcium_dynamics__k1__d2 * cmx_ipow<2>(g[3]) * g[4])) * (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3]))) - (calcium_dynamics__k1 * cmx_ipow<2>(g[3]) * g[4]) * (((d_calcium_dynamics__k1__d2 * cmx_ipow<2>(g[3]))))), ((calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3]))) * (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3])))));
    const R d_calcium_dynamics__O__d3 = cmx_div(((((R(2.0) * cmx_ipow<1>(g[3]) * R(1)) * calcium_dynamics__k1 * g[4])) * (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3]))) - (calcium_dynamics__k1 * cmx_ipow<2>(g[3]) * g[4]) * ((((R(2.0) * cmx_ipow<1>(g[3]) * R(1)) * calcium_dynamics__k1)))), ((calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3]))) * (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3])))));
    const R d_calcium_dynamics__O__d4 = cmx_div(((R(1) * calcium_dynamics__k1 * cmx_ipow<2>(g[3]))), (calcium_dynamics__k3 + (calcium_dynamics__k1 * cmx_ipow<2>(g[3]))));
    const R d_calcium_dynamics__i_rel__d2 = ((d_calcium_dynamics__O__d2 * calcium_dynamics__V_rel * (g[2] - g[3])) + (R(1) * calcium_dynamics__V_rel * calcium_dynamics__O));
    const R d_calcium_dynamics__i_rel__d3 = ((d_calcium_dynamics__O__d3 * calcium_dynamics__V_rel * (g[2] - g[3])) + ((-R(1)) * calcium_dynamics__V_rel * calcium_dynamics__O));
    const R d_calcium_dynamics__i_rel__d4 = ((d_calcium_dynamics__O__d4 * calcium_dynamics__V_rel * (g[2] - g[3])));
    const R d_calcium_dynamics__i_up__d1 = cmx_div((-(calcium_dynamics__Vmax_up * (cmx_div((-(cmx_ipow<2>(calcium_dynamics__K_up) * (R(2.0) * cmx_ipow<1>(g[1]) * R(1)))), (cmx_ipow<2>(g[1]) * cmx_ipow<2>(g[1])))))), ((R(1.0) + cmx_div(cmx_ipow<2>(calcium_dynamics__K_up), cmx_ipow<2>(g[1]))) * (R(1.0) + cmx_div(cmx_ipow<2>(calcium_dynamics__K_up), cmx_ipow<2>(g[1])))));
    const R d_calcium_dynamics__i_xfer__d1 = (((-R(1)) * calcium_dynamics__V_xfer));
    const R d_calcium_dynamics__i_xfer__d3 = ((R(1) * calcium_dynamics__V_xfer));
    const R d_calcium_dynamics__k2__d2 = ((d_calcium_dynamics__kcasr__d2 * calcium_dynamics__k2_prime));
    const R d_calcium_pump_current__i_p_Ca__d1 = cmx_div((((R(1) * calcium_pump_current__g_pCa)) * (g[1] + calcium_pump_current__K_pCa) - (calcium_pump_current__g_pCa * g[1]) * (R(1))), ((g[1] + calcium_pump_current__K_pCa) * (g[1] + calcium_pump_current__K_pCa)));
    const R d_reversal_potentials__E_Na__d5 = ((cmx_div(cmx_div((-(sodium_dynamics__Na_o * R(1))), (g[5] * g[5])), cmx_div(sodium_dynamics__Na_o, g[5])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_fast_sodium_current__i_Na__d5 = (((-d_reversal_potentials__E_Na__d5) * p[0] * cmx_ipow<3>(y[4]) * y[5] * y[6]));
    const R d_reversal_potentials__E_K__d6 = ((cmx_div(cmx_div((-(potassium_dynamics__K_o * R(1))), (g[6] * g[6])), cmx_div(potassium_dynamics__K_o, g[6])) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_inward_rectifier_potassium_current__alpha_K1__d6 = cmx_div((-(R(0.1) * ((cmx_exp((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0)))) * (((-d_reversal_potentials__E_K__d6) * R(0.06))))))), ((R(1.0) + cmx_exp_c((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0))), (((V - reversal_potentials__E_K) - R(200.0))), 0.0865617024533378)) * (R(1.0) + cmx_exp_c((R(0.06) * ((V - reversal_potentials__E_K) - R(200.0))), (((V - reversal_potentials__E_K) - R(200.0))), 0.0865617024533378))));
    const R d_inward_rectifier_potassium_current__beta_K1__d6 = cmx_div((((((cmx_exp((R(0.0002) * ((V - reversal_potentials__E_K) + R(100.0)))) * ((((-d_reversal_potentials__E_K__d6)) * R(0.0002)))) * R(3.0))) + (cmx_exp((R(0.1) * ((V - reversal_potentials__E_K) - R(10.0)))) * (((-d_reversal_potentials__E_K__d6) * R(0.1))))) * (R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817)) - ((R(3.0) * cmx_exp_c((R(0.0002) * ((V - reversal_potentials__E_K) + R(100.0))), (((V - reversal_potentials__E_K) + R(100.0))), 0.0002885390081777927)) + cmx_exp_c((R(0.1) * ((V - reversal_potentials__E_K) - R(10.0))), (((V - reversal_potentials__E_K) - R(10.0))), 0.14426950408889636)) * ((cmx_exp(((-R(0.5)) * (V - reversal_potentials__E_K))) * (((-d_reversal_potentials__E_K__d6) * (-R(0.5))))))), ((R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817)) * (R(1.0) + cmx_exp_c(((-R(0.5)) * (V - reversal_potentials__E_K)), ((V - reversal_potentials__E_K)), -0.7213475204444817))));
    const R d_inward_rectifier_potassium_current__xK1_inf__d6 = cmx_div((d_inward_rectifier_potassium_current__alpha_K1__d6 * (inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1) - inward_rectifier_potassium_current__alpha_K1 * (d_inward_rectifier_potassium_current__alpha_K1__d6 + d_inward_rectifier_potassium_current__beta_K1__d6)), ((inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1) * (inward_rectifier_potassium_current__alpha_K1 + inward_rectifier_potassium_current__beta_K1)));
    const R d_inward_rectifier_potassium_current__i_K1__d6 = ((d_inward_rectifier_potassium_current__xK1_inf__d6 * p[4] * (V - reversal_potentials__E_K)) + ((-d_reversal_potentials__E_K__d6) * p[4] * inward_rectifier_potassium_current__xK1_inf));
    const R d_potassium_pump_current__i_p_K__d6 = cmx_div((((-d_reversal_potentials__E_K__d6) * potassium_pump_current__g_pK)), (R(1.0) + cmx_exp_v(((R(25.0) - V) * (R(1) / R(5.98))), V, -0.24125335131922465, 6.0313337829806155)));
    const R d_rapid_time_dependent_potassium_current__i_Kr__d6 = (((-d_reversal_potentials__E_K__d6) * rapid_time_dependent_potassium_current__ScaleFactorGkr * p[2] * sqrt((potassium_dynamics__K_o * (R(1) / R(5.4)))) * y[1] * y[2]));
    const R d_reversal_potentials__E_Ks__d5 = ((cmx_div(cmx_div((-((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)) * (((R(1) * reversal_potentials__P_kna))))), ((g[6] + (reversal_potentials__P_kna * g[5])) * (g[6] + (reversal_potentials__P_kna * g[5])))), cmx_div((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)), (g[6] + (reversal_potentials__P_kna * g[5])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_reversal_potentials__E_Ks__d6 = ((cmx_div(cmx_div((-((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)) * (R(1)))), ((g[6] + (reversal_potentials__P_kna * g[5])) * (g[6] + (reversal_potentials__P_kna * g[5])))), cmx_div((potassium_dynamics__K_o + (reversal_potentials__P_kna * sodium_dynamics__Na_o)), (g[6] + (reversal_potentials__P_kna * g[5])))) * ((membrane__R * membrane__T) * (R(1) / membrane__F))));
    const R d_slow_time_dependent_potassium_current__i_Ks__d5 = (((-d_reversal_potentials__E_Ks__d5) * slow_time_dependent_potassium_current__ScaleFactorGks * p[3] * cmx_ipow<2>(y[3])));
    const R d_slow_time_dependent_potassium_current__i_Ks__d6 = (((-d_reversal_potentials__E_Ks__d6) * slow_time_dependent_potassium_current__ScaleFactorGks * p[3] * cmx_ipow<2>(y[3])));
    const R d_sodium_background_current__i_b_Na__d5 = (((-d_reversal_potentials__E_Na__d5) * sodium_background_current__g_bna));
    const R d_sodium_calcium_exchanger_current__i_NaCa__d1 = cmx_div((((-((R(1) * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0) * cmx_ipow<3>(sodium_dynamics__Na_o) * sodium_calcium_exchanger_current__alpha))) * sodium_calcium_exchanger_current__K_NaCa)), ((cmx_ipow<3>(sodium_calcium_exchanger_current__Km_Nai) + cmx_ipow<3>(sodium_dynamics__Na_o)) * (sodium_calcium_exchanger_current__Km_Ca + calcium_dynamics__Ca_o) * (R(1.0) + (sodium_calcium_exchanger_current__K_sat * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0)))));
    const R d_sodium_calcium_exchanger_current__i_NaCa__d5 = cmx_div((((((R(3.0) * cmx_ipow<2>(g[5]) * R(1)) * cmx_exp_v(((sodium_calcium_exchanger_current__gamma * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, 0.01890199102790382, 0.0) * calcium_dynamics__Ca_o)) * sodium_calcium_exchanger_current__K_NaCa)), ((cmx_ipow<3>(sodium_calcium_exchanger_current__Km_Nai) + cmx_ipow<3>(sodium_dynamics__Na_o)) * (sodium_calcium_exchanger_current__Km_Ca + calcium_dynamics__Ca_o) * (R(1.0) + (sodium_calcium_exchanger_current__K_sat * cmx_exp_v((((sodium_calcium_exchanger_current__gamma - R(1.0)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.03510369762324995, -0.0)))));
    const R d_sodium_potassium_pump_current__i_NaK__d5 = cmx_div(cmx_div((((R(1) * ((sodium_potassium_pump_current__P_NaK * potassium_dynamics__K_o) * (R(1) / (potassium_dynamics__K_o + sodium_potassium_pump_current__K_mk))))) * (g[5] + sodium_potassium_pump_current__K_mNa) - (((sodium_potassium_pump_current__P_NaK * potassium_dynamics__K_o) * (R(1) / (potassium_dynamics__K_o + sodium_potassium_pump_current__K_mk))) * g[5]) * (R(1))), ((g[5] + sodium_potassium_pump_current__K_mNa) * (g[5] + sodium_potassium_pump_current__K_mNa))), (R(1.0) + (R(0.1245) * cmx_exp_v((((-R(0.1)) * V * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.005400568865115377, -0.0)) + (R(0.0353) * cmx_exp_v((((-V) * membrane__F) * (R(1) / (membrane__R * membrane__T))), V, -0.05400568865115376, -0.0))));
    const R d_transient_outward_current__i_to__d6 = (((-d_reversal_potentials__E_K__d6) * transient_outward_current__ScaleFactorIto * p[5] * y[12] * y[11]));
    J[0] = R(1) - dt * cmx_div((-R(1)), L_type_Ca_current_fCass_gate__tau_fCass);
    J[1] = R(0);
    J[2] = R(0);
    J[3] = R(0) - dt * cmx_div((d_L_type_Ca_current_fCass_gate__fCass_inf__d3 * L_type_Ca_current_fCass_gate__tau_fCass - (L_type_Ca_current_fCass_gate__fCass_inf - g[0]) * d_L_type_Ca_current_fCass_gate__tau_fCass__d3), (L_type_Ca_current_fCass_gate__tau_fCass * L_type_Ca_current_fCass_gate__tau_fCass));
    J[4] = R(0);
    J[5] = R(0);
    J[6] = R(0);
    J[7] = R(0);
    J[8] = R(1) - dt * ((d_calcium_dynamics__Ca_i_bufc__d1 * (((((calcium_dynamics__i_leak - calcium_dynamics__i_up) * calcium_dynamics__V_sr) * (R(1) / membrane__V_c)) + calcium_dynamics__i_xfer) - ((((calcium_background_current__i_b_Ca + calcium_pump_current__i_p_Ca) - (R(2.0) * sodium_calcium_exchanger_current__i_NaCa)) * membrane__Cm) * (R(1) / (R(2.0) * membrane__V_c * membrane__F))))) + (((cmx_div((((d_calcium_dynamics__i_leak__d1 - d_calcium_dynamics__i_up__d1) * calcium_dynamics__V_sr)), membrane__V_c) + d_calcium_dynamics__i_xfer__d1) - cmx_div(((((d_calcium_background_current__i_b_Ca__d1 + d_calcium_pump_current__i_p_Ca__d1) - ((d_sodium_calcium_exchanger_current__i_NaCa__d1 * R(2.0)))) * membrane__Cm)), (R(2.0) * membrane__V_c * membrane__F))) * calcium_dynamics__Ca_i_bufc));
    J[9] = R(0) - dt * (((cmx_div(((d_calcium_dynamics__i_leak__d2 * calcium_dynamics__V_sr)), membrane__V_c)) * calcium_dynamics__Ca_i_bufc));
    J[10] = R(0) - dt * (((d_calcium_dynamics__i_xfer__d3) * calcium_dynamics__Ca_i_bufc));
    J[11] = R(0);
    J[12] = R(0) - dt * (((-cmx_div((((-((d_sodium_calcium_exchanger_current__i_NaCa__d5 * R(2.0)))) * membrane__Cm)), (R(2.0) * membrane__V_c * membrane__F))) * calcium_dynamics__Ca_i_bufc));
    J[13] = R(0);
    J[14] = R(0);
    J[15] = R(0) - dt * (((d_calcium_dynamics__i_up__d1 - (d_calcium_dynamics__i_leak__d1)) * calcium_dynamics__Ca_sr_bufsr));
    J[16] = R(1) - dt * ((d_calcium_dynamics__Ca_sr_bufsr__d2 * (calcium_dynamics__i_up - (calcium_dynamics__i_rel + calcium_dynamics__i_leak))) + ((-(d_calcium_dynamics__i_rel__d2 + d_calcium_dynamics__i_leak__d2)) * calcium_dynamics__Ca_sr_bufsr));
    J[17] = R(0) - dt * (((-(d_calcium_dynamics__i_rel__d3)) * calcium_dynamics__Ca_sr_bufsr));
    J[18] = R(0) - dt * (((-(d_calcium_dynamics__i_rel__d4)) * calcium_dynamics__Ca_sr_bufsr));
    J[19] = R(0);
    J[20] = R(0);
    J[21] = R(0) - dt * (((cmx_div((((-d_L_type_Ca_current__i_CaL__d0) * membrane__Cm)), (R(2.0) * calcium_dynamics__V_ss * membrane__F))) * calcium_dynamics__Ca_ss_bufss));
    J[22] = R(0) - dt * (((-cmx_div(((d_calcium_dynamics__i_xfer__d1 * membrane__V_c)), calcium_dynamics__V_ss)) * calcium_dynamics__Ca_ss_bufss));
    J[23] = R(0) - dt * (((cmx_div(((d_calcium_dynamics__i_rel__d2 * calcium_dynamics__V_sr)), calcium_dynamics__V_ss)) * calcium_dynamics__Ca_ss_bufss));
    J[24] = R(1) - dt * ((d_calcium_dynamics__Ca_ss_bufss__d3 * (((((-L_type_Ca_current__i_CaL) * membrane__Cm) * (R(1) / (R(2.0) * calcium_dynamics__V_ss * membrane__F))) + ((calcium_dynamics__i_rel * calcium_dynamics__V_sr) * (R(1) / calcium_dynamics__V_ss))) - ((calcium_dynamics__i_xfer * membrane__V_c) * (R(1) / calcium_dynamics__V_ss)))) + (((cmx_div((((-d_L_type_Ca_current__i_CaL__d3) * membrane__Cm)), (R(2.0) * calcium_dynamics__V_ss * membrane__F)) + cmx_div(((d_calcium_dynamics__i_rel__d3 * calcium_dynamics__V_sr)), calcium_dynamics__V_ss)) - cmx_div(((d_calcium_dynamics__i_xfer__d3 * membrane__V_c)), calcium_dynamics__V_ss)) * calcium_dynamics__Ca_ss_bufss));
    J[25] = R(0) - dt * (((cmx_div(((d_calcium_dynamics__i_rel__d4 * calcium_dynamics__V_sr)), calcium_dynamics__V_ss)) * calcium_dynamics__Ca_ss_bufss));
    J[26] = R(0);
    J[27] = R(0);
    J[28] = R(0);
    J[29] = R(0);
    J[30] = R(0) - dt * ((((-d_calcium_dynamics__k2__d2) * g[3] * g[4])));
    J[31] = R(0) - dt * (((R(1) * (-calcium_dynamics__k2) * g[4])));
    J[32] = R(1) - dt * (((R(1) * (-calcium_dynamics__k2) * g[3])) + (((-R(1)) * calcium_dynamics__k4)));
    J[33] = R(0);
    J[34] = R(0);
    J[35] = R(0);
    J[36] = R(0) - dt * ((cmx_div((-(((d_sodium_calcium_exchanger_current__i_NaCa__d1 * R(3.0))))), (membrane__V_c * membrane__F)) * membrane__Cm));
    J[37] = R(0);
    J[38] = R(0);
    J[39] = R(0);
    J[40] = R(1) - dt * ((cmx_div((-(d_fast_sodium_current__i_Na__d5 + d_sodium_background_current__i_b_Na__d5 + ((d_sodium_potassium_pump_current__i_NaK__d5 * R(3.0))) + ((d_sodium_calcium_exchanger_current__i_NaCa__d5 * R(3.0))))), (membrane__V_c * membrane__F)) * membrane__Cm));
    J[41] = R(0);
    J[42] = R(0);
    J[43] = R(0);
    J[44] = R(0);
    J[45] = R(0);
    J[46] = R(0);
    J[47] = R(0) - dt * ((cmx_div((-((d_slow_time_dependent_potassium_current__i_Ks__d5) - ((d_sodium_potassium_pump_current__i_NaK__d5 * R(2.0))))), (membrane__V_c * membrane__F)) * membrane__Cm));
    J[48] = R(1) - dt * ((cmx_div((-(d_inward_rectifier_potassium_current__i_K1__d6 + d_transient_outward_current__i_to__d6 + d_rapid_time_dependent_potassium_current__i_Kr__d6 + d_slow_time_dependent_potassium_current__i_Ks__d6 + d_potassium_pump_current__i_p_K__d6)), (membrane__V_c * membrane__F)) * membrane__Cm));
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

}}  // namespace cmx::tt06_epi
