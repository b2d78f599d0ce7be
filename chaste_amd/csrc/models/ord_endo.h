// O'Hara-Rudy 2011 human ventricular model, endocardial cell (O'Hara T,
// Virag L, Varro A, Rudy Y, "Simulation of the undiseased human cardiac
// ventricular action potential: model formulation and experimental
// validation", PLoS Comput Biol 7:e1002061, 2011), written by hand from the
// published equations in the interface of the generated model headers.  The
// reference tree holds no CellML or code for this model.
//
// Currents are per membrane capacitance (uA/uF); V in mV, time in ms,
// concentrations in mM.  Rush-Larsen gates are the states whose derivative is
// (inf(V) - y) / tau(V) with V-dependent inf and tau only; nca, the release
// fluxes Jrel_np / Jrel_p and CaMK_trap depend on other states and step by
// forward Euler.  The GHK-type fluxes are 0/0 at V = 0 and are interpolated
// across a window of 1e-7 in their exponent, as the generated models do.
#pragma once
#include "model_common.h"

namespace cmx { namespace ord_endo {

constexpr int kNumStates = 41;
constexpr int kVIndex = 0;
constexpr int kNumParams = 7;
constexpr bool kPerCapacitanceUnits = true;

static const char* const kStateNames[kNumStates] = {
    "membrane__V", "intracellular_ions__nai", "intracellular_ions__nass", "intracellular_ions__ki",
    "intracellular_ions__kss", "intracellular_ions__cai", "intracellular_ions__cass", "intracellular_ions__cansr",
    "intracellular_ions__cajsr", "INa__m", "INa__hf", "INa__hs", "INa__j", "INa__hsp", "INa__jp", "INaL__mL",
    "INaL__hL", "INaL__hLp", "Ito__a", "Ito__iF", "Ito__iS", "Ito__ap", "Ito__iFp", "Ito__iSp", "ICaL__d",
    "ICaL__ff", "ICaL__fs", "ICaL__fcaf", "ICaL__fcas", "ICaL__jca", "ICaL__nca", "ICaL__ffp", "ICaL__fcafp",
    "IKr__xrf", "IKr__xrs", "IKs__xs1", "IKs__xs2", "IK1__xk1", "ryr__Jrelnp", "ryr__Jrelp", "CaMK__CaMKt",
};
// The published model's initial values.
static constexpr double kInitialState[kNumStates] = {
    -87.5, 7.0, 7.0, 145.0, 145.0, 1.0e-4, 1.0e-4, 1.2, 1.2,      // V, nai, nass, ki, kss, cai, cass, cansr, cajsr
    0.0, 1.0, 1.0, 1.0, 1.0, 1.0,                                  // m, hf, hs, j, hsp, jp
    0.0, 1.0, 1.0,                                                 // mL, hL, hLp
    0.0, 1.0, 1.0, 0.0, 1.0, 1.0,                                  // a, iF, iS, ap, iFp, iSp
    0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 1.0,                   // d, ff, fs, fcaf, fcas, jca, nca, ffp, fcafp
    0.0, 0.0, 0.0, 0.0, 1.0,                                       // xrf, xrs, xs1, xs2, xk1
    0.0, 0.0, 0.0,                                                 // Jrelnp, Jrelp, CaMKt
};
// Scalable conductances / permeabilities (endocardial values).
static const char* const kParamNames[kNumParams] = {
    "INa__GNa", "ICaL__PCa", "IKr__GKr", "IKs__GKs", "IK1__GK1", "Ito__Gto", "INaL__GNaL",
};
// (no CellML file, so no annotated names)
static const char* const kParamAliases[kNumParams] = {"", "", "", "", "", "", ""};
static constexpr double kParamDefaults[kNumParams] = {75.0, 1.0e-4, 0.046, 0.0034, 0.1908, 0.02, 0.0075};
static constexpr int kIsGate[kNumStates] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0,
    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,  // m .. jca (9-29)
    0,                                                              // nca
    1, 1, 1, 1, 1, 1, 1,                                            // ffp .. xk1 (31-37)
    0, 0, 0,                                                        // Jrelnp, Jrelp, CaMKt
};

// Physical constants and cell geometry.
template <class R>
struct Consts {
    static constexpr double nao = 140.0, cao = 1.8, ko = 5.4;
    static constexpr double Rg = 8314.0, T = 310.0, F = 96485.0;
    static constexpr double L = 0.01, rad = 0.0011;
    static constexpr double vcell = 1000.0 * 3.14 * rad * rad * L;
    static constexpr double Ageo = 2.0 * 3.14 * rad * rad + 2.0 * 3.14 * rad * L;
    static constexpr double Acap = 2.0 * Ageo;
    static constexpr double vmyo = 0.68 * vcell, vnsr = 0.0552 * vcell, vjsr = 0.0048 * vcell, vss = 0.02 * vcell;
    static constexpr double KmCaMK = 0.15, aCaMK = 0.05, bCaMK = 0.00068, CaMKo = 0.05, KmCaM = 0.0015;
    static constexpr double RTF = Rg * T / F;          // mV
};

// x^n for small integer n
template <class R>
CMX_HD inline R sq(R x) { return x * x; }

// GHK-type flux  z^2 V F^2/(R T) (c_in e^{zVF/RT} - g c_out) / (e^{zVF/RT} - 1)
// (the factor P and the concentrations given), with the removable 0/0 at V = 0
// interpolated across |zVF/RT| < 1e-7.
template <class R>
CMX_HD inline R ghk(R V, R z, R cin, R gcout)
{
    using C = Consts<R>;
    const R dv = R(1e-7) * R(C::RTF) / z;
    return cmx_remove_singularity(V, R(0), dv, [&](const R v) {
        const R e = z * v * R(1.0 / C::RTF);
        return cmx_div(z * z * v * R(C::F / C::RTF) * (cin * cmx_exp(e) - gcout), cmx_expm1(e));
    });
}

// Na/Ca exchanger fluxes (Kang-Rudy four-state form) for one compartment:
// returns zna JncxNa + zca JncxCa scaled by the allosteric factor.
template <class R>
CMX_HD inline R inaca_flux(R V, R na, R ca)
{
    using C = Consts<R>;
    const R kna1 = R(15.0), kna2 = R(5.0), kna3 = R(88.12), kasymm = R(12.5), wna = R(6.0e4), wca = R(6.0e4),
            wnaca = R(5.0e3), kcaon = R(1.5e6), kcaoff = R(5.0e3), qna = R(0.5224), qca = R(0.1670);
    const R nao = R(C::nao), cao = R(C::cao);
    const R hca = cmx_exp(qca * V * R(1.0 / C::RTF));
    const R hna = cmx_exp(qna * V * R(1.0 / C::RTF));
    const R h1 = R(1.0) + cmx_div(na, kna3) * (R(1.0) + hna);
    const R h2 = cmx_div(na * hna, kna3 * h1);
    const R h3 = cmx_div(R(1.0), h1);
    const R h4 = R(1.0) + cmx_div(na, kna1) * (R(1.0) + cmx_div(na, kna2));
    const R h5 = cmx_div(na * na, h4 * kna1 * kna2);
    const R h6 = cmx_div(R(1.0), h4);
    const R h7 = R(1.0) + cmx_div(nao, kna3) * (R(1.0) + cmx_div(R(1.0), hna));
    const R h8 = cmx_div(nao, kna3 * hna * h7);
    const R h9 = cmx_div(R(1.0), h7);
    const R h10 = kasymm + R(1.0) + cmx_div(nao, kna1) * (R(1.0) + cmx_div(nao, kna2));
    const R h11 = cmx_div(nao * nao, h10 * kna1 * kna2);
    const R h12 = cmx_div(R(1.0), h10);
    const R k1 = h12 * cao * kcaon;
    const R k2 = kcaoff;
    const R k3p = h9 * wca;
    const R k3pp = h8 * wnaca;
    const R k3 = k3p + k3pp;
    const R k4p = cmx_div(h3 * wca, hca);
    const R k4pp = h2 * wnaca;
    const R k4 = k4p + k4pp;
    const R k5 = kcaoff;
    const R k6 = h6 * ca * kcaon;
    const R k7 = h5 * h2 * wna;
    const R k8 = h8 * h11 * wna;
    const R x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
    const R x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
    const R x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
    const R x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
    const R xs = cmx_div(R(1.0), x1 + x2 + x3 + x4);
    const R E1 = x1 * xs, E2 = x2 * xs, E3 = x3 * xs, E4 = x4 * xs;
    const R KmCaAct = R(150.0e-6);
    const R allo = cmx_div(R(1.0), R(1.0) + sq(cmx_div(KmCaAct, ca)));
    const R JncxNa = R(3.0) * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
    const R JncxCa = E2 * k2 - E1 * k1;
    return allo * (JncxNa + R(2.0) * JncxCa);
}

// Every current and flux at state y (V given separately).
template <class R>
struct Terms {
    R INa, INaL, Ito, ICaL, ICaNa, ICaK, IKr, IKs, IK1, INaCa_i, INaCa_ss, INaK, INab, ICab, IKb, IpCa;
    R JdiffNa, JdiffK, Jdiff, Jrel, Jup, Jtr, dCaMKt, dnca;
    R Jrel_inf, tau_rel, Jrel_infp, tau_relp;
    R fICaLp, PhiCaL;
};

template <class R>
CMX_HD inline void terms(const R* y, R V, const R* p, Terms<R>& t)
{
    CMX_MATH_USING
    using C = Consts<R>;
    const R nai = y[1], nass = y[2], ki = y[3], kss = y[4], cai = y[5], cass = y[6], cansr = y[7], cajsr = y[8];
    const R nao = R(C::nao), cao = R(C::cao), ko = R(C::ko);
    // CaMK
    const R CaMKb = cmx_div(R(C::CaMKo) * (R(1.0) - y[40]), R(1.0) + cmx_div(R(C::KmCaM), cass));
    const R CaMKa = CaMKb + y[40];
    t.dCaMKt = R(C::aCaMK) * CaMKb * (CaMKb + y[40]) - R(C::bCaMK) * y[40];
    const R fp_camk = cmx_div(R(1.0), R(1.0) + cmx_div(R(C::KmCaMK), CaMKa));
    // reversal potentials
    const R ENa = R(C::RTF) * cmx_log(cmx_div(nao, nai));
    const R EK = R(C::RTF) * cmx_log(cmx_div(ko, ki));
    const R PKNa = R(0.01833);
    const R EKs = R(C::RTF) * cmx_log(cmx_div(ko + PKNa * nao, ki + PKNa * nai));
    // INa
    const R Ahf = R(0.99), Ahs = R(1.0) - Ahf;
    const R h = Ahf * y[10] + Ahs * y[11];
    const R hp = Ahf * y[10] + Ahs * y[13];
    t.INa = p[0] * (V - ENa) * y[9] * y[9] * y[9] * ((R(1.0) - fp_camk) * h * y[12] + fp_camk * hp * y[14]);
    // INaL
    t.INaL = p[6] * (V - ENa) * y[15] * ((R(1.0) - fp_camk) * y[16] + fp_camk * y[17]);
    // Ito
    const R AiF = cmx_div(R(1.0), R(1.0) + cmx_exp((V - R(213.6)) * R(1.0 / 151.2)));
    const R AiS = R(1.0) - AiF;
    const R ii = AiF * y[19] + AiS * y[20];
    const R ip = AiF * y[22] + AiS * y[23];
    t.Ito = p[5] * (V - EK) * ((R(1.0) - fp_camk) * y[18] * ii + fp_camk * y[21] * ip);
    // ICaL, ICaNa, ICaK
    const R Aff = R(0.6), Afs = R(1.0) - Aff;
    const R f = Aff * y[25] + Afs * y[26];
    const R Afcaf = R(0.3) + cmx_div(R(0.6), R(1.0) + cmx_exp((V - R(10.0)) * R(0.1)));
    const R Afcas = R(1.0) - Afcaf;
    const R fca = Afcaf * y[27] + Afcas * y[28];
    const R fpp = Aff * y[31] + Afs * y[26];
    const R fcap = Afcaf * y[32] + Afcas * y[28];
    const R jca = y[29], nca = y[30];
    const R Kmn = R(0.002), k2n = R(1000.0);
    const R km2n = jca * R(1.0);
    const R anca = cmx_div(R(1.0), cmx_div(k2n, km2n) + sq(sq(R(1.0) + cmx_div(Kmn, cass))));
    t.dnca = anca * k2n - nca * km2n;
    t.PhiCaL = ghk<R>(V, R(2.0), cass, R(0.341) * cao);
    const R PhiCaNa = ghk<R>(V, R(1.0), R(0.75) * nass, R(0.75) * nao);
    const R PhiCaK = ghk<R>(V, R(1.0), R(0.75) * kss, R(0.75) * ko);
    const R PCa = p[1];
    const R PCap = R(1.1) * PCa;
    const R PCaNa = R(0.00125) * PCa, PCaK = R(3.574e-4) * PCa;
    const R PCaNap = R(0.00125) * PCap, PCaKp = R(3.574e-4) * PCap;
    t.fICaLp = fp_camk;
    const R gate_np = y[24] * (f * (R(1.0) - nca) + jca * fca * nca);
    const R gate_p = y[24] * (fpp * (R(1.0) - nca) + jca * fcap * nca);
    t.ICaL = (R(1.0) - fp_camk) * PCa * t.PhiCaL * gate_np + fp_camk * PCap * t.PhiCaL * gate_p;
    t.ICaNa = (R(1.0) - fp_camk) * PCaNa * PhiCaNa * gate_np + fp_camk * PCaNap * PhiCaNa * gate_p;
    t.ICaK = (R(1.0) - fp_camk) * PCaK * PhiCaK * gate_np + fp_camk * PCaKp * PhiCaK * gate_p;
    // IKr
    const R Axrf = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(54.81)) * R(1.0 / 38.21)));
    const R Axrs = R(1.0) - Axrf;
    const R xr = Axrf * y[33] + Axrs * y[34];
    const R rkr = cmx_div(R(1.0), (R(1.0) + cmx_exp((V + R(55.0)) * R(1.0 / 75.0))) *
                                      (R(1.0) + cmx_exp((V - R(10.0)) * R(1.0 / 30.0))));
    t.IKr = p[2] * R(1.0) * xr * rkr * (V - EK);            // sqrt(ko / 5.4) = 1
    // IKs
    const R KsCa = R(1.0) + cmx_div(R(0.6), R(1.0) + cmx_exp(R(1.4) * cmx_log(cmx_div(R(3.8e-5), cai))));
    t.IKs = p[3] * KsCa * y[35] * y[36] * (V - EKs);
    // IK1
    const R rk1 = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(105.8) - R(2.6) * ko) * R(1.0 / 9.493)));
    t.IK1 = p[4] * R(2.32379000772445) * rk1 * y[37] * (V - EK);   // sqrt(ko)
    // INaCa
    const R Gncx = R(0.0008);
    t.INaCa_i = R(0.8) * Gncx * inaca_flux<R>(V, nai, cai);
    t.INaCa_ss = R(0.2) * Gncx * inaca_flux<R>(V, nass, cass);
    // INaK
    {
        const R k1p = R(949.5), k1m = R(182.4), k2p = R(687.2), k2m = R(39.4), k3p = R(1899.0), k3m = R(79300.0),
                k4p = R(639.0), k4m = R(40.0);
        const R Knai0 = R(9.073), Knao0 = R(27.78), delta = R(-0.1550);
        const R Knai = Knai0 * cmx_exp(delta * V * R(1.0 / (3.0 * C::RTF)));
        const R Knao = Knao0 * cmx_exp((R(1.0) - delta) * V * R(1.0 / (3.0 * C::RTF)));
        const R Kki = R(0.5), Kko = R(0.3582), MgADP = R(0.05), MgATP = R(9.8), Kmgatp = R(1.698e-7), H = R(1.0e-7),
                eP = R(4.2), Khp = R(1.698e-7), Knap = R(224.0), Kxkur = R(292.0);
        const R P = cmx_div(eP, R(1.0) + cmx_div(H, Khp) + cmx_div(nai, Knap) + cmx_div(ki, Kxkur));
        const R rnai = cmx_div(nai, Knai), rnao = cmx_div(nao, Knao), rki = cmx_div(ki, Kki), rko = cmx_div(ko, Kko);
        const R den_i = (R(1.0) + rnai) * (R(1.0) + rnai) * (R(1.0) + rnai) + sq(R(1.0) + rki) - R(1.0);
        const R den_o = (R(1.0) + rnao) * (R(1.0) + rnao) * (R(1.0) + rnao) + sq(R(1.0) + rko) - R(1.0);
        const R a1 = cmx_div(k1p * rnai * rnai * rnai, den_i);
        const R b1 = k1m * MgADP;
        const R a2 = k2p;
        const R b2 = cmx_div(k2m * rnao * rnao * rnao, den_o);
        const R a3 = cmx_div(k3p * rko * rko, den_o);
        const R b3 = cmx_div(k3m * P * H, R(1.0) + cmx_div(MgATP, Kmgatp));
        const R a4 = cmx_div(k4p * cmx_div(MgATP, Kmgatp), R(1.0) + cmx_div(MgATP, Kmgatp));
        const R b4 = cmx_div(k4m * rki * rki, den_i);
        const R x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
        const R x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
        const R x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
        const R x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
        const R xs = cmx_div(R(1.0), x1 + x2 + x3 + x4);
        const R E1 = x1 * xs, E2 = x2 * xs, E3 = x3 * xs, E4 = x4 * xs;
        const R JnakNa = R(3.0) * (E1 * a3 - E2 * b3);
        const R JnakK = R(2.0) * (E4 * b1 - E3 * a1);
        const R Pnak = R(30.0);
        t.INaK = Pnak * (JnakNa + JnakK);
    }
    // background currents and the sarcolemmal Ca pump
    const R xkb = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V - R(14.48)) * R(1.0 / 18.34)));
    t.IKb = R(0.003) * xkb * (V - EK);
    t.INab = R(3.75e-10) * ghk<R>(V, R(1.0), nai, nao);
    t.ICab = R(2.5e-8) * ghk<R>(V, R(2.0), cai, R(0.341) * cao);
    t.IpCa = cmx_div(R(0.0005) * cai, R(0.0005) + cai);
    // diffusion, release, uptake, transfer
    t.JdiffNa = (nass - nai) * R(0.5);
    t.JdiffK = (kss - ki) * R(0.5);
    t.Jdiff = (cass - cai) * R(5.0);
    const R bt = R(4.75);
    const R a_rel = R(0.5) * bt;
    const R cj8 = sq(sq(sq(cmx_div(R(1.5), cajsr))));
    t.Jrel_inf = cmx_div(a_rel * (-t.ICaL), R(1.0) + cj8);
    const R tr = cmx_div(bt, R(1.0) + cmx_div(R(0.0123), cajsr));
    t.tau_rel = tr < R(0.001) ? R(0.001) : tr;
    const R btp = R(1.25) * bt;
    const R a_relp = R(0.5) * btp;
    t.Jrel_infp = cmx_div(a_relp * (-t.ICaL), R(1.0) + cj8);
    const R trp = cmx_div(btp, R(1.0) + cmx_div(R(0.0123), cajsr));
    t.tau_relp = trp < R(0.001) ? R(0.001) : trp;
    t.Jrel = (R(1.0) - fp_camk) * y[38] + fp_camk * y[39];
    const R Jupnp = cmx_div(R(0.004375) * cai, cai + R(0.00092));
    const R Jupp = cmx_div(R(2.75) * R(0.004375) * cai, cai + R(0.00092) - R(0.00017));
    const R Jleak = R(0.0039375) * cansr * R(1.0 / 15.0);
    t.Jup = (R(1.0) - fp_camk) * Jupnp + fp_camk * Jupp - Jleak;
    t.Jtr = (cansr - cajsr) * R(0.01);
}

template <class R>
CMX_HD inline R total_current(const Terms<R>& t)
{
    return t.INa + t.INaL + t.Ito + t.ICaL + t.ICaNa + t.ICaK + t.IKr + t.IKs + t.IK1 + t.INaCa_i + t.INaCa_ss +
           t.INaK + t.INab + t.IKb + t.IpCa + t.ICab;
}

// Steady states and time constants of the 28 gates at V.
template <class R>
CMX_HD inline void gates(R V, R* inf, R* tau)
{
    // INa
    inf[9] = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V + R(39.57)) * R(1.0 / 9.871)));
    tau[9] = cmx_div(R(1.0), R(6.765) * cmx_exp((V + R(11.64)) * R(1.0 / 34.77)) +
                                  R(8.552) * cmx_exp(-(V + R(77.42)) * R(1.0 / 5.955)));
    const R hss = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(82.90)) * R(1.0 / 6.086)));
    inf[10] = hss;
    tau[10] = cmx_div(R(1.0), R(1.432e-5) * cmx_exp(-(V + R(1.196)) * R(1.0 / 6.285)) +
                                   R(6.149) * cmx_exp((V + R(0.5096)) * R(1.0 / 20.27)));
    inf[11] = hss;
    const R ths = cmx_div(R(1.0), R(0.009794) * cmx_exp(-(V + R(17.95)) * R(1.0 / 28.05)) +
                                       R(0.3343) * cmx_exp((V + R(5.730)) * R(1.0 / 56.66)));
    tau[11] = ths;
    inf[12] = hss;
    const R tj = R(2.038) + cmx_div(R(1.0), R(0.02136) * cmx_exp(-(V + R(100.6)) * R(1.0 / 8.281)) +
                                                  R(0.3052) * cmx_exp((V + R(0.9941)) * R(1.0 / 38.45)));
    tau[12] = tj;
    inf[13] = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(89.1)) * R(1.0 / 6.086)));
    tau[13] = R(3.0) * ths;
    inf[14] = hss;
    tau[14] = R(1.46) * tj;
    // INaL
    inf[15] = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V + R(42.85)) * R(1.0 / 5.264)));
    tau[15] = tau[9];
    inf[16] = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(87.61)) * R(1.0 / 7.488)));
    tau[16] = R(200.0);
    inf[17] = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(93.81)) * R(1.0 / 7.488)));
    tau[17] = R(600.0);
    // Ito (endocardial: no delta_epi factor)
    inf[18] = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V - R(14.34)) * R(1.0 / 14.82)));
    const R ta = cmx_div(R(1.0515), cmx_div(R(1.0), R(1.2089) * (R(1.0) + cmx_exp(-(V - R(18.4099)) * R(1.0 / 29.3814)))) +
                                        cmx_div(R(3.5), R(1.0) + cmx_exp((V + R(100.0)) * R(1.0 / 29.3814))));
    tau[18] = ta;
    const R iss = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(43.94)) * R(1.0 / 5.711)));
    const R tiF = R(4.562) + cmx_div(R(1.0), R(0.3933) * cmx_exp(-(V + R(100.0)) * R(0.01)) +
                                                   R(0.08004) * cmx_exp((V + R(50.0)) * R(1.0 / 16.59)));
    const R tiS = R(23.62) + cmx_div(R(1.0), R(0.001416) * cmx_exp(-(V + R(96.52)) * R(1.0 / 59.05)) +
                                                   R(1.780e-8) * cmx_exp((V + R(114.1)) * R(1.0 / 8.079)));
    inf[19] = iss;
    tau[19] = tiF;
    inf[20] = iss;
    tau[20] = tiS;
    inf[21] = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V - R(24.34)) * R(1.0 / 14.82)));
    tau[21] = ta;
    const R dti_develop = R(1.354) + cmx_div(R(1.0e-4), cmx_exp((V - R(167.4)) * R(1.0 / 15.89)) +
                                                            cmx_exp(-(V - R(12.23)) * R(1.0 / 0.2154)));
    const R dti_recover = R(1.0) - cmx_div(R(0.5), R(1.0) + cmx_exp((V + R(70.0)) * R(0.05)));
    inf[22] = iss;
    tau[22] = dti_develop * dti_recover * tiF;
    inf[23] = iss;
    tau[23] = dti_develop * dti_recover * tiS;
    // ICaL
    inf[24] = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V + R(3.940)) * R(1.0 / 4.230)));
    tau[24] = R(0.6) + cmx_div(R(1.0), cmx_exp(R(-0.05) * (V + R(6.0))) + cmx_exp(R(0.09) * (V + R(14.0))));
    const R fss = cmx_div(R(1.0), R(1.0) + cmx_exp((V + R(19.58)) * R(1.0 / 3.696)));
    const R tff = R(7.0) + cmx_div(R(1.0), R(0.0045) * cmx_exp(-(V + R(20.0)) * R(0.1)) +
                                                 R(0.0045) * cmx_exp((V + R(20.0)) * R(0.1)));
    inf[25] = fss;
    tau[25] = tff;
    inf[26] = fss;
    tau[26] = R(1000.0) + cmx_div(R(1.0), R(0.000035) * cmx_exp(-(V + R(5.0)) * R(0.25)) +
                                                R(0.000035) * cmx_exp((V + R(5.0)) * R(1.0 / 6.0)));
    const R tfcaf = R(7.0) + cmx_div(R(1.0), R(0.04) * cmx_exp(-(V - R(4.0)) * R(1.0 / 7.0)) +
                                                   R(0.04) * cmx_exp((V - R(4.0)) * R(1.0 / 7.0)));
    inf[27] = fss;
    tau[27] = tfcaf;
    inf[28] = fss;
    tau[28] = R(100.0) + cmx_div(R(1.0), R(0.00012) * cmx_exp(-V * R(1.0 / 3.0)) +
                                               R(0.00012) * cmx_exp(V * R(1.0 / 7.0)));
    inf[29] = fss;
    tau[29] = R(75.0);
    inf[31] = fss;
    tau[31] = R(2.5) * tff;
    inf[32] = fss;
    tau[32] = R(2.5) * tfcaf;
    // IKr
    const R xrss = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V + R(8.337)) * R(1.0 / 6.789)));
    inf[33] = xrss;
    tau[33] = R(12.98) + cmx_div(R(1.0), R(0.3652) * cmx_exp((V - R(31.66)) * R(1.0 / 3.869)) +
                                               R(4.123e-5) * cmx_exp(-(V - R(47.78)) * R(1.0 / 20.38)));
    inf[34] = xrss;
    tau[34] = R(1.865) + cmx_div(R(1.0), R(0.06629) * cmx_exp((V - R(34.70)) * R(1.0 / 7.355)) +
                                               R(1.128e-5) * cmx_exp(-(V - R(29.74)) * R(1.0 / 25.94)));
    // IKs
    const R xs1ss = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V + R(11.60)) * R(1.0 / 8.932)));
    inf[35] = xs1ss;
    tau[35] = R(817.3) + cmx_div(R(1.0), R(2.326e-4) * cmx_exp((V + R(48.28)) * R(1.0 / 17.80)) +
                                               R(0.001292) * cmx_exp(-(V + R(210.0)) * R(1.0 / 230.0)));
    inf[36] = xs1ss;
    tau[36] = cmx_div(R(1.0), R(0.01) * cmx_exp((V - R(50.0)) * R(0.05)) +
                                   R(0.0193) * cmx_exp(-(V + R(66.54)) * R(1.0 / 31.0)));
    // IK1 (ko = 5.4)
    inf[37] = cmx_div(R(1.0), R(1.0) + cmx_exp(-(V + R(2.5538 * 5.4 + 144.59)) * R(1.0 / (1.5692 * 5.4 + 3.8115))));
    tau[37] = cmx_div(R(122.2), cmx_exp(-(V + R(127.2)) * R(1.0 / 20.36)) + cmx_exp((V + R(236.8)) * R(1.0 / 69.33)));
}

// Derivatives of the non-gate states (V excluded) from the terms.
template <class R>
CMX_HD inline void nongate_rates(const R* y, const Terms<R>& t, R i_stim, R* dy)
{
    using C = Consts<R>;
    const R cai = y[5], cass = y[6], cajsr = y[8];
    const R AF_myo = R(C::Acap / (C::F * C::vmyo)), AF_ss = R(C::Acap / (C::F * C::vss));
    const R ss_myo = R(C::vss / C::vmyo);
    dy[1] = -(t.INa + t.INaL + R(3.0) * t.INaCa_i + R(3.0) * t.INaK + t.INab) * AF_myo + t.JdiffNa * ss_myo;
    dy[2] = -(t.ICaNa + R(3.0) * t.INaCa_ss) * AF_ss - t.JdiffNa;
    dy[3] = -(t.Ito + t.IKr + t.IKs + t.IK1 + t.IKb + i_stim - R(2.0) * t.INaK) * AF_myo + t.JdiffK * ss_myo;
    dy[4] = -t.ICaK * AF_ss - t.JdiffK;
    const R cmdnmax = R(0.05), kmcmdn = R(0.00238), trpnmax = R(0.07), kmtrpn = R(0.0005), BSRmax = R(0.047),
            KmBSR = R(0.00087), BSLmax = R(1.124), KmBSL = R(0.0087), csqnmax = R(10.0), kmcsqn = R(0.8);
    const R Bcai = cmx_div(R(1.0), R(1.0) + cmx_div(cmdnmax * kmcmdn, sq(kmcmdn + cai)) +
                                       cmx_div(trpnmax * kmtrpn, sq(kmtrpn + cai)));
    dy[5] = Bcai * (-(t.IpCa + t.ICab - R(2.0) * t.INaCa_i) * R(0.5) * AF_myo - t.Jup * R(C::vnsr / C::vmyo) +
                    t.Jdiff * ss_myo);
    const R Bcass = cmx_div(R(1.0), R(1.0) + cmx_div(BSRmax * KmBSR, sq(KmBSR + cass)) +
                                        cmx_div(BSLmax * KmBSL, sq(KmBSL + cass)));
    dy[6] = Bcass * (-(t.ICaL - R(2.0) * t.INaCa_ss) * R(0.5) * AF_ss + t.Jrel * R(C::vjsr / C::vss) - t.Jdiff);
    dy[7] = t.Jup - t.Jtr * R(C::vjsr / C::vnsr);
    const R Bcajsr = cmx_div(R(1.0), R(1.0) + cmx_div(csqnmax * kmcsqn, sq(kmcsqn + cajsr)));
    dy[8] = Bcajsr * (t.Jtr - t.Jrel);
    dy[30] = t.dnca;
    dy[38] = cmx_div(t.Jrel_inf - y[38], t.tau_rel);
    dy[39] = cmx_div(t.Jrel_infp - y[39], t.tau_relp);
    dy[40] = t.dCaMKt;
}

// Right-hand side with V clamped, in the generated headers' convention:
// non-gates (and, with RL=false, every state) get dy[i]; with RL=true gate i
// gets g_inf[i], g_k[i] with y_i(t+dt) = g_inf + (y_i - g_inf) exp(g_k dt).
// i_stim_area in uA/cm^2; cm the capacitance in uF/cm^2.
template <class R, bool RL>
CMX_HD inline void evaluate(const R* y, R V, R i_stim_area, const R* p, R cm, R* dy, R* g_inf, R* g_k)
{
    Terms<R> t;
    terms<R>(y, V, p, t);
    R inf[kNumStates], tau[kNumStates];
    gates<R>(V, inf, tau);
    nongate_rates<R>(y, t, cmx_div(i_stim_area, cm), dy);
CMX_UNROLL
    for (int s = 1; s < kNumStates; ++s) {
        if (!kIsGate[s]) continue;
        if (RL) {
            g_inf[s] = inf[s];
            g_k[s] = cmx_div(R(-1), tau[s]);
        } else {
            dy[s] = cmx_div(inf[s] - y[s], tau[s]);
        }
    }
}

// dV/dt including the stimulus (single-cell integration).
template <class R>
CMX_HD inline R dvdt(const R* y, R V, R i_stim_area, const R* p, R cm)
{
    Terms<R> t;
    terms<R>(y, V, p, t);
    return -(total_current<R>(t) + cmx_div(i_stim_area, cm));
}

// Total ionic current in uA/cm^2 (GetIIonic): the per-capacitance currents
// times the capacitance.
template <class R>
CMX_HD inline R iionic(const R* y, R V, const R* p, R cm)
{
    Terms<R> t;
    terms<R>(y, V, p, t);
    return total_current<R>(t) * cm;
}

}}  // namespace cmx::ord_endo
