// gw_model.h -- the Greenstein-Winslow global model on the device (and host):
// derived constants, the per-step constant tables the SSA reads, and the
// explicit-Euler global ODE step. Every expression keeps the reference's
// operation order so that results are bit-identical to the CPU oracle.
#pragma once

#include "gw_math.h"
#include "../../include/gw_capi.h"

namespace gw {

constexpr double kFaraday = 96.5;   // src/includes/common.hpp:9
constexpr double kGasConst = 8.314; // src/includes/common.hpp:10

GW_HD double sq(double x) { return x * x; }
GW_HD double cube(double x) { return x * x * x; }

// Derived constants of GW::Constants (src/includes/GW_utils.tpp:67-116).
struct CellConsts {
    double RT_F, F_RT, CRU_factor, CSA_FVcyto, VSS_Vcyto, Vcyto_VNSR, VJSR_VNSR;
    double gamma0a, gamma0a2, gamma0a3, gamma0a4, binv, omega_b, omega_b2, omega_b3, omega_b4;
    double BSR_const, BSL_const, VSS_VJSR, CSQN_const, JLCC_const, Cao_scaled;
    double Nao3, INaCa_const, sigma, INaK_const, sqrtKo, PKv14_Csc, Ito2_const, IK1_const, ICaL_const, CMDN_const;
};

GW_HD void cell_consts_init(CellConsts &c, const gw_params_t &p, int nCRU)
{
    c.RT_F = kGasConst * p.T / kFaraday;
    c.F_RT = 1.0 / c.RT_F;
    c.CRU_factor = static_cast<double>(p.NCaRU) / static_cast<double>(nCRU);
    c.CSA_FVcyto = p.CSA / (1000 * p.Vcyto * kFaraday);
    c.VSS_Vcyto = p.VSS / p.Vcyto;
    c.Vcyto_VNSR = p.Vcyto / p.VNSR;
    c.VJSR_VNSR = p.VJSR / p.VNSR;
    c.gamma0a = p.gamma0 * p.a;
    c.gamma0a2 = c.gamma0a * p.a;
    c.gamma0a3 = c.gamma0a2 * p.a;
    c.gamma0a4 = c.gamma0a3 * p.a;
    c.binv = 1.0 / p.b;
    c.omega_b = p.omega * c.binv;
    c.omega_b2 = c.omega_b * c.binv;
    c.omega_b3 = c.omega_b2 * c.binv;
    c.omega_b4 = c.omega_b3 * c.binv;
    c.BSR_const = p.KBSR * p.BSRT;
    c.BSL_const = p.KBSL * p.BSLT;
    c.VSS_VJSR = p.VSS / p.VJSR;
    c.CSQN_const = p.KCSQN * p.CSQNT;
    c.JLCC_const = 2.0e6 * p.PCaL / p.VSS;
    c.Cao_scaled = 0.341 * p.Cao;
    c.Nao3 = p.Nao * p.Nao * p.Nao;
    c.INaCa_const = 5000.0 * p.kNaCa / ((p.KmNa * p.KmNa * p.KmNa + p.Nao * p.Nao * p.Nao) * (p.KmCa + p.Cao));
    c.sigma = (gw::exp(p.Nao / 67.3) - 1.0) / 7.0;
    c.INaK_const = p.INaKmax * p.Ko / (p.Ko + p.KmKo);
    c.sqrtKo = __builtin_sqrt(p.Ko);
    c.PKv14_Csc = p.PKv14 / p.Csc;
    c.Ito2_const = 1.0e9 * p.Pto2 * kFaraday * c.CRU_factor / p.CSA;
    c.IK1_const = p.Ko / (p.Ko + p.KmK1);
    c.ICaL_const = -1000.0 * (2.0 * kFaraday * p.VSS) * c.CRU_factor / p.CSA;
    c.CMDN_const = p.KCMDN * p.CMDNT;
}

// Constants the SSA reads during one step (frozen inside GW_model::SSA,
// src/includes/GW.tpp:214-233). The LCC propensities of a subunit in state s
// are tabulated so that (r0 + r1) + r2 = lccP[s] + lccQ[s] * CaSS bit for bit;
// R0, R1 hold r0, r1 with the CaSS-proportional slot set to zero (r2 is
// never needed on its own: the transition scan stops before comparing it).
template <class R>
struct LccTableT {
    R P[13], Q[13], R0[13], R1[13];
};
using LccTable = LccTableT<double>;
using LccTableF = LccTableT<float>;

struct StepConsts {
    double Cai, CaNSR;
    double inact0, inact1;   // (y == 0) ? yinf/tau : (1 - yinf)/tau  (GW_utils.hpp:381)
    double JLCC_mult, JLCC_exp;
    uint32_t step_lo;
    uint32_t _pad;
    LccTable lcc;            // read per lane (indexed by the LCC state)
};

// GW::alphaLCC .. tauLCC (GW_utils.hpp:365-375)
// The exponentials of the per-step constants: alpha, beta, y_inf and tau of
// the LCC (GW_utils.hpp:365-375, as exp(arg)), then exp(-V F/RT).
enum ScSlot : int { SC_AL, SC_BE, SC_YINF, SC_TAU, SC_EXPM, SC_N };

GW_HD void sc_args(double V, double F_RT, double *arg)
{
    arg[SC_AL] = 0.012 * (V - 35.0);
    arg[SC_BE] = -0.05 * (V - 35.0);
    arg[SC_YINF] = (V + 12.5) / 5.0;
    arg[SC_TAU] = (V + 30.0) / 12.0;
    arg[SC_EXPM] = -(V * F_RT);
}

// Per-step constants of GW_model::SSA (GW.tpp:216-221) from the exponentials
// X[SC_*] of sc_args, without the LCC rate table. At V == 0 the multiplier
// takes its analytic limit JLCC_const / 2 instead of 0/0.
GW_HD void step_consts_scalars(StepConsts &k, const CellConsts &c, double V, const double *X, double Cai, double CaNSR,
                               uint32_t step_lo)
{
    const double expmVF_RT = X[SC_EXPM];
    const double yinf = 0.4 / (1.0 + X[SC_YINF]) + 0.6;
    const double tau = 340.0 / (1.0 + X[SC_TAU]) + 60.0;
    k.Cai = Cai;
    k.CaNSR = CaNSR;
    k.inact0 = yinf / tau;
    k.inact1 = (1.0 - yinf) / tau;
    k.JLCC_exp = sq(1.0 / expmVF_RT);
    k.JLCC_mult = (V == 0.0) ? c.JLCC_const / 2.0 : c.JLCC_const * V * c.F_RT / (k.JLCC_exp - 1);
    k.step_lo = step_lo;
    k._pad = 0;
}

struct LccRow {
    double P, Q, R0, R1;
};

// Row s of the LCC rate table of update_LCC_rates (GW_utils.tpp:118-186):
// r0, r1, r2 of state s, the CaSS-proportional rate's coefficient in Q.
// State 1 carries its CaSS term in r1: (r0 + g0 CaSS) + 0 == P + Q CaSS with
// P = r0; states 2..5 carry it in r2: (r0 + r1) + Q CaSS; the others
// (r0 + r1) + r2 + 0.
GW_HD LccRow lcc_row(int s, const gw_params_t &p, const CellConsts &c, const double *X)
{
    const double al = 2.0 * X[SC_AL];
    const double be = 0.0882 * X[SC_BE];
    double r0 = 0, r1 = 0, r2 = 0, q = 0;
    switch (s) {
    case 1: r0 = 4 * al; q = p.gamma0; break;
    case 2: r0 = be; r1 = 3 * al; q = c.gamma0a; break;
    case 3: r0 = 2 * be; r1 = 2 * al; q = c.gamma0a2; break;
    case 4: r0 = 3 * be; r1 = al; q = c.gamma0a3; break;
    case 5: r0 = 4 * be; r1 = p.f; q = c.gamma0a4; break;
    case 6: r0 = p.g; break;
    case 7: r0 = p.omega; r1 = 4 * al * p.a; break;
    case 8: r0 = c.omega_b; r1 = be * c.binv; r2 = 3 * al * p.a; break;
    case 9: r0 = c.omega_b2; r1 = 2 * be * c.binv; r2 = 2 * al * p.a; break;
    case 10: r0 = c.omega_b3; r1 = 3 * be * c.binv; r2 = al * p.a; break;
    case 11: r0 = c.omega_b4; r1 = 4 * be * c.binv; r2 = p.f1; break;
    case 12: r0 = p.g1; break;
    default: break;
    }
    LccRow w;
    w.R0 = r0;
    w.R1 = r1;
    w.Q = q;
    if (s == 1) w.P = r0;
    else if (s >= 2 && s <= 5) w.P = r0 + r1;
    else w.P = r0 + r1 + r2;
    return w;
}

template <class R>
GW_HD void lcc_row_store(LccTableT<R> &t, int s, const LccRow &w)
{
    t.P[s] = static_cast<R>(w.P);
    t.Q[s] = static_cast<R>(w.Q);
    t.R0[s] = static_cast<R>(w.R0);
    t.R1[s] = static_cast<R>(w.R1);
}

// Per-step constants of GW_model::SSA and the LCC rate table, from the
// exponentials X[SC_*] of sc_args.
GW_HD void step_consts_x(StepConsts &k, const gw_params_t &p, const CellConsts &c, double V, const double *X,
                         double Cai, double CaNSR, uint32_t step_lo)
{
    step_consts_scalars(k, c, V, X, Cai, CaNSR, step_lo);
    for (int s = 0; s < 13; ++s) lcc_row_store(k.lcc, s, lcc_row(s, p, c, X));
}

// step_consts_x with its exponentials evaluated in turn (exp(-V F/RT) given).
GW_HD void step_consts(StepConsts &k, const gw_params_t &p, const CellConsts &c, double V, double expmVF_RT,
                       double Cai, double CaNSR, uint32_t step_lo)
{
    double X[SC_N];
    sc_args(V, c.F_RT, X);
    for (int i = 0; i < SC_EXPM; ++i) X[i] = gw::exp(X[i]);
    X[SC_EXPM] = expmVF_RT;
    step_consts_x(k, p, c, V, X, Cai, CaNSR, step_lo);
}

// The SSA constants of a step rounded to FP32, for the FP32-local SSA.
struct StepConstsF {
    double Cai_d, CaNSR_d;   // kept in FP64: they enter the SSA only through differences
    float Cai, CaNSR;
    float inact0, inact1;
    float JLCC_mult, JLCC_exp;
    LccTableF lcc;
};

GW_HD void step_consts_f32_scalars(StepConstsF &f, const StepConsts &k)
{
    f.Cai_d = k.Cai;
    f.CaNSR_d = k.CaNSR;
    f.Cai = static_cast<float>(k.Cai);
    f.CaNSR = static_cast<float>(k.CaNSR);
    f.inact0 = static_cast<float>(k.inact0);
    f.inact1 = static_cast<float>(k.inact1);
    f.JLCC_mult = static_cast<float>(k.JLCC_mult);
    f.JLCC_exp = static_cast<float>(k.JLCC_exp);
}

GW_HD void step_consts_f32(StepConstsF &f, const StepConsts &k)
{
    step_consts_f32_scalars(f, k);
    for (int s = 0; s < 13; ++s) lcc_row_store(f.lcc, s, LccRow{k.lcc.P[s], k.lcc.Q[s], k.lcc.R0[s], k.lcc.R1[s]});
}

// ---------------------------------------------------------------------------
// Global ODE: update_QKr / update_QKv (GW.tpp:25-44, 64-123) and the
// derivatives of update_V_and_concentration_derivatives, update_gate_derivatives,
// update_Kr/Kv_derivatives (GW.tpp:126-212). dg receives dt * f(state) as the
// reference forms it. Returns GW_OK or GW_E_NONPOSITIVE_CONC.
// ---------------------------------------------------------------------------
struct FluxSums {
    double JLCC, Jxfer, Jtr, ClCh;
};

// The exponentials, logarithms and powers of global_derivatives whose
// arguments depend on the state alone, in a fixed slot order. gd_args forms
// each argument exactly as the formula that uses it; gd_finish reads the
// values X[slot]. The values may come from one evaluation after another
// (global_derivatives) or from many lanes at once: the same function of the
// same argument either way, so the derivatives are the same bits.
enum GdSlot : int {
    GD_EXPM,                                  // exp(-VF_RT)
    GD_NACA, GD_NAK, GD_KR_R, GD_KP,          // INaCa, INaK, IKr rectification, IKp
    GD_AM, GD_BM, GD_AH, GD_BH1, GD_BH2, GD_BH3, GD_AJ1, GD_AJ2, GD_AJ3,
    GD_BJ1, GD_BJ2, GD_BJ3, GD_BJ4, GD_XKS, GD_TXKS1, GD_TXKS2,   // gates (common.hpp:175-218)
    GD_KR0, GD_KR1, GD_KR2, GD_KR3, GD_KR4, GD_KR5, GD_KR6,       // update_QKr (GW.tpp:25-44)
    GD_KV0, GD_KV1, GD_KV2, GD_KV3, GD_KV4, GD_KV5, GD_KV6, GD_KV7,   // update_QKv (GW.tpp:64-123)
    GD_LOG0,                                  // logarithms from here
    GD_LNA = GD_LOG0, GD_LK, GD_LKS,          // log(Nao/Nai), log(Ko/Ki), the IKs reversal
    GD_POW0,                                  // powers from here (exponent py[slot - GD_POW0])
    GD_PNAK = GD_POW0, GD_PF, GD_PR,          // INaK, Jup forward and reverse
    GD_N
};

GW_HD double gd_eval(int slot, double arg, double y)
{
    return slot < GD_LOG0 ? gw::exp(arg) : slot < GD_POW0 ? gw::log(arg) : gw::pow(arg, y);
}

// The arguments of the slots before GD_KR0 (gd_args_lo) and from it on
// (gd_args_hi); gd_args is both.
GW_HD void gd_args_lo(const gw_params_t &p, const gw_globals_t &g, double VF_RT, double *arg)
{
    const double V = g.V;
    arg[GD_EXPM] = -VF_RT;
    arg[GD_NACA] = p.eta * VF_RT;
    arg[GD_NAK] = -0.1 * VF_RT;
    arg[GD_KR_R] = 0.0446 * V;
    arg[GD_KP] = (7.488 - V) / 5.98;
    arg[GD_AM] = -0.1 * (V + 47.13);
    arg[GD_BM] = -V / 11.0;
    arg[GD_AH] = -(V + 80.0) / 6.8;
    arg[GD_BH1] = -(V + 10.66) / 11.1;
    arg[GD_BH2] = 0.079 * V;
    arg[GD_BH3] = 0.35 * V;
    arg[GD_AJ1] = 0.2444 * V;
    arg[GD_AJ2] = -0.04391 * V;
    arg[GD_AJ3] = 0.311 * (V + 79.23);
    arg[GD_BJ1] = -2.535e-7 * V;
    arg[GD_BJ2] = -0.1 * (V + 32.0);
    arg[GD_BJ3] = -0.01052 * V;
    arg[GD_BJ4] = -0.1378 * (V + 40.14);
    arg[GD_XKS] = -(V - 24.7) / 13.6;
    arg[GD_TXKS1] = -0.148 * (V - 10.0);
    arg[GD_TXKS2] = 0.0687 * (V - 10.0);
}

GW_HD void gd_args_hi(const gw_params_t &p, const gw_globals_t &g, double *arg, double *py)
{
    const double V = g.V;
    arg[GD_KR0] = 0.0272 * V;
    arg[GD_KR1] = -0.0431 * V;
    arg[GD_KR2] = 0.0262 * V;
    arg[GD_KR3] = 2.71e-6 * V;
    arg[GD_KR4] = -0.0269 * V;
    arg[GD_KR5] = 0.0120 * V;
    arg[GD_KR6] = -0.0443 * V;
    arg[GD_KV0] = p.aaKv14 * V;
    arg[GD_KV1] = p.aaKv43 * V;
    arg[GD_KV2] = -p.aiKv14 * V;
    arg[GD_KV3] = -p.aiKv43 * V;
    arg[GD_KV4] = -p.baKv14 * V;
    arg[GD_KV5] = -p.baKv43 * V;
    arg[GD_KV6] = p.biKv14 * V;
    arg[GD_KV7] = p.biKv43 * V;
    arg[GD_LNA] = p.Nao / g.Nai;
    arg[GD_LK] = p.Ko / g.Ki;
    arg[GD_LKS] = (p.Ko + 0.01833 * p.Nao) / (g.Ki + 0.01833 * g.Nai);
    arg[GD_PNAK] = p.KmNai / g.Nai;
    py[0] = 1.5;
    arg[GD_PF] = g.Cai / p.Kmf;
    py[1] = p.Hf;
    arg[GD_PR] = g.CaNSR / p.Kmr;
    py[2] = p.Hr;
}

GW_HD void gd_args(const gw_params_t &p, const gw_globals_t &g, double VF_RT, double *arg, double *py)
{
    gd_args_lo(p, g, VF_RT, arg);
    gd_args_hi(p, g, arg, py);
}

// Nernst potential (common.hpp:70-75) from lg = log(Xo / Xi); the reference
// exits on a non-positive concentration, this clears ok instead.
GW_HD double nernst(double Xi, double Xo, double lg, double RT_F, double z, bool &ok)
{
    if (Xi <= 0.0 || Xo <= 0.0) {
        ok = false;
        return 0.0;
    }
    return lg * RT_F / z;
}

// Gate rates (common.hpp:175-218) from the exponentials of gd_args.
GW_HD double alpha_m(double V, const double *X) { return (V == -47.13) ? 3.2 : 0.32 * (V + 47.13) / (1.0 - X[GD_AM]); }
GW_HD double beta_m(const double *X) { return 0.08 * X[GD_BM]; }
GW_HD double alpha_h(double V, const double *X) { return (V >= -40.0) ? 0.0 : 0.135 * X[GD_AH]; }
GW_HD double beta_h(double V, const double *X)
{
    return (V >= -40.0) ? 1.0 / (0.13 * (1.0 + X[GD_BH1])) : 3.56 * X[GD_BH2] + 3.1e5 * X[GD_BH3];
}
GW_HD double alpha_j(double V, const double *X)
{
    return (V >= -40.0) ? 0.0 : (-127140.0 * X[GD_AJ1] - 3.474e-5 * X[GD_AJ2]) * (V + 37.78) / (1.0 + X[GD_AJ3]);
}
GW_HD double beta_j(double V, const double *X)
{
    return (V >= -40.0) ? 0.3 * X[GD_BJ1] / (1.0 + X[GD_BJ2]) : 0.1212 * X[GD_BJ3] / (1.0 + X[GD_BJ4]);
}
GW_HD double XKsinf(const double *X) { return 1.0 / (1.0 + X[GD_XKS]); }
GW_HD double tauXKs_inv(double V, const double *X)
{
    if (V == 10.0) return 0.0000719 / 0.148 + 0.000131 / 0.0687;  // analytic limit of 0/0
    return 0.0000719 * (V - 10.0) / (1.0 - X[GD_TXKS1]) + 0.000131 * (V - 10.0) / (X[GD_TXKS2] - 1.0);
}

// IKr chain: generator entries of update_QKr (GW.tpp:25-44) and the
// derivatives of update_Kr_derivatives (GW.tpp:126-133), in registers.
GW_HD void kr_chain(const gw_params_t &p, const double *X, const double x[5], double dt, double d[5])
{
    const double q12 = p.Kf, q21 = p.Kb;   // initialise_QKr (GW.hpp:75)
    const double q01 = 0.0069 * X[GD_KR0];
    const double q00 = -q01;
    const double q10 = 0.0227 * X[GD_KR1];
    const double q11 = -(q10 + q12);
    const double q23 = 0.0218 * X[GD_KR2];
    const double q24 = 1.29e-5 * X[GD_KR3];
    const double q22 = -(q21 + q23 + q24);
    const double q32 = 0.0009 * X[GD_KR4];
    const double q34 = 0.0622 * X[GD_KR5];
    const double q33 = -(q32 + q34);
    const double q43 = 0.0059 * X[GD_KR6];
    const double q42 = q32 * q43 * q24 / (q23 * q34);
    const double q44 = -(q42 + q43);
    d[0] = dt * (q10 * x[1] + q00 * x[0]);
    d[1] = dt * (q01 * x[0] + q21 * x[2] + q11 * x[1]);
    d[2] = dt * (q12 * x[1] + q32 * x[3] + q42 * x[4] + q22 * x[2]);
    d[3] = dt * (q23 * x[2] + q43 * x[4] + q33 * x[3]);
    d[4] = dt * (q24 * x[2] + q34 * x[3] + q44 * x[4]);
}

// One Kv chain (Kv1.4 or Kv4.3): generator entries of update_QKv (GW.tpp:64-123)
// and the derivatives of update_Kv_derivatives (GW.tpp:135-160), in registers.
GW_HD void kv_chain(double aa, double ai, double ba, double bi, double f1, double f2, double f3, double f4,
                    double b1, double b2, double b3, double b4, const double y[10], double dt, double d[10])
{
    const double q01 = 4 * aa, q05 = bi, q00 = -(q01 + q05);
    const double q10 = ba, q12 = 3 * aa, q16 = f1 * bi, q11 = -(q10 + q12 + q16);
    const double q21 = 2 * ba, q23 = 2 * aa, q27 = f2 * bi, q22 = -(q21 + q23 + q27);
    const double q32 = 3 * ba, q34 = aa, q38 = f3 * bi, q33 = -(q32 + q34 + q38);
    const double q43 = 4 * ba, q49 = f4 * bi, q44 = -(q43 + q49);
    const double q56 = 4 * aa * b1, q50 = ai, q55 = -(q56 + q50);
    const double q65 = ba / f1, q67 = 3 * aa * b2 / b1, q61 = ai / b1, q66 = -(q65 + q67 + q61);
    const double q76 = 2 * ba * f1 / f2, q78 = 2 * aa * b3 / b2, q72 = ai / b2, q77 = -(q76 + q78 + q72);
    const double q87 = 3 * ba * f2 / f3, q89 = aa * b4 / b3, q83 = ai / b3, q88 = -(q87 + q89 + q83);
    const double q98 = 4 * ba * f3 / f4, q94 = ai / b4, q99 = -(q98 + q94);
    d[0] = dt * (q10 * y[1] + q50 * y[5] + q00 * y[0]);
    d[1] = dt * (q01 * y[0] + q21 * y[2] + q61 * y[6] + q11 * y[1]);
    d[2] = dt * (q12 * y[1] + q32 * y[3] + q72 * y[7] + q22 * y[2]);
    d[3] = dt * (q23 * y[2] + q43 * y[4] + q83 * y[8] + q33 * y[3]);
    d[4] = dt * (q34 * y[3] + q94 * y[9] + q44 * y[4]);
    d[5] = dt * (q65 * y[6] + q05 * y[0] + q55 * y[5]);
    d[6] = dt * (q56 * y[5] + q76 * y[7] + q16 * y[1] + q66 * y[6]);
    d[7] = dt * (q67 * y[6] + q87 * y[8] + q27 * y[2] + q77 * y[7]);
    d[8] = dt * (q78 * y[7] + q98 * y[9] + q38 * y[3] + q88 * y[8]);
    d[9] = dt * (q89 * y[8] + q49 * y[4] + q99 * y[9]);
}

// Currents of one step, in the order of GW::Currents (GW.hpp:25-42).
struct Currents {
    double INa, INab, INaCa, INaK, IKr, IKs, Ito1, Ito2, IK1, IKp, ICaL, ICab, IpCa, Jup, Jtr_tot, Jxfer_tot;
};

// The derivatives from the values X[slot] of gd_args' arguments, except the
// gates and the IKr and Kv chains (gd_gates, gd_chain_kr and gd_chain_kv
// below); the one exponential whose argument needs a logarithm (IK1) is
// evaluated here.
GW_HD int gd_finish_main(const gw_params_t &p, const CellConsts &c, const gw_globals_t &g, double VF_RT, const double *X,
                    double dt, double Istim, const FluxSums &S, gw_globals_t &dg, Currents &cur)
{
    const double V = g.V;
    const double expmVF_RT = X[GD_EXPM];
    bool ok = true;
    const double ENa = nernst(g.Nai, p.Nao, X[GD_LNA], c.RT_F, 1.0, ok);
    const double EK = nernst(g.Ki, p.Ko, X[GD_LK], c.RT_F, 1.0, ok);
    const double ECa = nernst(g.Ki, p.Ko, X[GD_LK], c.RT_F, 2.0, ok);  // K-based, as GW.tpp:176
    if (!ok) return GW_E_NONPOSITIVE_CONC;

    cur.INa = p.GNa * cube(g.m) * g.h * g.j * (V - ENa);                       // common.hpp:89-92
    cur.INab = p.GNab * (V - ENa);                                             // common.hpp:100-101
    {   // common.hpp:163-168
        const double e1 = X[GD_NACA];
        const double e2 = e1 * expmVF_RT;
        cur.INaCa = c.INaCa_const * (e1 * cube(g.Nai) * p.Cao - e2 * c.Nao3 * g.Cai) / (1.0 + p.ksat * e2);
    }
    {   // common.hpp:121-124, 140-143
        const double fNaK = 1.0 / (1.0 + 0.1245 * X[GD_NAK] + 0.0365 * c.sigma * expmVF_RT);
        cur.INaK = c.INaK_const * fNaK / (1.0 + X[GD_PNAK]);
    }
    {   // GW_utils.hpp:273-279
        const double RKr = 1.0 / (1.0 + 1.4945 * X[GD_KR_R]);
        cur.IKr = p.GKr * c.sqrtKo * RKr * g.Kr[3] * (V - EK) * 0.5;
    }
    {   // GW_utils.hpp:282-291
        const double EKs = X[GD_LKS] * c.RT_F;
        cur.IKs = p.GKs * sq(g.xKs) * (V - EKs);
    }
    double IKv14;
    {   // GW_utils.hpp:299-303 with exp_term = expmVF_RT (GW.tpp:185); limit at V == 0
        const double e = expmVF_RT;
        double mm;
        if (VF_RT == 0.0) mm = (c.PKv14_Csc) * kFaraday * g.Kv14[4];
        else mm = (c.PKv14_Csc) * (kFaraday * VF_RT) * g.Kv14[4] / (1 - e);
        IKv14 = mm * ((g.Ki - p.Ko * e) + 0.02 * (g.Nai - p.Nao * e)) * 1.0e9;
    }
    const double IKv43 = p.GKv43 * g.Kv43[4] * (V - EK);                       // GW_utils.hpp:294-297
    cur.Ito1 = IKv14 + IKv43;
    if (VF_RT == 0.0) cur.Ito2 = S.ClCh * c.Ito2_const * (p.Clo - p.Clcyto);    // limit of GW_utils.hpp:330
    else cur.Ito2 = S.ClCh * c.Ito2_const * VF_RT * (p.Clcyto * expmVF_RT - p.Clo) / (expmVF_RT - 1.0);
    {   // GW_utils.hpp:306-312
        const double K1inf = 1.0 / (2.0 + gw::exp(1.5 * (V - EK) * c.F_RT));
        cur.IK1 = p.GK1 * K1inf * c.IK1_const * (V - EK);
    }
    {   // GW_utils.hpp:315-319
        const double Kp = 1. / (1. + X[GD_KP]);
        cur.IKp = p.GKp * Kp * (V - EK);
    }
    cur.ICaL = S.JLCC * c.ICaL_const;                                           // GW_utils.hpp:322-325
    cur.ICab = p.GCab * (V - ECa);
    cur.IpCa = p.IpCamax * (g.Cai / (p.KmpCa + g.Cai));                         // common.hpp:109-110
    {   // GW_utils.hpp:339-344
        const double f = X[GD_PF];
        const double r = X[GD_PR];
        cur.Jup = (p.Vmaxf * f - p.Vmaxr * r) / (1.0 + f + r);
    }
    cur.Jtr_tot = S.Jtr * c.CRU_factor;                                         // GW_utils.hpp:354-355
    cur.Jxfer_tot = S.Jxfer * c.CRU_factor;
    double beta_cyto;
    {   // GW_utils.hpp:347-351
        const double x = p.KCMDN + g.Cai;
        beta_cyto = 1.0 / (1.0 + c.CMDN_const / sq(x));
    }
    // GW.tpp:201-211
    dg.CaLTRPN = p.kLTRPNp * g.Cai * (p.LTRPNtot - g.CaLTRPN) - p.kLTRPNm * g.CaLTRPN;
    dg.CaHTRPN = p.kHTRPNp * g.Cai * (p.HTRPNtot - g.CaHTRPN) - p.kHTRPNm * g.CaHTRPN;
    dg.Nai = -dt * c.CSA_FVcyto * (cur.INa + cur.INab + 3 * cur.INaCa + 3 * cur.INaK);
    dg.Ki = -dt * c.CSA_FVcyto * (cur.IKr + cur.IKs + cur.Ito1 + cur.IK1 + cur.IKp - 2 * cur.INaK);
    dg.Cai = dt * beta_cyto *
             (-0.5 * c.CSA_FVcyto * (cur.ICab + cur.IpCa - 2 * cur.INaCa) + c.VSS_Vcyto * cur.Jxfer_tot - cur.Jup -
              (dg.CaLTRPN + dg.CaHTRPN));
    dg.CaNSR = dt * (c.Vcyto_VNSR * cur.Jup - c.VJSR_VNSR * cur.Jtr_tot);
    dg.CaLTRPN *= dt;
    dg.CaHTRPN *= dt;
    dg.V = dt * (Istim - (cur.INa + cur.ICaL + cur.IKr + cur.IKs + cur.Ito1 + cur.IK1 + cur.IKp + cur.Ito2 +
                          cur.INaK + cur.INaCa + cur.IpCa + cur.ICab + cur.INab));
    return GW_OK;
}

// GW.tpp:162-168: the derivatives of the gates m, h, j and xKs, into
// d[0..3] (gw_globals_t order).
GW_HD void gd_gates(const double *X, const gw_globals_t &g, double dt, double d[4])
{
    const double V = g.V;
    d[0] = dt * (alpha_m(V, X) * (1.0 - g.m) - beta_m(X) * g.m);
    d[1] = dt * (alpha_h(V, X) * (1.0 - g.h) - beta_h(V, X) * g.h);
    d[2] = dt * (alpha_j(V, X) * (1.0 - g.j) - beta_j(V, X) * g.j);
    d[3] = dt * (XKsinf(X) - g.xKs) * tauXKs_inv(V, X);
}

// GW.tpp:25-44, 126-133: the IKr chain's derivatives.
GW_HD void gd_chain_kr(const gw_params_t &p, const double *X, const gw_globals_t &g, double dt, double dKr[5])
{
    kr_chain(p, X, g.Kr, dt, dKr);
}

// GW.tpp:64-123, 135-160: the derivatives of the Kv1.4 (which == 0) or Kv4.3
// (which == 1) chain.
GW_HD void gd_chain_kv(const gw_params_t &p, const double *X, const gw_globals_t &g, double dt, int which,
                       double dKv[10])
{
    if (which == 0) {
        const double aa14 = p.alphaa0Kv14 * X[GD_KV0];
        const double ai14 = p.alphai0Kv14 * X[GD_KV2];
        const double ba14 = p.betaa0Kv14 * X[GD_KV4];
        const double bi14 = p.betai0Kv14 * X[GD_KV6];
        kv_chain(aa14, ai14, ba14, bi14, p.f1Kv14, p.f2Kv14, p.f3Kv14, p.f4Kv14, p.b1Kv14, p.b2Kv14, p.b3Kv14,
                 p.b4Kv14, g.Kv14, dt, dKv);
    } else {
        const double aa43 = p.alphaa0Kv43 * X[GD_KV1];
        const double ai43 = p.alphai0Kv43 * X[GD_KV3];
        const double ba43 = p.betaa0Kv43 * X[GD_KV5];
        const double bi43 = p.betai0Kv43 * X[GD_KV7];
        kv_chain(aa43, ai43, ba43, bi43, p.f1Kv43, p.f2Kv43, p.f3Kv43, p.f4Kv43, p.b1Kv43, p.b2Kv43, p.b3Kv43,
                 p.b4Kv43, g.Kv43, dt, dKv);
    }
}

// All the derivatives from the values X[slot] of gd_args' arguments.
GW_HD int gd_finish(const gw_params_t &p, const CellConsts &c, const gw_globals_t &g, double VF_RT, const double *X,
                    double dt, double Istim, const FluxSums &S, gw_globals_t &dg, Currents &cur)
{
    const int rc = gd_finish_main(p, c, g, VF_RT, X, dt, Istim, S, dg, cur);
    if (rc != GW_OK) return rc;
    double gates[4];
    gd_gates(X, g, dt, gates);
    dg.m = gates[0];
    dg.h = gates[1];
    dg.j = gates[2];
    dg.xKs = gates[3];
    gd_chain_kr(p, X, g, dt, dg.Kr);
    gd_chain_kv(p, X, g, dt, 0, dg.Kv14);
    gd_chain_kv(p, X, g, dt, 1, dg.Kv43);
    return GW_OK;
}

// gd_finish with gd_args' functions evaluated one after another.
GW_HD int global_derivatives(const gw_params_t &p, const CellConsts &c, const gw_globals_t &g, double VF_RT,
                             double dt, double Istim, const FluxSums &S, gw_globals_t &dg, Currents &cur)
{
    double X[GD_N], py[GD_N - GD_POW0];
    gd_args(p, g, VF_RT, X, py);
    for (int i = 0; i < GD_N; ++i) X[i] = gd_eval(i, X[i], i >= GD_POW0 ? py[i - GD_POW0] : 0.0);
    return gd_finish(p, c, g, VF_RT, X, dt, Istim, S, dg, cur);
}

// globals += dGlobals (GW.tpp:252-271); V is held when clamped.
GW_HD void globals_apply(gw_globals_t &g, const gw_globals_t &dg, bool clamp)
{
    if (!clamp) g.V += dg.V;
    g.Nai += dg.Nai;
    g.Ki += dg.Ki;
    g.Cai += dg.Cai;
    g.CaNSR += dg.CaNSR;
    g.CaLTRPN += dg.CaLTRPN;
    g.CaHTRPN += dg.CaHTRPN;
    g.m += dg.m;
    g.h += dg.h;
    g.j += dg.j;
    g.xKs += dg.xKs;
    for (int i = 0; i < 5; ++i) g.Kr[i] += dg.Kr[i];
    for (int i = 0; i < 10; ++i) {
        g.Kv14[i] += dg.Kv14[i];
        g.Kv43[i] += dg.Kv43[i];
    }
}

}  // namespace gw
