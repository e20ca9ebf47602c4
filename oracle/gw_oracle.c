/*
 * gw_oracle.c -- CPU restatement of the Greenstein-Winslow coupled integrator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path. The product never
 * links, loads or calls this file. Built twice by oracle/Makefile:
 *   liboracle.so       strict IEEE (-O2 -ffp-contract=off): the bit-exact oracle
 *   liboracle_omp.so   reference release flags (-Ofast ... -fopenmp -funroll-loops
 *                      -flto, src/CMakeLists.txt:27-31): the timed CPU baseline
 *
 * Where the restatement departs from the reference, on purpose:
 *   D1 libm: exp/log/pow are this file's own (Cody-Waite reduction + Taylor
 *      polynomial; atanh series), not glibc, so the GPU can reproduce them
 *      bit for bit. They are within 2 ulp of glibc (tests/test_oracle_libm.py).
 *   D2 flux sums: sum(JLCC), sum(Jxfer), sum(Jtr) (src/includes/GW.tpp:188-198,
 *      Eigen .sum()) are formed per CRU in subunit order, then added exactly
 *      in 128-bit fixed point (2^-80 resolution) and rounded once. The result
 *      does not depend on summation order; Eigen's vectorised order cannot be
 *      reproduced without Eigen.
 *   D3 PRNG streams: one xoshiro stream per CRU, key = seed + (i << 32),
 *      instead of one std::random_device-seeded stream per OpenMP thread
 *      (src/includes/common.hpp:19-23). The draw order inside a CRU is the
 *      reference's.
 *   D4 removable singularities: at V == 0 exactly (JLCC multiplier, Ito2,
 *      IKv14) and V == 10 exactly (tauXKs_inv) the analytic limit is used
 *      where the reference divides 0/0.
 *   D5 Nernst with a non-positive argument returns an error status instead of
 *      exit() (src/includes/common.hpp:72-75).
 *
 * Pinning (how each part is checked beyond this restatement):
 *   xoshiro/splitmix/uniform map/sample_weights/sample_binomial -> golden vectors
 *       from the reference's own xoshiro.hpp and common.hpp compiled in oracle/_ref
 *   common.hpp currents and gates (INa, INaK, INaCa, IpCa, Ib, Nernst, m/h/j
 *       rates) -> golden vectors from oracle/_ref (<= 8 ulp, because of D1)
 *   libm -> numpy/glibc over sweeps (<= 2 ulp) and exact identities
 *   GW_utils.hpp gating functions (LCC alpha/beta/yinf/tau, XKsinf, tauXKs)
 *       -> special values (Boltzmann midpoints, e-fold points, limits and
 *       one-term asymptotes), tests/test_oracle_pins.py
 *   GW_utils.hpp currents (IKr, IK1, IKp, IKv43, IKv14, Jup) -> special values
 *       and limits; reversal potentials; GHK zero flux (test_oracle_pins.py,
 *       test_oracle_model.py)
 *   LCC / RyR rate tables -> the four-sensor structure entry by entry;
 *       detailed balance (Kolmogorov cycle criterion); the stationary LCC law
 *       at rest against the published rest sampler; CTMC statistics of the SSA
 *   Q matrices -> every rate at V = 0 and its voltage ratios; the Kv
 *       four-sensor structure; zero row sums; the published initial Kr/Kv
 *       vectors near the stationary law at the published rest potential
 *   global derivatives -> membrane charge = cytosolic ion charge; pump and
 *       exchanger stoichiometry; compartment Ca balance of the SR fluxes; the
 *       master equation of each chain; the published rest state near
 *       equilibrium; whole-cell Ca conservation during release
 *   SSA waiting times -> mean 1/total rate; stationary occupancy vs CTMC
 *   The pins catch planted mistakes: tests/test_oracle_mutations.py builds 20
 *   mutated copies of this file and requires a pin to fail on each.
 */
#include "gw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GWO_FARADAY 96.5      /* src/includes/common.hpp:9 */
#define GWO_GAS_CONST 8.314   /* src/includes/common.hpp:10 */

static inline uint64_t as_u64(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static inline double as_f64(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static inline uint64_t rotl64(uint64_t x, int s) { return (x << s) | (x >> (64 - s)); }

/* ======================================================================= */
/* PRNG                                                                    */
/* ======================================================================= */

/* SplitMix64::operator() (src/includes/xoshiro.hpp:684-690) */
uint64_t gwo_splitmix64_next(uint64_t *state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* Xoshiro256*(uint64 seed) / Xoroshiro128*(uint64 seed) constructors:
 * SplitMix64{seed}.generateSeedSequence<4 or 2>() (src/includes/xoshiro.hpp:935-936,
 * 729-730, 832-833, 1038-1039, 1127-1128, 1216-1217). */
void gwo_rng_seed(int prng_id, uint64_t s[4], uint64_t seed)
{
    uint64_t sm = seed;
    int n = (prng_id >= GW_PRNG_XOROSHIRO128SS) ? 2 : 4;
    s[0] = s[1] = s[2] = s[3] = 0;
    for (int i = 0; i < n; ++i) s[i] = gwo_splitmix64_next(&sm);
}

/* operator() of the six 64-bit engines (src/includes/xoshiro.hpp:735-746, 838-849,
 * 941-952, 1044-1053, 1133-1142, 1222-1231). */
uint64_t gwo_rng_next(int prng_id, uint64_t s[4])
{
    uint64_t result;
    if (prng_id <= GW_PRNG_XOSHIRO256P) {
        if (prng_id == GW_PRNG_XOSHIRO256SS) result = rotl64(s[1] * 5, 7) * 9;
        else if (prng_id == GW_PRNG_XOSHIRO256PP) result = rotl64(s[0] + s[3], 23) + s[0];
        else result = s[0] + s[3];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl64(s[3], 45);
        return result;
    }
    const uint64_t s0 = s[0];
    uint64_t s1 = s[1];
    if (prng_id == GW_PRNG_XOROSHIRO128SS) {
        result = rotl64(s0 * 5, 7) * 9;
        s1 ^= s0;
        s[0] = rotl64(s0, 24) ^ s1 ^ (s1 << 16);
        s[1] = rotl64(s1, 37);
    } else if (prng_id == GW_PRNG_XOROSHIRO128PP) {
        result = rotl64(s0 + s1, 17) + s0;
        s1 ^= s0;
        s[0] = rotl64(s0, 49) ^ s1 ^ (s1 << 21);
        s[1] = rotl64(s1, 28);
    } else {
        result = s0 + s1;
        s1 ^= s0;
        s[0] = rotl64(s0, 24) ^ s1 ^ (s1 << 16);
        s[1] = rotl64(s1, 37);
    }
    return result;
}

/* urand (src/includes/common.hpp:25-29): uniform_real_distribution<double>(0,1)
 * over generate_canonical<double,53> (libstdc++ random.tcc:3348-3380 of GCC 11):
 * u = double(x) / 2^64 with one round-to-nearest, clamped below 1. */
double gwo_u01(uint64_t x)
{
    double u = (double)x * 0x1p-64;
    if (u >= 1.0) u = 0x1.fffffffffffffp-1;
    return u;
}

uint64_t gwo_stream_key(uint64_t seed, int64_t i) { return seed + ((uint64_t)i << 32); }

/* ======================================================================= */
/* libm (D1)                                                               */
/* ======================================================================= */
#define LN2_HI 0x1.62e42fefa39efp-1
#define LN2_LO 0x1.abc9e3b39803fp-56
#define INV_LN2 0x1.71547652b82fep+0
#define SQRT2 0x1.6a09e667f3bcdp+0

static inline double pow2i(int k) { return as_f64((uint64_t)(k + 1023) << 52); }

/* exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, then the degree-13
 * Taylor polynomial of e^r in Horner form with fused multiply-adds. */
double gwo_exp(double x)
{
    if (x != x) return x;
    if (x > 709.79) return INFINITY;
    if (x < -745.2) return 0.0;
    const double kd = nearbyint(x * INV_LN2);
    double r = fma(-kd, LN2_HI, x);
    r = fma(-kd, LN2_LO, r);
    double p = 0x1.6124613a86d09p-33;      /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);  /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);  /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);  /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);  /* 1/9!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);  /* 1/8!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);  /* 1/7!  */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);  /* 1/6!  */
    p = fma(p, r, 0x1.1111111111111p-7);   /* 1/5!  */
    p = fma(p, r, 0x1.5555555555555p-5);   /* 1/4!  */
    p = fma(p, r, 0x1.5555555555555p-3);   /* 1/3!  */
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const int k = (int)kd;
    if (k > 1023) return (p * pow2i(1023)) * pow2i(k - 1023);
    if (k < -1021) return (p * pow2i(k + 1000)) * pow2i(-1000);
    return p * pow2i(k);
}

/* log: x = 2^e m, m in [sqrt(2)/2, sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
 * atanh series to s^25 in Horner form with fused multiply-adds. */
double gwo_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    uint64_t u = as_u64(x);
    int e = (int)((u >> 52) & 0x7ff);
    if (e == 0) {
        u = as_u64(x * 0x1p54);
        e = (int)((u >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = as_f64((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > SQRT2) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.47ae147ae147bp-5;       /* 1/25 */
    R = fma(R, z, 0x1.642c8590b2164p-5);   /* 1/23 */
    R = fma(R, z, 0x1.8618618618618p-5);   /* 1/21 */
    R = fma(R, z, 0x1.af286bca1af28p-5);   /* 1/19 */
    R = fma(R, z, 0x1.e1e1e1e1e1e1ep-5);   /* 1/17 */
    R = fma(R, z, 0x1.1111111111111p-4);   /* 1/15 */
    R = fma(R, z, 0x1.3b13b13b13b14p-4);   /* 1/13 */
    R = fma(R, z, 0x1.745d1745d1746p-4);   /* 1/11 */
    R = fma(R, z, 0x1.c71c71c71c71cp-4);   /* 1/9  */
    R = fma(R, z, 0x1.2492492492492p-3);   /* 1/7  */
    R = fma(R, z, 0x1.999999999999ap-3);   /* 1/5  */
    R = fma(R, z, 0x1.5555555555555p-2);   /* 1/3  */
    R = R * z;
    const double s2 = 2.0 * s;
    const double lm = fma(s2, R, s2);
    const double ed = (double)e;
    return fma(ed, LN2_HI, fma(ed, LN2_LO, lm));
}

double gwo_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return NAN;
    if (x < 0.0) return NAN;
    if (x == 0.0) return (y > 0.0) ? 0.0 : INFINITY;
    return gwo_exp(y * gwo_log(x));
}

/* ======================================================================= */
/* exact flux accumulator (D2)                                             */
/* ======================================================================= */
typedef unsigned __int128 u128;
typedef __int128 s128;

void gwo_fx_add(gwo_fx *a, double x, int *flag)
{
    const uint64_t u = as_u64(x);
    const int E = (int)((u >> 52) & 0x7ff);
    if (E == 0) return;                      /* zero / subnormal: below 2^-80 */
    if (E == 0x7ff) { *flag |= 1; return; }  /* NaN / Inf */
    if (E >= 1047) { *flag |= 2; return; }   /* |x| >= 2^24 */
    const uint64_t M = (u & 0x000fffffffffffffULL) | (1ULL << 52);
    const int sh = E - 995;                  /* x * 2^80 = M * 2^(E - 1075 + 80) */
    u128 mag;
    if (sh >= 0) mag = ((u128)M) << sh;
    else mag = (sh > -53) ? (u128)(M >> (-sh)) : 0;
    u128 acc = ((u128)a->hi << 64) | a->lo;
    acc = (u >> 63) ? acc - mag : acc + mag;
    a->lo = (uint64_t)acc;
    a->hi = (uint64_t)(acc >> 64);
}

double gwo_fx_to_double(const gwo_fx *a)
{
    const s128 s = (s128)(((u128)a->hi << 64) | a->lo);
    if (s == 0) return 0.0;
    const int neg = s < 0;
    u128 mag = neg ? (u128)(-s) : (u128)s;
    int p = 127;
    while (!((mag >> p) & 1)) --p;
    double d;
    if (p <= 52) {
        d = ldexp((double)(uint64_t)mag, -80);
    } else {
        int shift = p - 52;
        uint64_t m = (uint64_t)(mag >> shift);
        const u128 rem = mag & ((((u128)1) << shift) - 1);
        const u128 half = ((u128)1) << (shift - 1);
        if (rem > half || (rem == half && (m & 1))) {
            ++m;
            if (m == (1ULL << 53)) { m >>= 1; ++shift; }
        }
        d = ldexp((double)m, shift - 80);
    }
    return neg ? -d : d;
}

/* ======================================================================= */
/* sampling helpers (src/includes/common.hpp:37-59)                        */
/* ======================================================================= */

/* sample_weights given the uniform draw (src/includes/common.hpp:49-59). */
int gwo_sample_weights(const double *w, double total, int n, double u01)
{
    const double u = u01 * total;
    double cum = w[0];
    int i = 0;
    while (cum < u && i < (n - 1)) {
        ++i;
        cum += w[i];
    }
    return i;
}

/* ======================================================================= */
/* point functions (src/includes/common.hpp:70-218, GW_utils.hpp:273-375)  */
/* ======================================================================= */
static inline double sq(double x) { return x * x; }          /* common.hpp:12-13 */
static inline double cube(double x) { return x * x * x; }    /* common.hpp:15-16 */

/* common::Nernst (common.hpp:70-77); D5 */
double gwo_Nernst(double Xi, double Xo, double RT_F, double z, int *ok)
{
    if (Xi <= 0.0 || Xo <= 0.0) { if (ok) *ok = 0; return NAN; }
    return gwo_log(Xo / Xi) * RT_F / z;
}
/* common::INa (common.hpp:89-92) */
double gwo_INa(double V, double m, double h, double j, double GNa, double ENa)
{ return GNa * cube(m) * h * j * (V - ENa); }
/* common::Ib (common.hpp:100-101) */
double gwo_Ib(double V, double Gx, double Ex) { return Gx * (V - Ex); }
/* common::IpCa (common.hpp:109-110) */
double gwo_IpCa(double Cai, double IpCamax, double KmpCa) { return IpCamax * (Cai / (KmpCa + Cai)); }
/* common::fNaK, expmVF_RT form (common.hpp:121-124) */
double gwo_fNaK(double VF_RT, double expmVF_RT, double sigma)
{ return 1.0 / (1.0 + 0.1245 * gwo_exp(-0.1 * VF_RT) + 0.0365 * sigma * expmVF_RT); }
/* common::INaK, expmVF_RT form (common.hpp:140-143) */
double gwo_INaK(double VF_RT, double expmVF_RT, double Nai, double sigma, double KmNai, double mult)
{ return mult * gwo_fNaK(VF_RT, expmVF_RT, sigma) / (1.0 + gwo_pow(KmNai / Nai, 1.5)); }
/* common::INaCa, expmVF_RT form (common.hpp:163-168) */
double gwo_INaCa(double VF_RT, double expmVF_RT, double Nai, double Cai, double Nao3, double Cao,
                 double eta, double ksat, double mult)
{
    const double e1 = gwo_exp(eta * VF_RT);
    const double e2 = e1 * expmVF_RT;
    return mult * (e1 * cube(Nai) * Cao - e2 * Nao3 * Cai) / (1.0 + ksat * e2);
}
/* common::alpha_m .. beta_j (common.hpp:175-218) */
double gwo_alpha_m(double V)
{ return (V == -47.13) ? 3.2 : 0.32 * (V + 47.13) / (1.0 - gwo_exp(-0.1 * (V + 47.13))); }
double gwo_beta_m(double V) { return 0.08 * gwo_exp(-V / 11.0); }
double gwo_alpha_h(double V) { return (V >= -40.0) ? 0.0 : 0.135 * gwo_exp(-(V + 80.0) / 6.8); }
double gwo_beta_h(double V)
{
    return (V >= -40.0) ? 1.0 / (0.13 * (1.0 + gwo_exp(-(V + 10.66) / 11.1)))
                        : 3.56 * gwo_exp(0.079 * V) + 3.1e5 * gwo_exp(0.35 * V);
}
double gwo_alpha_j(double V)
{
    return (V >= -40.0) ? 0.0
        : (-127140.0 * gwo_exp(0.2444 * V) - 3.474e-5 * gwo_exp(-0.04391 * V)) * (V + 37.78)
              / (1.0 + gwo_exp(0.311 * (V + 79.23)));
}
double gwo_beta_j(double V)
{
    return (V >= -40.0) ? 0.3 * gwo_exp(-2.535e-7 * V) / (1.0 + gwo_exp(-0.1 * (V + 32.0)))
                        : 0.1212 * gwo_exp(-0.01052 * V) / (1.0 + gwo_exp(-0.1378 * (V + 40.14)));
}
/* GW::XKsinf, tauXKs_inv (GW_utils.hpp:358-362); D4 at V == 10 */
double gwo_XKsinf(double V) { return 1.0 / (1.0 + gwo_exp(-(V - 24.7) / 13.6)); }
double gwo_tauXKs_inv(double V)
{
    if (V == 10.0) return 0.0000719 / 0.148 + 0.000131 / 0.0687;
    return 0.0000719 * (V - 10.0) / (1.0 - gwo_exp(-0.148 * (V - 10.0)))
         + 0.000131 * (V - 10.0) / (gwo_exp(0.0687 * (V - 10.0)) - 1.0);
}
/* GW::alphaLCC .. tauLCC (GW_utils.hpp:365-375) */
double gwo_alphaLCC(double V) { return 2.0 * gwo_exp(0.012 * (V - 35.0)); }
double gwo_betaLCC(double V) { return 0.0882 * gwo_exp(-0.05 * (V - 35.0)); }
double gwo_yinfLCC(double V) { return 0.4 / (1.0 + gwo_exp((V + 12.5) / 5.0)) + 0.6; }
double gwo_tauLCC(double V) { return 340.0 / (1.0 + gwo_exp((V + 30.0) / 12.0)) + 60.0; }

/* ======================================================================= */
/* Constants (src/includes/GW_utils.hpp:162-227, GW_utils.tpp:67-116)     */
/* ======================================================================= */
typedef struct {
    double RT_F, F_RT, CRU_factor, CSA_FVcyto, VSS_Vcyto, Vcyto_VNSR, VJSR_VNSR;
    double a, gamma0, gamma0a, gamma0a2, gamma0a3, gamma0a4, binv, omega;
    double omega_b, omega_b2, omega_b3, omega_b4;
    double BSR_const, BSL_const, VSS_VJSR, CSQN_const, JLCC_const, Cao_scaled;
    double Nao3, INaCa_const, sigma, INaK_const, sqrtKo, PKv14_Csc, Ito2_const;
    double IK1_const, ICaL_const, CMDN_const;
    double VF_RT, expmVF_RT, ENa, EK, ECa, beta_cyto, JLCC_multiplier, JLCC_exp;
    double alphaLCC, betaLCC, yinfLCC, tauLCC;
} gwo_consts;

static void consts_init(gwo_consts *c, const gw_params_t *p, int nCRU)
{
    memset(c, 0, sizeof(*c));
    c->RT_F = GWO_GAS_CONST * p->T / GWO_FARADAY;
    c->F_RT = 1.0 / c->RT_F;
    c->CRU_factor = (double)p->NCaRU / (double)nCRU;
    c->CSA_FVcyto = p->CSA / (1000 * p->Vcyto * GWO_FARADAY);
    c->VSS_Vcyto = p->VSS / p->Vcyto;
    c->Vcyto_VNSR = p->Vcyto / p->VNSR;
    c->VJSR_VNSR = p->VJSR / p->VNSR;
    c->a = p->a;
    c->gamma0 = p->gamma0;
    c->gamma0a = c->gamma0 * p->a;
    c->gamma0a2 = c->gamma0a * p->a;
    c->gamma0a3 = c->gamma0a2 * p->a;
    c->gamma0a4 = c->gamma0a3 * p->a;
    c->binv = 1.0 / p->b;
    c->omega = p->omega;
    c->omega_b = c->omega * c->binv;
    c->omega_b2 = c->omega_b * c->binv;
    c->omega_b3 = c->omega_b2 * c->binv;
    c->omega_b4 = c->omega_b3 * c->binv;
    c->BSR_const = p->KBSR * p->BSRT;
    c->BSL_const = p->KBSL * p->BSLT;
    c->VSS_VJSR = p->VSS / p->VJSR;
    c->CSQN_const = p->KCSQN * p->CSQNT;
    c->JLCC_const = 2.0e6 * p->PCaL / p->VSS;
    c->Cao_scaled = 0.341 * p->Cao;
    c->Nao3 = p->Nao * p->Nao * p->Nao;
    c->INaCa_const = 5000.0 * p->kNaCa / ((p->KmNa * p->KmNa * p->KmNa + p->Nao * p->Nao * p->Nao) * (p->KmCa + p->Cao));
    c->sigma = (gwo_exp(p->Nao / 67.3) - 1.0) / 7.0;
    c->INaK_const = p->INaKmax * p->Ko / (p->Ko + p->KmKo);
    c->sqrtKo = sqrt(p->Ko);
    c->PKv14_Csc = p->PKv14 / p->Csc;
    c->Ito2_const = 1.0e9 * p->Pto2 * GWO_FARADAY * c->CRU_factor / p->CSA;
    c->IK1_const = p->Ko / (p->Ko + p->KmK1);
    c->ICaL_const = -1000.0 * (2.0 * GWO_FARADAY * p->VSS) * c->CRU_factor / p->CSA;
    c->CMDN_const = p->KCMDN * p->CMDNT;
}

/* Per-step LCC constants of GW_model::SSA (src/includes/GW.tpp:216-221); D4 at V == 0. */
static void ssa_consts(gwo_consts *c, double V)
{
    c->alphaLCC = gwo_alphaLCC(V);
    c->betaLCC = gwo_betaLCC(V);
    c->yinfLCC = gwo_yinfLCC(V);
    c->tauLCC = gwo_tauLCC(V);
    c->JLCC_exp = sq(1.0 / c->expmVF_RT);
    if (V == 0.0) c->JLCC_multiplier = c->JLCC_const / 2.0;
    else c->JLCC_multiplier = c->JLCC_const * V * c->F_RT / (c->JLCC_exp - 1);
}

/* ======================================================================= */
/* CRU working copy and SSA (src/includes/SSA.hpp:15-42, SSA.tpp)         */
/* ======================================================================= */
typedef struct {
    int LCC[4];
    int LCC_inactivation[4];
    int RyR[24];
    double open_RyR[4];
    int ClCh[4];
    double CaSS[4];
    double CaJSR;
    double JLCC[4];
    double Jrel[4];
    double Jxfer[4];
    double Jiss[4];
    double Jtr;
    double LCC_rates[12];
    double LCC_inactivation_rates[4];
    double RyR_rates[48];
    double ClCh_rates[4];
    double subunit_rates[4];
    /* stream and replay digest of this CRU */
    uint64_t *rng;
    int prng;
    uint64_t ev_count, ev_hash;
    uint32_t *ev_log;
    int ev_cap;
    uint32_t step_lo;
    uint64_t iters;
    const uint64_t *replay;   /* test hook: raw draws replayed instead of the stream */
    int replay_pos;
} cru_tmp;

static inline double urand(cru_tmp *s)
{
    if (s->replay) return gwo_u01(s->replay[s->replay_pos++]);
    return gwo_u01(gwo_rng_next(s->prng, s->rng));
}

/* Replay digest: FNV-1a over 32-bit event codes. */
enum { EV_LCC = 0, EV_Y = 1, EV_RYR = 2, EV_CLCH = 3, EV_SPLIT56 = 4, EV_SPLIT34 = 5 };
static inline void ev_record(cru_tmp *s, int subunit, int kind, int detail)
{
    const uint32_t code = (uint32_t)subunit | ((uint32_t)kind << 2) | (((uint32_t)detail & 15u) << 5)
                        | (s->step_lo << 16);
    if (s->ev_count < (uint64_t)s->ev_cap) s->ev_log[s->ev_count] = code;
    s->ev_count++;
    s->ev_hash = (s->ev_hash ^ code) * 0x100000001b3ULL;
}

/* sample_binomial (common.hpp:37-47) */
static int sample_binomial(cru_tmp *s, double p, int N)
{
    int X = 0;
    for (int i = 0; i < N; i++) {
        const double u = urand(s);
        if (u < p) X++;
    }
    return X;
}

/* sample_binomial over replayed raw draws (test hook); returns draws consumed. */
int gwo_sample_binomial_replay(double p, int N, const uint64_t *raw, int *X)
{
    cru_tmp s;
    memset(&s, 0, sizeof(s));
    s.replay = raw;
    *X = sample_binomial(&s, p, N);
    return s.replay_pos;
}

/* update_LCC_rates (GW_utils.tpp:118-186) */
static void update_LCC_rates(double *LCC_rates, double *subunit_rates, const int *LCC, const double *CaSS,
                             int idx, const gw_params_t *p, const gwo_consts *c)
{
    double *r = LCC_rates + 3 * idx;
    switch (LCC[idx]) {
    case 1: r[0] = 4 * c->alphaLCC; r[1] = c->gamma0 * CaSS[idx]; r[2] = 0; break;
    case 2: r[0] = c->betaLCC; r[1] = 3 * c->alphaLCC; r[2] = c->gamma0a * CaSS[idx]; break;
    case 3: r[0] = 2 * c->betaLCC; r[1] = 2 * c->alphaLCC; r[2] = c->gamma0a2 * CaSS[idx]; break;
    case 4: r[0] = 3 * c->betaLCC; r[1] = c->alphaLCC; r[2] = c->gamma0a3 * CaSS[idx]; break;
    case 5: r[0] = 4 * c->betaLCC; r[1] = p->f; r[2] = c->gamma0a4 * CaSS[idx]; break;
    case 6: r[0] = p->g; r[1] = 0; r[2] = 0; break;
    case 7: r[0] = c->omega; r[1] = 4 * c->alphaLCC * c->a; r[2] = 0; break;
    case 8: r[0] = c->omega_b; r[1] = c->betaLCC * c->binv; r[2] = 3 * c->alphaLCC * c->a; break;
    case 9: r[0] = c->omega_b2; r[1] = 2 * c->betaLCC * c->binv; r[2] = 2 * c->alphaLCC * c->a; break;
    case 10: r[0] = c->omega_b3; r[1] = 3 * c->betaLCC * c->binv; r[2] = c->alphaLCC * c->a; break;
    case 11: r[0] = c->omega_b4; r[1] = 4 * c->betaLCC * c->binv; r[2] = p->f1; break;
    case 12: r[0] = p->g1; r[1] = 0; r[2] = 0; break;
    default: break;
    }
    subunit_rates[idx] += (r[0] + r[1] + r[2]);
}

/* update_LCC_inactivation_rates (GW_utils.hpp:378-383) */
static void update_LCC_inactivation_rates(double *rates, double *subunit_rates, const int *y,
                                          double yinf, double tau, int idx)
{
    rates[idx] = (y[idx] == 0) ? yinf / tau : (1.0 - yinf) / tau;
    subunit_rates[idx] += rates[idx];
}

/* update_ClCh_rates (GW_utils.hpp:386-391) */
static void update_ClCh_rates(double *rates, double *subunit_rates, const int *ClCh, const double *CaSS,
                              double kf, double kb, int idx)
{
    rates[idx] = (ClCh[idx] == 0) ? kf * CaSS[idx] : kb;
    subunit_rates[idx] += rates[idx];
}

/* update_RyR_rates (GW_utils.tpp:222-255) */
static void update_RyR_rates(double *R, double *subunit_rates, const int *RyR, const double *CaSS,
                             int idx, const gw_params_t *p)
{
    const double CaSS2 = CaSS[idx] * CaSS[idx];
    const double eq56 = p->k65 / (p->k56 * CaSS2 + p->k65);
    const double tau34 = 1.0 / (p->k34 * CaSS2 + p->k43);
    const int *n = RyR + 6 * idx;
    double *r = R + 12 * idx;
    r[0] = n[0] * p->k12 * CaSS2;
    r[1] = n[1] * p->k23 * CaSS2;
    r[2] = n[1] * p->k25 * CaSS2;
    r[3] = CaSS[idx] > 3.685e-2 ? 0 : n[2] * p->k34 * CaSS2;
    r[4] = CaSS[idx] > 3.685e-2 ? (n[2] + n[3]) * p->k45 * p->k34 * CaSS2 * tau34 : n[3] * p->k45;
    r[5] = CaSS[idx] > 1.15e-4 ? 0 : n[4] * p->k56 * CaSS2;
    r[6] = n[1] * p->k21;
    r[7] = CaSS[idx] > 3.685e-2 ? (n[2] + n[3]) * p->k32 * p->k43 * tau34 : n[2] * p->k32;
    r[8] = CaSS[idx] > 3.685e-2 ? 0 : n[3] * p->k43;
    r[9] = CaSS[idx] > 1.15e-4 ? (n[4] + n[5]) * p->k52 * eq56 : n[4] * p->k52;
    r[10] = CaSS[idx] > 1.15e-4 ? (n[4] + n[5]) * p->k54 * CaSS2 * eq56 : n[4] * p->k54 * CaSS2;
    r[11] = CaSS[idx] > 1.15e-4 ? 0 : n[5] * p->k65;
    for (int k = 0; k < 12; k++) subunit_rates[idx] += r[k];
}

/* copy_from_CRUState (SSA.tpp:10-33) */
static void copy_from_CRUState(cru_tmp *s, const gwo_model *m, int idx);

/* update_fluxes (SSA.tpp:37-45) */
static void update_fluxes(cru_tmp *s, double Cai, double CaNSR, const gw_params_t *p)
{
    s->Jtr = p->rtr * (CaNSR - s->CaJSR);
    s->Jiss[0] = p->riss * (s->CaSS[1] + s->CaSS[3] - 2 * s->CaSS[0]);
    s->Jiss[1] = p->riss * (s->CaSS[2] + s->CaSS[0] - 2 * s->CaSS[1]);
    s->Jiss[2] = p->riss * (s->CaSS[3] + s->CaSS[1] - 2 * s->CaSS[2]);
    s->Jiss[3] = p->riss * (s->CaSS[0] + s->CaSS[2] - 2 * s->CaSS[3]);
    for (int j = 0; j < 4; j++) s->Jxfer[j] = p->rxfer * (s->CaSS[j] - Cai);
}

/* update_rates (SSA.tpp:48-56) */
static void update_rates(cru_tmp *s, const gw_params_t *p, const gwo_consts *c)
{
    for (int j = 0; j < 4; j++) {
        s->subunit_rates[j] = 0;
        update_LCC_rates(s->LCC_rates, s->subunit_rates, s->LCC, s->CaSS, j, p, c);
        update_LCC_inactivation_rates(s->LCC_inactivation_rates, s->subunit_rates, s->LCC_inactivation,
                                      c->yinfLCC, c->tauLCC, j);
        update_RyR_rates(s->RyR_rates, s->subunit_rates, s->RyR, s->CaSS, j, p);
        update_ClCh_rates(s->ClCh_rates, s->subunit_rates, s->ClCh, s->CaSS, p->kfClCh, p->kbClCh, j);
    }
}

/* sample_RyR56 / sample_RyR34 (SSA.tpp:59-84) */
static void sample_RyR56(cru_tmp *s, int idx, const gw_params_t *p)
{
    const int n56 = s->RyR[4 + 6 * idx] + s->RyR[5 + 6 * idx];
    if (n56 > 0) {
        const double pr = p->k65 / (p->k65 + p->k56 * sq(s->CaSS[idx]));
        s->RyR[4 + 6 * idx] = sample_binomial(s, pr, n56);
        s->RyR[5 + 6 * idx] = n56 - s->RyR[4 + 6 * idx];
    } else {
        s->RyR[4 + 6 * idx] = 0;
        s->RyR[5 + 6 * idx] = 0;
    }
    ev_record(s, idx, EV_SPLIT56, s->RyR[4 + 6 * idx]);
}
static void sample_RyR34(cru_tmp *s, int idx, const gw_params_t *p)
{
    const int n34 = s->RyR[2 + 6 * idx] + s->RyR[3 + 6 * idx];
    if (n34 > 0) {
        const double pr = p->k43 / (p->k43 + p->k34 * sq(s->CaSS[idx]));
        s->RyR[2 + 6 * idx] = sample_binomial(s, pr, n34);
        s->RyR[3 + 6 * idx] = n34 - s->RyR[2 + 6 * idx];
    } else {
        s->RyR[2 + 6 * idx] = 0;
        s->RyR[3 + 6 * idx] = 0;
    }
    ev_record(s, idx, EV_SPLIT34, s->RyR[2 + 6 * idx]);
}

/* update_CaSS (SSA.tpp:87-99) */
static void update_CaSS(cru_tmp *s, double dt, const gw_params_t *p, const gwo_consts *c)
{
    double betaSS, dCaSS, CaSS_tmp;
    for (int j = 0; j < 4; j++) {
        betaSS = 1.0 / (1 + (c->BSR_const / sq(p->KBSR + s->CaSS[j])) + (c->BSL_const / sq(p->KBSL + s->CaSS[j])));
        dCaSS = dt * (s->JLCC[j] + s->Jrel[j] - s->Jxfer[j] + s->Jiss[j]) * betaSS;
        CaSS_tmp = s->CaSS[j] + dCaSS;
        if (s->CaSS[j] > 1.15e-4 && CaSS_tmp <= 1.15e-4)
            sample_RyR56(s, j, p);
        else if (s->CaSS[j] > 0.03685 && CaSS_tmp <= 0.03685)
            sample_RyR34(s, j, p);
        s->CaSS[j] = CaSS_tmp;
    }
}

/* update_CaJSR (SSA.hpp:67-71) */
static void update_CaJSR(cru_tmp *s, double dt, const gw_params_t *p, const gwo_consts *c)
{
    const double betaJSR = 1.0 / (1 + c->CSQN_const / sq(p->KCSQN + s->CaJSR));
    s->CaJSR += (dt * betaJSR * (s->Jtr - c->VSS_VJSR * (s->Jrel[0] + s->Jrel[1] + s->Jrel[2] + s->Jrel[3])));
}

/* sample_LCC (SSA.tpp:103-197) */
static void sample_LCC(cru_tmp *s, double sum_LCC_rates, int j, const gwo_consts *c)
{
    const int tr = gwo_sample_weights(s->LCC_rates + 3 * j, sum_LCC_rates, 3, urand(s));
    int *L = &s->LCC[j];
    switch (*L) {
    case 1: *L = (tr == 0) ? 2 : 7; break;
    case 2: *L = (tr == 0) ? 1 : (tr == 1) ? 3 : 8; break;
    case 3: *L = (tr == 0) ? 2 : (tr == 1) ? 4 : 9; break;
    case 4: *L = (tr == 0) ? 3 : (tr == 1) ? 5 : 10; break;
    case 5:
        if (tr == 0) *L = 4;
        else if (tr == 1) {
            *L = 6;
            s->JLCC[j] = (s->LCC_inactivation[j] == 1) ? c->JLCC_multiplier * (c->Cao_scaled - c->JLCC_exp * s->CaSS[j]) : 0;
        } else *L = 11;
        break;
    case 6: *L = 5; s->JLCC[j] = 0; break;
    case 7: *L = (tr == 0) ? 1 : 8; break;
    case 8: *L = (tr == 0) ? 2 : (tr == 1) ? 7 : 9; break;
    case 9: *L = (tr == 0) ? 3 : (tr == 1) ? 8 : 10; break;
    case 10: *L = (tr == 0) ? 4 : (tr == 1) ? 9 : 11; break;
    case 11:
        if (tr == 0) *L = 5;
        else if (tr == 1) *L = 10;
        else {
            *L = 12;
            s->JLCC[j] = (s->LCC_inactivation[j] == 1) ? c->JLCC_multiplier * (c->Cao_scaled - c->JLCC_exp * s->CaSS[j]) : 0.0;
        }
        break;
    case 12: *L = 11; s->JLCC[j] = 0.0; break;
    default: break;
    }
    ev_record(s, j, EV_LCC, *L);
}

/* sample_RyR (SSA.tpp:200-259) */
static void sample_RyR(cru_tmp *s, double sum_RyR_rates, int j, const gw_params_t *p)
{
    const int tr = gwo_sample_weights(s->RyR_rates + 12 * j, sum_RyR_rates, 12, urand(s));
    int *n = s->RyR + 6 * j;
    switch (tr) {
    case 0: --n[0]; ++n[1]; break;
    case 1: --n[1]; ++n[2]; ++s->open_RyR[j]; break;
    case 2: --n[1]; ++n[4]; break;
    case 3: --n[2]; ++n[3]; break;
    case 4: --n[3]; ++n[4]; --s->open_RyR[j]; break;
    case 5: --n[4]; ++n[5]; break;
    case 6: --n[1]; ++n[0]; break;
    case 7: --n[2]; ++n[1]; --s->open_RyR[j]; break;
    case 8: --n[3]; ++n[2]; break;
    case 9: --n[4]; ++n[1]; break;
    case 10: --n[4]; ++n[3]; ++s->open_RyR[j]; break;
    case 11: --n[5]; ++n[4]; break;
    default: break;
    }
    s->Jrel[j] = p->rRyR * s->open_RyR[j] * (s->CaJSR - s->CaSS[j]);
    ev_record(s, j, EV_RYR, tr);
}

/* sample_new_state (SSA.tpp:262-283) */
static void sample_new_state(cru_tmp *s, int j, const gw_params_t *p, const gwo_consts *c)
{
    const double subunit_rate = s->subunit_rates[j];
    const double sum_LCC_rates = s->LCC_rates[3 * j] + s->LCC_rates[3 * j + 1] + s->LCC_rates[3 * j + 2];
    const double inact = s->LCC_inactivation_rates[j];
    const double clch = s->ClCh_rates[j];
    const double sum_RyR_rates = subunit_rate - (sum_LCC_rates + inact + clch);
    const double u = urand(s) * subunit_rate;
    if (u < sum_LCC_rates)
        sample_LCC(s, sum_LCC_rates, j, c);
    else if (u < (sum_LCC_rates + inact)) {
        s->LCC_inactivation[j] = 1 - s->LCC_inactivation[j];
        if (s->LCC_inactivation[j] == 0)
            s->JLCC[j] = 0;
        else if ((s->LCC[j] == 6) || (s->LCC[j] == 12))
            s->JLCC[j] = c->JLCC_multiplier * (c->Cao_scaled - c->JLCC_exp * s->CaSS[j]);
        ev_record(s, j, EV_Y, s->LCC_inactivation[j]);
    } else if (u < (sum_LCC_rates + inact + sum_RyR_rates))
        sample_RyR(s, sum_RyR_rates, j, p);
    else {
        s->ClCh[j] = 1 - s->ClCh[j];
        ev_record(s, j, EV_CLCH, s->ClCh[j]);
    }
}

/* SSA_single_CRU (SSA.tpp:286-310) */
static void SSA_single_CRU(cru_tmp *s, double Cai, double CaNSR, double dt, const gw_params_t *p, const gwo_consts *c)
{
    int j;
    double t = 0, jump_t = 0, total_rate;
    while (1) {
        s->iters++;
        update_fluxes(s, Cai, CaNSR, p);
        update_rates(s, p, c);
        total_rate = s->subunit_rates[0] + s->subunit_rates[1] + s->subunit_rates[2] + s->subunit_rates[3];
        jump_t = -gwo_log(urand(s)) / total_rate;
        if (t + jump_t < dt) {
            update_CaSS(s, jump_t, p, c);
            update_CaJSR(s, jump_t, p, c);
        } else {
            update_CaSS(s, dt - t, p, c);
            update_CaJSR(s, dt - t, p, c);
            break;
        }
        t += jump_t;
        j = gwo_sample_weights(s->subunit_rates, total_rate, 4, urand(s));
        sample_new_state(s, j, p, c);
    }
}

/* ======================================================================= */
/* the model (src/includes/GW.hpp:45-124, GW.tpp)                          */
/* ======================================================================= */
struct gwo_model {
    gw_params_t p;
    gwo_consts c;
    gw_globals_t g, dg;
    int n, prng;
    uint64_t seed;
    double *CaSS, *CaJSR, *JLCC, *Jxfer, *Jtr;
    int *LCC, *LCCi, *RyR, *ClCh;
    uint64_t *rng;
    uint64_t *ev_count, *ev_hash;
    uint32_t *ev_log;
    int ev_cap;
    uint64_t iters;
    int64_t step;
    double QKr[25], QKv14[100], QKv43[100];
    double Istim;
    int clamp, frozen;
    double vclamp;
    double last_sums[5];
};

static void copy_from_CRUState(cru_tmp *s, const gwo_model *m, int idx)
{
    for (int j = 0; j < 4; j++) {
        for (int k = 0; k < 6; k++) s->RyR[6 * j + k] = m->RyR[24 * idx + 6 * j + k];
        s->LCC[j] = m->LCC[4 * idx + j];
        s->LCC_inactivation[j] = m->LCCi[4 * idx + j];
        s->ClCh[j] = m->ClCh[4 * idx + j];
        s->open_RyR[j] = (double)(m->RyR[24 * idx + 6 * j + 2] + m->RyR[24 * idx + 6 * j + 3]);
        s->CaSS[j] = m->CaSS[4 * idx + j];
        s->JLCC[j] = m->JLCC[4 * idx + j];
        s->Jrel[j] = m->p.rRyR * s->open_RyR[j] * (m->CaJSR[idx] - s->CaSS[j]);
    }
    memset(s->LCC_rates, 0, sizeof(s->LCC_rates));
    memset(s->LCC_inactivation_rates, 0, sizeof(s->LCC_inactivation_rates));
    memset(s->RyR_rates, 0, sizeof(s->RyR_rates));
    memset(s->ClCh_rates, 0, sizeof(s->ClCh_rates));
    memset(s->subunit_rates, 0, sizeof(s->subunit_rates));
    memset(s->Jxfer, 0, sizeof(s->Jxfer));
    memset(s->Jiss, 0, sizeof(s->Jiss));
    s->CaJSR = m->CaJSR[idx];
    s->Jtr = 0;
}

/* update_CRUstate_from_temp (GW.tpp:47-62) */
static void update_CRUstate_from_temp(gwo_model *m, const cru_tmp *s, int idx)
{
    for (int j = 0; j < 4; j++) {
        m->LCC[4 * idx + j] = s->LCC[j];
        m->LCCi[4 * idx + j] = s->LCC_inactivation[j];
        for (int k = 0; k < 6; k++) m->RyR[24 * idx + 6 * j + k] = s->RyR[6 * j + k];
        m->ClCh[4 * idx + j] = s->ClCh[j];
        m->CaSS[4 * idx + j] = s->CaSS[j];
        m->JLCC[4 * idx + j] = s->JLCC[j];
        m->Jxfer[4 * idx + j] = s->Jxfer[j];
    }
    m->CaJSR[idx] = s->CaJSR;
    m->Jtr[idx] = s->Jtr;
}

/* GW_model ctor (GW.hpp:104-113): constants, then JLCC/Jxfer/Jtr from the
 * constructor's own default CRU state (CaSS = 1.45370e-4, CaJSR = 0.908408,
 * GW_utils.tpp:212-213) and default globals -- not from the state passed later. */
gwo_model *gwo_model_create(const gw_params_t *p, int nCRU, int prng_id, uint64_t seed)
{
    if (nCRU <= 0 || prng_id < 0 || prng_id >= GW_PRNG_COUNT) return NULL;
    gwo_model *m = (gwo_model *)calloc(1, sizeof(gwo_model));
    m->p = *p;
    m->n = nCRU;
    m->prng = prng_id;
    m->seed = seed;
    consts_init(&m->c, p, nCRU);
    gw_globals_t g0;
    memset(&g0, 0, sizeof(g0));
    {   /* GlobalState defaults (GW_utils.hpp:231-247) */
        const double Kr[5] = {0.999503, 4.13720e-4, 7.27568e-5, 8.73984e-6, 1.36159e-6};
        const double Kv43[10] = {0.953060, 0.0253906, 2.53848e-4, 1.12796e-6, 1.87950e-9,
                                 0.0151370, 0.00517622, 8.96600e-4, 8.17569e-5, 2.24032e-6};
        const double Kv14[10] = {0.722328, 0.101971, 0.00539932, 1.27081e-4, 1.82742e-6,
                                 0.152769, 0.00962328, 0.00439043, 0.00195348, 0.00143629};
        g0.V = -91.382; g0.Nai = 10.0; g0.Ki = 131.84; g0.Cai = 1.45273e-4; g0.CaNSR = 0.908882;
        g0.CaLTRPN = 8.98282e-3; g0.CaHTRPN = 0.137617; g0.m = 5.33837e-4; g0.h = 0.996345;
        g0.j = 0.997315; g0.xKs = 2.04171e-4;
        memcpy(g0.Kr, Kr, sizeof(Kr)); memcpy(g0.Kv43, Kv43, sizeof(Kv43)); memcpy(g0.Kv14, Kv14, sizeof(Kv14));
    }
    m->g = g0;
    const size_t n = (size_t)nCRU;
    m->CaSS = (double *)calloc(4 * n, sizeof(double));
    m->CaJSR = (double *)calloc(n, sizeof(double));
    m->JLCC = (double *)calloc(4 * n, sizeof(double));
    m->Jxfer = (double *)calloc(4 * n, sizeof(double));
    m->Jtr = (double *)calloc(n, sizeof(double));
    m->LCC = (int *)calloc(4 * n, sizeof(int));
    m->LCCi = (int *)calloc(4 * n, sizeof(int));
    m->RyR = (int *)calloc(24 * n, sizeof(int));
    m->ClCh = (int *)calloc(4 * n, sizeof(int));
    m->rng = (uint64_t *)calloc(4 * n, sizeof(uint64_t));
    m->ev_count = (uint64_t *)calloc(n, sizeof(uint64_t));
    m->ev_hash = (uint64_t *)calloc(n, sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) {
        gwo_rng_seed(prng_id, m->rng + 4 * i, gwo_stream_key(seed, (int64_t)i));
        m->ev_hash[i] = 0xcbf29ce484222325ULL;
        for (int j = 0; j < 4; ++j) m->CaSS[4 * i + j] = 1.45370e-4;
        m->CaJSR[i] = 0.908408;
    }
    /* initialise_JLCC (GW.tpp:12-23): the constructor's CRUs hold no open LCC, so JLCC = 0. */
    m->c.VF_RT = m->g.V * m->c.F_RT;
    m->c.JLCC_exp = gwo_exp(2 * m->c.VF_RT);
    /* initialise_Jxfer / initialise_Jtr (GW.hpp:73-74) */
    for (size_t i = 0; i < n; ++i) {
        for (int j = 0; j < 4; ++j) m->Jxfer[4 * i + j] = m->p.rxfer * (m->CaSS[4 * i + j] - m->g.Cai);
        m->Jtr[i] = m->p.rtr * (m->g.CaNSR - m->CaJSR[i]);
    }
    /* initialise_QKr (GW.hpp:75) */
    m->QKr[1 * 5 + 2] = p->Kf;
    m->QKr[2 * 5 + 1] = p->Kb;
    /* default CRU channel state until set_state: LCC closed, y = 1, RyR all in state 1 */
    for (size_t i = 0; i < n; ++i)
        for (int j = 0; j < 4; ++j) {
            m->LCC[4 * i + j] = 1;
            m->LCCi[4 * i + j] = 1;
            m->RyR[24 * i + 6 * j] = 5;
        }
    return m;
}

void gwo_model_destroy(gwo_model *m)
{
    if (!m) return;
    free(m->CaSS); free(m->CaJSR); free(m->JLCC); free(m->Jxfer); free(m->Jtr);
    free(m->LCC); free(m->LCCi); free(m->RyR); free(m->ClCh); free(m->rng);
    free(m->ev_count); free(m->ev_hash); free(m->ev_log);
    free(m);
}

/* set_initial_value (GW.tpp:5-10) */
int gwo_model_set_state(gwo_model *m, const gw_globals_t *g, const double *CaSS, const double *CaJSR,
                        const int32_t *LCC, const int32_t *LCC_i, const int32_t *RyR, const int32_t *ClCh,
                        const double *JLCC)
{
    if (g) m->g = *g;
    const size_t n = (size_t)m->n;
    if (CaSS) for (size_t i = 0; i < 4 * n; ++i) m->CaSS[i] = CaSS[i];
    if (CaJSR) for (size_t i = 0; i < n; ++i) m->CaJSR[i] = CaJSR[i];
    if (LCC) for (size_t i = 0; i < 4 * n; ++i) m->LCC[i] = LCC[i];
    if (LCC_i) for (size_t i = 0; i < 4 * n; ++i) m->LCCi[i] = LCC_i[i];
    if (RyR) for (size_t i = 0; i < 24 * n; ++i) m->RyR[i] = RyR[i];
    if (ClCh) for (size_t i = 0; i < 4 * n; ++i) m->ClCh[i] = ClCh[i];
    if (JLCC) for (size_t i = 0; i < 4 * n; ++i) m->JLCC[i] = JLCC[i];
    return 0;
}

/* Rest-state sampler: initialise_LCC / _LCC_i / _RyR / _ClCh (GW_utils.tpp:5-64)
 * and the CRUState constants (GW_utils.tpp:212-213), drawing from a per-CRU
 * stream keyed ~stream_key. Per CRU: LCC[0..3], y[0..3], RyR[0..3][0..4], ClCh[0..3]. */
void gwo_model_sample_rest_state(gwo_model *m)
{
    const double wL[3] = {0.958, 0.038, 0.004};
    const double wY[2] = {0.9425, 0.0575};
    const double wR[3] = {0.609, 0.5 * 0.391, 0.5 * 0.391};
    const double wC[2] = {0.998, 0.002};
    for (int i = 0; i < m->n; ++i) {
        uint64_t s[4];
        gwo_rng_seed(m->prng, s, ~gwo_stream_key(m->seed, i));
        for (int j = 0; j < 4; ++j) {
            m->CaSS[4 * i + j] = 1.45370e-4;
            const int idx = gwo_sample_weights(wL, 1.0, 3, gwo_u01(gwo_rng_next(m->prng, s)));
            m->LCC[4 * i + j] = idx == 0 ? 1 : (idx == 1 ? 2 : 7);
        }
        m->CaJSR[i] = 0.908408;
        for (int j = 0; j < 4; ++j) {
            const int idx = gwo_sample_weights(wY, 1.0, 2, gwo_u01(gwo_rng_next(m->prng, s)));
            m->LCCi[4 * i + j] = (idx == 0) ? 1 : 0;
        }
        for (int j = 0; j < 4; ++j) {
            int *n6 = m->RyR + 24 * i + 6 * j;
            for (int k = 0; k < 6; ++k) n6[k] = 0;
            for (int k = 0; k < 5; ++k) {
                const int idx = gwo_sample_weights(wR, 1.0, 3, gwo_u01(gwo_rng_next(m->prng, s)));
                if (idx == 0) ++n6[0];
                else if (idx == 1) ++n6[4];
                else ++n6[5];
            }
        }
        for (int j = 0; j < 4; ++j)
            m->ClCh[4 * i + j] = gwo_sample_weights(wC, 1.0, 2, gwo_u01(gwo_rng_next(m->prng, s)));
        for (int j = 0; j < 4; ++j) m->JLCC[4 * i + j] = 0.0;
    }
}

void gwo_model_set_clamp(gwo_model *m, int clamp, double vclamp) { m->clamp = clamp; m->vclamp = vclamp; }
void gwo_model_set_frozen(gwo_model *m, int frozen) { m->frozen = frozen; }
void gwo_model_set_event_log(gwo_model *m, int cap)
{
    free(m->ev_log);
    m->ev_log = NULL;
    m->ev_cap = cap > 0 ? cap : 0;
    if (m->ev_cap) m->ev_log = (uint32_t *)calloc((size_t)m->n * m->ev_cap, sizeof(uint32_t));
}

/* update_QKr (GW.tpp:25-44) and update_QKv (GW.tpp:64-123) on row-major storage. */
static void update_Q(const gw_params_t *p, double V, double *QKr, double *Q14, double *Q43)
{
#define KR(i, j) QKr[(i) * 5 + (j)]
#define A14(i, j) Q14[(i) * 10 + (j)]
#define A43(i, j) Q43[(i) * 10 + (j)]
    KR(0, 1) = 0.0069 * gwo_exp(0.0272 * V);
    KR(0, 0) = -KR(0, 1);
    KR(1, 0) = 0.0227 * gwo_exp(-0.0431 * V);
    KR(1, 1) = -(KR(1, 0) + KR(1, 2));
    KR(2, 3) = 0.0218 * gwo_exp(0.0262 * V);
    KR(2, 4) = 1.29e-5 * gwo_exp(2.71e-6 * V);
    KR(2, 2) = -(KR(2, 1) + KR(2, 3) + KR(2, 4));
    KR(3, 2) = 0.0009 * gwo_exp(-0.0269 * V);
    KR(3, 4) = 0.0622 * gwo_exp(0.0120 * V);
    KR(3, 3) = -(KR(3, 2) + KR(3, 4));
    KR(4, 3) = 0.0059 * gwo_exp(-0.0443 * V);
    KR(4, 2) = KR(3, 2) * KR(4, 3) * KR(2, 4) / (KR(2, 3) * KR(3, 4));
    KR(4, 4) = -(KR(4, 2) + KR(4, 3));

    const double alphaa14 = p->alphaa0Kv14 * gwo_exp(p->aaKv14 * V);
    const double alphaa43 = p->alphaa0Kv43 * gwo_exp(p->aaKv43 * V);
    const double alphai14 = p->alphai0Kv14 * gwo_exp(-p->aiKv14 * V);
    const double alphai43 = p->alphai0Kv43 * gwo_exp(-p->aiKv43 * V);
    const double betaa14 = p->betaa0Kv14 * gwo_exp(-p->baKv14 * V);
    const double betaa43 = p->betaa0Kv43 * gwo_exp(-p->baKv43 * V);
    const double betai14 = p->betai0Kv14 * gwo_exp(p->biKv14 * V);
    const double betai43 = p->betai0Kv43 * gwo_exp(p->biKv43 * V);
#define QKV(Q, aa, ai, ba, bi, f1, f2, f3, f4, b1, b2, b3, b4)                     \
    Q(0, 1) = 4 * aa; Q(0, 5) = bi; Q(0, 0) = -(Q(0, 1) + Q(0, 5));               \
    Q(1, 0) = ba; Q(1, 2) = 3 * aa; Q(1, 6) = f1 * bi;                           \
    Q(1, 1) = -(Q(1, 0) + Q(1, 2) + Q(1, 6));                                     \
    Q(2, 1) = 2 * ba; Q(2, 3) = 2 * aa; Q(2, 7) = f2 * bi;                       \
    Q(2, 2) = -(Q(2, 1) + Q(2, 3) + Q(2, 7));                                     \
    Q(3, 2) = 3 * ba; Q(3, 4) = aa; Q(3, 8) = f3 * bi;                           \
    Q(3, 3) = -(Q(3, 2) + Q(3, 4) + Q(3, 8));                                     \
    Q(4, 3) = 4 * ba; Q(4, 9) = f4 * bi; Q(4, 4) = -(Q(4, 3) + Q(4, 9));         \
    Q(5, 6) = 4 * aa * b1; Q(5, 0) = ai; Q(5, 5) = -(Q(5, 6) + Q(5, 0));         \
    Q(6, 5) = ba / f1; Q(6, 7) = 3 * aa * b2 / b1; Q(6, 1) = ai / b1;            \
    Q(6, 6) = -(Q(6, 5) + Q(6, 7) + Q(6, 1));                                     \
    Q(7, 6) = 2 * ba * f1 / f2; Q(7, 8) = 2 * aa * b3 / b2; Q(7, 2) = ai / b2;   \
    Q(7, 7) = -(Q(7, 6) + Q(7, 8) + Q(7, 2));                                     \
    Q(8, 7) = 3 * ba * f2 / f3; Q(8, 9) = aa * b4 / b3; Q(8, 3) = ai / b3;       \
    Q(8, 8) = -(Q(8, 7) + Q(8, 9) + Q(8, 3));                                     \
    Q(9, 8) = 4 * ba * f3 / f4; Q(9, 4) = ai / b4; Q(9, 9) = -(Q(9, 8) + Q(9, 4));
    QKV(A14, alphaa14, alphai14, betaa14, betai14, p->f1Kv14, p->f2Kv14, p->f3Kv14, p->f4Kv14,
        p->b1Kv14, p->b2Kv14, p->b3Kv14, p->b4Kv14)
    QKV(A43, alphaa43, alphai43, betaa43, betai43, p->f1Kv43, p->f2Kv43, p->f3Kv43, p->f4Kv43,
        p->b1Kv43, p->b2Kv43, p->b3Kv43, p->b4Kv43)
#undef QKV
#undef KR
#undef A14
#undef A43
}

void gwo_Q_matrices(const gw_params_t *p, double V, double QKr[25], double QKv14[100], double QKv43[100])
{
    memset(QKr, 0, 25 * sizeof(double));
    memset(QKv14, 0, 100 * sizeof(double));
    memset(QKv43, 0, 100 * sizeof(double));
    QKr[1 * 5 + 2] = p->Kf;
    QKr[2 * 5 + 1] = p->Kb;
    update_Q(p, V, QKr, QKv14, QKv43);
}

/*
 * update_V_and_concentration_derivatives, update_gate_derivatives,
 * update_Kr/Kv_derivatives (GW.tpp:126-212), given the four flux sums.
 * Fills dg (increments including the factor dt, as the reference does).
 */
static int derivatives(const gw_params_t *p, gwo_consts *c, const gw_globals_t *g, const double *QKr,
                       const double *Q14, const double *Q43, double dt, double Istim, double sJLCC,
                       double sJxfer, double sJtr, double sClCh, gw_globals_t *dg, double *cur)
{
    int ok = 1;
    const double V = g->V;
    double IKv14, IKv43;
    c->ENa = gwo_Nernst(g->Nai, p->Nao, c->RT_F, 1.0, &ok);
    c->EK = gwo_Nernst(g->Ki, p->Ko, c->RT_F, 1.0, &ok);
    c->ECa = gwo_Nernst(g->Ki, p->Ko, c->RT_F, 2.0, &ok);   /* K-based, as GW.tpp:176 */
    if (!ok) return GW_E_NONPOSITIVE_CONC;

    const double INa = gwo_INa(V, g->m, g->h, g->j, p->GNa, c->ENa);
    const double INab = gwo_Ib(V, p->GNab, c->ENa);
    const double INaCa = gwo_INaCa(c->VF_RT, c->expmVF_RT, g->Nai, g->Cai, c->Nao3, p->Cao, p->eta, p->ksat, c->INaCa_const);
    const double INaK = gwo_INaK(c->VF_RT, c->expmVF_RT, g->Nai, c->sigma, p->KmNai, c->INaK_const);
    /* IKr (GW_utils.hpp:273-279) */
    const double RKr = 1.0 / (1.0 + 1.4945 * gwo_exp(0.0446 * V));
    const double IKr = p->GKr * c->sqrtKo * RKr * g->Kr[3] * (V - c->EK) * 0.5;
    /* IKs (GW_utils.hpp:282-291) */
    const double EKs = gwo_log((p->Ko + 0.01833 * p->Nao) / (g->Ki + 0.01833 * g->Nai)) * c->RT_F;
    const double IKs = p->GKs * sq(g->xKs) * (V - EKs);
    /* IKv14 (GW_utils.hpp:299-303) with exp_term = expmVF_RT (GW.tpp:185); D4 at V == 0 */
    {
        const double e = c->expmVF_RT;
        double mm;
        if (c->VF_RT == 0.0) mm = (c->PKv14_Csc) * GWO_FARADAY * g->Kv14[4];
        else mm = (c->PKv14_Csc) * (GWO_FARADAY * c->VF_RT) * g->Kv14[4] / (1 - e);
        IKv14 = mm * ((g->Ki - p->Ko * e) + 0.02 * (g->Nai - p->Nao * e)) * 1.0e9;
    }
    IKv43 = p->GKv43 * g->Kv43[4] * (V - c->EK);   /* GW_utils.hpp:294-297 */
    const double Ito1 = IKv14 + IKv43;
    /* Ito2 (GW_utils.hpp:328-331); D4 at V == 0 */
    double Ito2;
    if (c->VF_RT == 0.0) Ito2 = sClCh * c->Ito2_const * (p->Clo - p->Clcyto);
    else Ito2 = sClCh * c->Ito2_const * c->VF_RT * (p->Clcyto * c->expmVF_RT - p->Clo) / (c->expmVF_RT - 1.0);
    /* IK1 (GW_utils.hpp:306-312) */
    const double K1inf = 1.0 / (2.0 + gwo_exp(1.5 * (V - c->EK) * c->F_RT));
    const double IK1 = p->GK1 * K1inf * c->IK1_const * (V - c->EK);
    /* IKp (GW_utils.hpp:315-319) */
    const double Kp = 1. / (1. + gwo_exp((7.488 - V) / 5.98));
    const double IKp = p->GKp * Kp * (V - c->EK);
    const double ICaL = sJLCC * c->ICaL_const;             /* GW_utils.hpp:322-325 */
    const double ICab = gwo_Ib(V, p->GCab, c->ECa);
    const double IpCa = gwo_IpCa(g->Cai, p->IpCamax, p->KmpCa);
    /* Jup (GW_utils.hpp:339-344) */
    const double fJ = gwo_pow(g->Cai / p->Kmf, p->Hf);
    const double rJ = gwo_pow(g->CaNSR / p->Kmr, p->Hr);
    const double Jup = (p->Vmaxf * fJ - p->Vmaxr * rJ) / (1.0 + fJ + rJ);
    const double Jtr_tot = sJtr * c->CRU_factor;           /* flux_average, GW_utils.hpp:354-355 */
    const double Jxfer_tot = sJxfer * c->CRU_factor;
    { const double x = p->KCMDN + g->Cai; c->beta_cyto = 1.0 / (1.0 + c->CMDN_const / sq(x)); }

    /* dTRPNCa (GW_utils.hpp:333-336) */
    dg->CaLTRPN = p->kLTRPNp * g->Cai * (p->LTRPNtot - g->CaLTRPN) - p->kLTRPNm * g->CaLTRPN;
    dg->CaHTRPN = p->kHTRPNp * g->Cai * (p->HTRPNtot - g->CaHTRPN) - p->kHTRPNm * g->CaHTRPN;
    dg->Nai = -dt * c->CSA_FVcyto * (INa + INab + 3 * INaCa + 3 * INaK);
    dg->Ki = -dt * c->CSA_FVcyto * (IKr + IKs + Ito1 + IK1 + IKp - 2 * INaK);
    dg->Cai = dt * c->beta_cyto * (-0.5 * c->CSA_FVcyto * (ICab + IpCa - 2 * INaCa) + c->VSS_Vcyto * Jxfer_tot - Jup
                                   - (dg->CaLTRPN + dg->CaHTRPN));
    dg->CaNSR = dt * (c->Vcyto_VNSR * Jup - c->VJSR_VNSR * Jtr_tot);
    dg->CaLTRPN *= dt;
    dg->CaHTRPN *= dt;
    dg->V = dt * (Istim - (INa + ICaL + IKr + IKs + Ito1 + IK1 + IKp + Ito2 + INaK + INaCa + IpCa + ICab + INab));

    /* update_gate_derivatives (GW.tpp:162-168) */
    dg->m = dt * (gwo_alpha_m(V) * (1.0 - g->m) - gwo_beta_m(V) * g->m);
    dg->h = dt * (gwo_alpha_h(V) * (1.0 - g->h) - gwo_beta_h(V) * g->h);
    dg->j = dt * (gwo_alpha_j(V) * (1.0 - g->j) - gwo_beta_j(V) * g->j);
    dg->xKs = dt * (gwo_XKsinf(V) - g->xKs) * gwo_tauXKs_inv(V);

    /* update_Kr_derivatives (GW.tpp:126-133) */
#define KR(i, j) QKr[(i) * 5 + (j)]
    const double *x = g->Kr;
    dg->Kr[0] = dt * (KR(1, 0) * x[1] + KR(0, 0) * x[0]);
    dg->Kr[1] = dt * (KR(0, 1) * x[0] + KR(2, 1) * x[2] + KR(1, 1) * x[1]);
    dg->Kr[2] = dt * (KR(1, 2) * x[1] + KR(3, 2) * x[3] + KR(4, 2) * x[4] + KR(2, 2) * x[2]);
    dg->Kr[3] = dt * (KR(2, 3) * x[2] + KR(4, 3) * x[4] + KR(3, 3) * x[3]);
    dg->Kr[4] = dt * (KR(2, 4) * x[2] + KR(3, 4) * x[3] + KR(4, 4) * x[4]);
#undef KR
    /* update_Kv_derivatives (GW.tpp:135-160) */
#define KV(Q, y, d)                                                                                  \
    d[0] = dt * (Q[1 * 10 + 0] * y[1] + Q[5 * 10 + 0] * y[5] + Q[0 * 10 + 0] * y[0]);             \
    d[1] = dt * (Q[0 * 10 + 1] * y[0] + Q[2 * 10 + 1] * y[2] + Q[6 * 10 + 1] * y[6] + Q[1 * 10 + 1] * y[1]); \
    d[2] = dt * (Q[1 * 10 + 2] * y[1] + Q[3 * 10 + 2] * y[3] + Q[7 * 10 + 2] * y[7] + Q[2 * 10 + 2] * y[2]); \
    d[3] = dt * (Q[2 * 10 + 3] * y[2] + Q[4 * 10 + 3] * y[4] + Q[8 * 10 + 3] * y[8] + Q[3 * 10 + 3] * y[3]); \
    d[4] = dt * (Q[3 * 10 + 4] * y[3] + Q[9 * 10 + 4] * y[9] + Q[4 * 10 + 4] * y[4]);             \
    d[5] = dt * (Q[6 * 10 + 5] * y[6] + Q[0 * 10 + 5] * y[0] + Q[5 * 10 + 5] * y[5]);             \
    d[6] = dt * (Q[5 * 10 + 6] * y[5] + Q[7 * 10 + 6] * y[7] + Q[1 * 10 + 6] * y[1] + Q[6 * 10 + 6] * y[6]); \
    d[7] = dt * (Q[6 * 10 + 7] * y[6] + Q[8 * 10 + 7] * y[8] + Q[2 * 10 + 7] * y[2] + Q[7 * 10 + 7] * y[7]); \
    d[8] = dt * (Q[7 * 10 + 8] * y[7] + Q[9 * 10 + 8] * y[9] + Q[3 * 10 + 8] * y[3] + Q[8 * 10 + 8] * y[8]); \
    d[9] = dt * (Q[8 * 10 + 9] * y[8] + Q[4 * 10 + 9] * y[4] + Q[9 * 10 + 9] * y[9]);
    KV(Q14, g->Kv14, dg->Kv14)
    KV(Q43, g->Kv43, dg->Kv43)
#undef KV
    if (cur) {
        cur[0] = INa; cur[1] = INab; cur[2] = INaCa; cur[3] = INaK; cur[4] = IKr; cur[5] = IKs;
        cur[6] = Ito1; cur[7] = Ito2; cur[8] = IK1; cur[9] = IKp; cur[10] = ICaL; cur[11] = ICab;
        cur[12] = IpCa; cur[13] = Jup; cur[14] = Jtr_tot; cur[15] = Jxfer_tot;
    }
    return GW_OK;
}

static void globals_to_array(const gw_globals_t *g, double out[36]) { memcpy(out, g, 36 * sizeof(double)); }

int gwo_global_rhs(const gw_params_t *p, int nCRU, const gw_globals_t *g, double Istim, double sum_JLCC,
                   double sum_Jxfer, double sum_Jtr, double sum_ClCh, double out[36], double currents[16])
{
    gwo_consts c;
    consts_init(&c, p, nCRU);
    c.VF_RT = g->V * c.F_RT;
    c.expmVF_RT = gwo_exp(-c.VF_RT);
    double QKr[25], Q14[100], Q43[100];
    gwo_Q_matrices(p, g->V, QKr, Q14, Q43);
    gw_globals_t dg;
    memset(&dg, 0, sizeof(dg));
    const int rc = derivatives(p, &c, g, QKr, Q14, Q43, 1.0, Istim, sum_JLCC, sum_Jxfer, sum_Jtr, sum_ClCh, &dg, currents);
    globals_to_array(&dg, out);
    return rc;
}

/* The four sums the derivatives read (GW.tpp:188,192,197-198), formed as D2. */
/* the exact sum of two accumulators (128-bit two's complement) */
static void fx_merge(gwo_fx *a, const gwo_fx *b)
{
    u128 acc = ((u128)a->hi << 64) | a->lo;
    acc += ((u128)b->hi << 64) | b->lo;
    a->lo = (uint64_t)acc;
    a->hi = (uint64_t)(acc >> 64);
}

/* The sums are exact, so they do not depend on the order of the CRUs: with
   OpenMP each thread sums its share of a large cell and the shares are added
   after (a small cell is summed by one thread: a parallel region per step
   would cost more than it saves). */
static int flux_sums(const gwo_model *m, double *sJLCC, double *sJxfer, double *sJtr, double *sClCh, double *sOpen)
{
    gwo_fx aL = {0, 0}, aX = {0, 0}, aT = {0, 0};
    int flag = 0;
    int64_t clch = 0, open = 0;
#pragma omp parallel if (m->n >= 4096)
    {
        gwo_fx tL = {0, 0}, tX = {0, 0}, tT = {0, 0};
        int tflag = 0;
        int64_t tclch = 0, topen = 0;
#pragma omp for schedule(static) nowait
        for (int i = 0; i < m->n; ++i) {
            const double *L = m->JLCC + 4 * i, *X = m->Jxfer + 4 * i;
            gwo_fx_add(&tL, ((L[0] + L[1]) + L[2]) + L[3], &tflag);
            gwo_fx_add(&tX, ((X[0] + X[1]) + X[2]) + X[3], &tflag);
            gwo_fx_add(&tT, m->Jtr[i], &tflag);
            for (int j = 0; j < 4; ++j) {
                tclch += m->ClCh[4 * i + j];
                topen += m->RyR[24 * i + 6 * j + 2] + m->RyR[24 * i + 6 * j + 3];
            }
        }
#pragma omp critical(gwo_flux_sums)
        {
            fx_merge(&aL, &tL);
            fx_merge(&aX, &tX);
            fx_merge(&aT, &tT);
            flag |= tflag;
            clch += tclch;
            open += topen;
        }
    }
    *sJLCC = gwo_fx_to_double(&aL);
    *sJxfer = gwo_fx_to_double(&aX);
    *sJtr = gwo_fx_to_double(&aT);
    *sClCh = (double)clch;
    if (sOpen) *sOpen = (double)open;
    return flag ? GW_E_RANGE : GW_OK;
}

/* GW_model::euler_step (GW.tpp:236-272) */
int gwo_model_euler_step(gwo_model *m, double dt, double Istim)
{
    gwo_consts *c = &m->c;
    const gw_params_t *p = &m->p;
    m->Istim = Istim;
    if (m->clamp) m->g.V = m->vclamp;
    c->VF_RT = m->g.V * c->F_RT;
    c->expmVF_RT = gwo_exp(-c->VF_RT);
    update_Q(p, m->g.V, m->QKr, m->QKv14, m->QKv43);

    double sL, sX, sT, sC, sO;
    int rc = flux_sums(m, &sL, &sX, &sT, &sC, &sO);
    m->last_sums[0] = sL; m->last_sums[1] = sX; m->last_sums[2] = sT; m->last_sums[3] = sC; m->last_sums[4] = sO;
    if (rc != GW_OK) return rc;
    gw_globals_t dg;
    memset(&dg, 0, sizeof(dg));
    rc = derivatives(p, c, &m->g, m->QKr, m->QKv14, m->QKv43, dt, Istim, sL, sX, sT, sC, &dg, NULL);
    if (rc != GW_OK) return rc;

    /* GW_model::SSA (GW.tpp:214-234) */
    ssa_consts(c, m->g.V);
    const double Cai = m->g.Cai, CaNSR = m->g.CaNSR;
    const uint32_t step_lo = (uint32_t)(m->step & 0xffff);
    uint64_t iters = 0;
#pragma omp parallel reduction(+ : iters)
    {
        cru_tmp s;
        memset(&s, 0, sizeof(s));
        s.prng = m->prng;
        s.ev_cap = m->ev_cap;
        s.step_lo = step_lo;
#pragma omp for schedule(static)
        for (int i = 0; i < m->n; i++) {
            copy_from_CRUState(&s, m, i);
            s.rng = m->rng + 4 * (size_t)i;
            s.ev_count = m->ev_count[i];
            s.ev_hash = m->ev_hash[i];
            s.ev_log = m->ev_log ? m->ev_log + (size_t)i * m->ev_cap : NULL;
            SSA_single_CRU(&s, Cai, CaNSR, dt, p, c);
            update_CRUstate_from_temp(m, &s, i);
            m->ev_count[i] = s.ev_count;
            m->ev_hash[i] = s.ev_hash;
        }
        iters += s.iters;
    }
    m->iters += iters;
    m->step++;

    if (!m->frozen) {
        gw_globals_t *g = &m->g;
        if (!m->clamp) g->V += dg.V;
        g->Nai += dg.Nai;
        g->Ki += dg.Ki;
        g->Cai += dg.Cai;
        g->CaNSR += dg.CaNSR;
        g->CaLTRPN += dg.CaLTRPN;
        g->CaHTRPN += dg.CaHTRPN;
        g->m += dg.m;
        g->h += dg.h;
        g->j += dg.j;
        g->xKs += dg.xKs;
        for (int k = 0; k < 5; ++k) g->Kr[k] += dg.Kr[k];
        for (int k = 0; k < 10; k++) {
            g->Kv14[k] += dg.Kv14[k];
            g->Kv43[k] += dg.Kv43[k];
        }
        if (!(g->V == g->V) || !(g->Cai == g->Cai)) return GW_E_NAN;
    }
    return GW_OK;
}

/* record_state (src/pyGW.hpp:56-99) */
static void record_state(const gwo_model *m, gw_record_t *out, int64_t k, double t)
{
    const gw_globals_t *g = &m->g;
    const size_t n = (size_t)m->n;
#define PUT(f, v) if (out->f) out->f[k] = (v)
    PUT(t, t); PUT(V, g->V); PUT(m, g->m); PUT(h, g->h); PUT(j, g->j); PUT(Nai, g->Nai); PUT(Ki, g->Ki);
    PUT(Cai, g->Cai); PUT(CaNSR, g->CaNSR); PUT(CaLTRPN, g->CaLTRPN); PUT(CaHTRPN, g->CaHTRPN); PUT(xKs, g->xKs);
#undef PUT
    if (out->XKr) for (int i = 0; i < 5; ++i) out->XKr[k * 5 + i] = g->Kr[i];
    if (out->XKv14) for (int i = 0; i < 10; ++i) out->XKv14[k * 10 + i] = g->Kv14[i];
    if (out->XKv43) for (int i = 0; i < 10; ++i) out->XKv43[k * 10 + i] = g->Kv43[i];
    if (out->CaJSR) memcpy(out->CaJSR + k * n, m->CaJSR, n * sizeof(double));
    if (out->CaSS) memcpy(out->CaSS + k * n * 4, m->CaSS, n * 4 * sizeof(double));
    for (size_t i = 0; i < 4 * n; ++i) {
        if (out->LCC) out->LCC[k * n * 4 + i] = m->LCC[i];
        if (out->LCC_inactivation) out->LCC_inactivation[k * n * 4 + i] = m->LCCi[i];
        if (out->ClCh) out->ClCh[k * n * 4 + i] = m->ClCh[i];
    }
    if (out->RyR) for (size_t i = 0; i < 24 * n; ++i) out->RyR[k * n * 24 + i] = m->RyR[i];
}

/* run<PRNG> (src/pyGW.hpp:102-122); the record buffer holds ceil(num_steps/record_every) rows. */
int gwo_model_run(gwo_model *m, double dt, int64_t num_steps, const double *istim, int32_t record_every,
                  gw_record_t *out)
{
    double t = 0.0;
    int64_t counter = 0;
    for (int64_t i = 0; i < num_steps; ++i) {
        const int rc = gwo_model_euler_step(m, dt, istim ? istim[i] : 0.0);
        if (rc != GW_OK) return rc;
        t += dt;
        if (out && record_every > 0 && i % record_every == 0) {
            if (counter < out->num_records) record_state(m, out, counter, t);
            ++counter;
        }
    }
    return GW_OK;
}

void gwo_model_get_globals(const gwo_model *m, gw_globals_t *g) { *g = m->g; }

void gwo_model_get_state(const gwo_model *m, double *CaSS, double *CaJSR, int32_t *LCC, int32_t *LCC_i,
                         int32_t *RyR, int32_t *ClCh, double *JLCC)
{
    const size_t n = (size_t)m->n;
    if (CaSS) memcpy(CaSS, m->CaSS, 4 * n * sizeof(double));
    if (CaJSR) memcpy(CaJSR, m->CaJSR, n * sizeof(double));
    if (JLCC) memcpy(JLCC, m->JLCC, 4 * n * sizeof(double));
    for (size_t i = 0; i < 4 * n; ++i) {
        if (LCC) LCC[i] = m->LCC[i];
        if (LCC_i) LCC_i[i] = m->LCCi[i];
        if (ClCh) ClCh[i] = m->ClCh[i];
    }
    if (RyR) for (size_t i = 0; i < 24 * n; ++i) RyR[i] = m->RyR[i];
}

void gwo_model_get_digest(const gwo_model *m, uint64_t *count, uint64_t *hash)
{
    memcpy(count, m->ev_count, (size_t)m->n * sizeof(uint64_t));
    memcpy(hash, m->ev_hash, (size_t)m->n * sizeof(uint64_t));
}

void gwo_model_get_event_log(const gwo_model *m, uint32_t *log)
{
    if (m->ev_log) memcpy(log, m->ev_log, (size_t)m->n * m->ev_cap * sizeof(uint32_t));
}

void gwo_model_get_sums(const gwo_model *m, double sums[5]) { memcpy(sums, m->last_sums, 5 * sizeof(double)); }
uint64_t gwo_model_iterations(const gwo_model *m) { return m->iters; }

int gwo_model_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Propensities of one subunit (GW_utils.tpp:118-186, 222-255; GW_utils.hpp:378-391). */
double gwo_subunit_rates(const gw_params_t *p, double V, int LCC, int y, const int RyR6[6], int ClCh,
                         double CaSS, double lcc3[3], double *inact, double ryr12[12], double *clch)
{
    gwo_consts c;
    consts_init(&c, p, 1);
    c.VF_RT = V * c.F_RT;
    c.expmVF_RT = gwo_exp(-c.VF_RT);
    ssa_consts(&c, V);
    double LCC_rates[12] = {0}, inact_rates[4] = {0}, R[48] = {0}, cl[4] = {0}, sub[4] = {0};
    int L[4] = {LCC, 1, 1, 1}, Y[4] = {y, 1, 1, 1}, C[4] = {ClCh, 0, 0, 0}, RR[24] = {0};
    double Ca[4] = {CaSS, CaSS, CaSS, CaSS};
    for (int k = 0; k < 6; ++k) RR[k] = RyR6[k];
    update_LCC_rates(LCC_rates, sub, L, Ca, 0, p, &c);
    update_LCC_inactivation_rates(inact_rates, sub, Y, c.yinfLCC, c.tauLCC, 0);
    update_RyR_rates(R, sub, RR, Ca, 0, p);
    update_ClCh_rates(cl, sub, C, Ca, p->kfClCh, p->kbClCh, 0);
    for (int k = 0; k < 3; ++k) lcc3[k] = LCC_rates[k];
    for (int k = 0; k < 12; ++k) ryr12[k] = R[k];
    *inact = inact_rates[0];
    *clch = cl[0];
    return sub[0];
}
