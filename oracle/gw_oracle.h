/*
 * gw_oracle.h -- CPU restatement of the reference Greenstein-Winslow hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * (gw-model_amd/). Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it.
 *
 * Every function cites the reference file:line it restates. Pinning of each
 * function (what checks it beyond this restatement) is listed at the top of
 * gw_oracle.c.
 */
#ifndef GW_ORACLE_H
#define GW_ORACLE_H

#include <stdint.h>
#include "../include/gw_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- PRNG (src/includes/xoshiro.hpp) ----------------------- */
uint64_t gwo_splitmix64_next(uint64_t *state);
void     gwo_rng_seed(int prng_id, uint64_t s[4], uint64_t seed);
uint64_t gwo_rng_next(int prng_id, uint64_t s[4]);
double   gwo_u01(uint64_t x);
/* Per-CRU stream key of CRU i of a cell with seed `seed`. */
uint64_t gwo_stream_key(uint64_t seed, int64_t i);

/* ---------------- libm shared with the device --------------------------- */
double gwo_exp(double x);
double gwo_log(double x);
double gwo_pow(double x, double y);

/* ---------------- exact flux accumulator -------------------------------- */
typedef struct gwo_fx { uint64_t lo, hi; } gwo_fx;
void   gwo_fx_add(gwo_fx *a, double x, int *flag);
double gwo_fx_to_double(const gwo_fx *a);

/* ---------------- point functions exported for KAT tests ---------------- */
double gwo_Nernst(double Xi, double Xo, double RT_F, double z, int *ok);
double gwo_INa(double V, double m, double h, double j, double GNa, double ENa);
double gwo_Ib(double V, double Gx, double Ex);
double gwo_IpCa(double Cai, double IpCamax, double KmpCa);
double gwo_fNaK(double VF_RT, double expmVF_RT, double sigma);
double gwo_INaK(double VF_RT, double expmVF_RT, double Nai, double sigma, double KmNai, double INaK_mult);
double gwo_INaCa(double VF_RT, double expmVF_RT, double Nai, double Cai, double Nao3, double Cao,
                 double eta, double ksat, double INaCa_mult);
double gwo_alpha_m(double V); double gwo_beta_m(double V);
double gwo_alpha_h(double V); double gwo_beta_h(double V);
double gwo_alpha_j(double V); double gwo_beta_j(double V);
double gwo_XKsinf(double V);  double gwo_tauXKs_inv(double V);
double gwo_alphaLCC(double V); double gwo_betaLCC(double V);
double gwo_yinfLCC(double V);  double gwo_tauLCC(double V);
int gwo_sample_weights(const double *w, double total, int n, double u01);
int gwo_sample_binomial_replay(double p, int N, const uint64_t *raw, int *X);

/* LCC (12), inactivation, RyR (12) and ClCh propensities of one subunit at
 * (V-constants, CaSS), in the reference's storage order. Returns the subunit
 * total accumulated in the reference order. */
double gwo_subunit_rates(const gw_params_t *p, double V, int LCC, int y, const int RyR6[6],
                         int ClCh, double CaSS, double lcc3[3], double *inact,
                         double ryr12[12], double *clch);

/* Q matrices of the IKr (5x5) and Kv1.4/Kv4.3 (10x10) chains, row-major Q[i*n+j]. */
void gwo_Q_matrices(const gw_params_t *p, double V, double QKr[25], double QKv14[100], double QKv43[100]);

/* d(state)/dt of the global ODE at a state, given the four flux sums. Output
 * order is that of gw_globals_t (36 doubles), before multiplication by dt. */
int gwo_global_rhs(const gw_params_t *p, int nCRU, const gw_globals_t *g, double Istim,
                   double sum_JLCC, double sum_Jxfer, double sum_Jtr, double sum_ClCh,
                   double out[36], double currents[16]);

/* ---------------- the coupled model ------------------------------------- */
typedef struct gwo_model gwo_model;

gwo_model *gwo_model_create(const gw_params_t *p, int nCRU, int prng_id, uint64_t seed);
void gwo_model_destroy(gwo_model *m);
/* set_initial_value (src/includes/GW.tpp:5-10); RyR is [n][4][6]. JLCC may be NULL. */
int  gwo_model_set_state(gwo_model *m, const gw_globals_t *g, const double *CaSS,
                         const double *CaJSR, const int32_t *LCC, const int32_t *LCC_i,
                         const int32_t *RyR, const int32_t *ClCh, const double *JLCC);
/* Seeded rest-state sampler (src/includes/GW_utils.tpp:5-64, per-CRU streams). */
void gwo_model_sample_rest_state(gwo_model *m);
/* Voltage clamp (additive): V held at vclamp when clamp != 0. */
void gwo_model_set_clamp(gwo_model *m, int clamp, double vclamp);
/* Freeze the globals: only the SSA advances (replay mode). */
void gwo_model_set_frozen(gwo_model *m, int frozen);
/* Keep per-CRU event logs of the first `cap` events (0 disables). */
void gwo_model_set_event_log(gwo_model *m, int cap);
int  gwo_model_euler_step(gwo_model *m, double dt, double Istim);
int  gwo_model_run(gwo_model *m, double dt, int64_t num_steps, const double *istim,
                   int32_t record_every, gw_record_t *out);
void gwo_model_get_globals(const gwo_model *m, gw_globals_t *g);
void gwo_model_get_state(const gwo_model *m, double *CaSS, double *CaJSR, int32_t *LCC,
                         int32_t *LCC_i, int32_t *RyR, int32_t *ClCh, double *JLCC);
void gwo_model_get_digest(const gwo_model *m, uint64_t *count, uint64_t *hash);
void gwo_model_get_event_log(const gwo_model *m, uint32_t *log /* [n][cap] */);
void gwo_model_get_sums(const gwo_model *m, double sums[5]); /* JLCC, Jxfer, Jtr, ClCh, open RyR */
uint64_t gwo_model_iterations(const gwo_model *m);
int  gwo_model_threads(void);

#ifdef __cplusplus
}
#endif
#endif
