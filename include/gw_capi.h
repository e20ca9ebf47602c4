/*
 * gw_capi.h -- C-ABI of the MI355X Greenstein-Winslow coupled integrator.
 *
 * The reference exposes this path only through its pybind11 module
 * `GreensteinWinslow` (src/pyGW.cpp:160-344). This header is the flat,
 * exception-free boundary underneath that module: plain structs, plain
 * pointers and sizes, integer status codes. Every entry point names the
 * reference interface it replaces.
 *
 * Conventions
 *   - All calls are blocking and return GW_OK (0) or a negative status.
 *   - Host buffers are owned by the caller; device memory is owned by the
 *     gw_ctx / gw_batch objects.
 *   - One gw_ctx per host thread; calls on one ctx are not re-entrant.
 *   - `err`/`errlen` receive a NUL-terminated message on failure (may be NULL).
 */
#ifndef GW_CAPI_H
#define GW_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GW_CAPI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define GW_OK                   0
#define GW_E_INVALID_ARG       -1  /* bad argument (reference: ValueError / std::invalid_argument, src/pyGW.cpp:156) */
#define GW_E_NONPOSITIVE_CONC  -2  /* Nernst argument <= 0; the reference calls exit() (src/includes/common.hpp:72-75) */
#define GW_E_NAN               -3  /* a global state variable became NaN/Inf */
#define GW_E_HIP               -4  /* HIP runtime error */
#define GW_E_RANGE             -5  /* a per-CRU flux or channel count left the range the device encoding holds */
#define GW_E_NO_DEVICE         -6  /* no usable gfx950 device */
#define GW_E_TIMEOUT           -7  /* a bounded device-side wait expired */

/* ---- precision modes -------------------------------------------------- */
#define GW_PREC_FP64           0   /* all FP64; bit-exact with the CPU restatement */
#define GW_PREC_FP32_LOCAL     1   /* CRU-local Ca/propensities in FP32 (Ca as float-float), globals FP64 */

/* ---- PRNG families (string names of src/pyGW.cpp:139-154) -------------- */
#define GW_PRNG_XOSHIRO256SS   0   /* "xoshiro256**"   */
#define GW_PRNG_XOSHIRO256PP   1   /* "xoshiro256++"   */
#define GW_PRNG_XOSHIRO256P    2   /* "xoshiro256+"    */
#define GW_PRNG_XOROSHIRO128SS 3   /* "xoroshiro128**" */
#define GW_PRNG_XOROSHIRO128PP 4   /* "xoroshiro128++" */
#define GW_PRNG_XOROSHIRO128P  5   /* "xoroshiro128+"  */
#define GW_PRNG_COUNT          6

/*
 * Model parameters: the 115 fields of GW::Parameters in declaration order
 * (src/includes/GW_utils.hpp:40-156), with their defaults. NCaRU is the only
 * integer field; it is followed by an explicit pad so the layout is the same
 * for every compiler.  X(type, name, default)
 */
#define GW_PARAMETER_FIELDS(X)                                               \
    X(double, T, 310.0) X(double, CSA, 153.4) X(double, Vcyto, 25.84)         \
    X(double, VNSR, 1.113) X(double, VJSR, 22.26e-6) X(double, VSS, 0.2303e-6) \
    X(int32_t, NCaRU, 12500) X(int32_t, _pad_NCaRU, 0)                         \
    X(double, Ko, 4.0) X(double, Nao, 138.0) X(double, Cao, 2.0)               \
    X(double, Clo, 150.0) X(double, Clcyto, 20.0) X(double, f, 0.85)           \
    X(double, g, 2.0) X(double, f1, 0.005) X(double, g1, 7.0)                  \
    X(double, a, 2.0) X(double, b, 1.9356) X(double, gamma0, 0.44)            \
    X(double, omega, 0.02158) X(double, PCaL, 9.13e-13)                       \
    X(double, kfClCh, 13.3156) X(double, kbClCh, 2.0) X(double, Pto2, 2.65e-15)\
    X(double, k12, 877.5) X(double, k21, 250.0) X(double, k23, 2.358e8)       \
    X(double, k32, 9.6) X(double, k34, 1.415e6) X(double, k43, 13.65)         \
    X(double, k45, 0.07) X(double, k54, 93.385) X(double, k56, 1.887e7)       \
    X(double, k65, 30.0) X(double, k25, 2.358e6) X(double, k52, 0.001235)     \
    X(double, rRyR, 3.92) X(double, rxfer, 200.0) X(double, rtr, 0.333)       \
    X(double, riss, 20.0) X(double, BSRT, 0.047) X(double, KBSR, 0.00087)     \
    X(double, BSLT, 1.124) X(double, KBSL, 0.0087) X(double, CSQNT, 13.5)     \
    X(double, KCSQN, 0.63) X(double, CMDNT, 0.05) X(double, KCMDN, 0.00238)   \
    X(double, GNa, 12.8) X(double, GKr, 0.024) X(double, Kf, 0.0266)          \
    X(double, Kb, 0.1348) X(double, GKs, 0.00271) X(double, GKv43, 0.1389)    \
    X(double, alphaa0Kv43, 0.5437) X(double, aaKv43, 0.02898)                 \
    X(double, betaa0Kv43, 0.08019) X(double, baKv43, 0.04684)                 \
    X(double, alphai0Kv43, 0.04984) X(double, aiKv43, 3.37302e-4)             \
    X(double, betai0Kv43, 8.1948e-4) X(double, biKv43, 5.374e-8)              \
    X(double, f1Kv43, 1.8936) X(double, f2Kv43, 14.225)                       \
    X(double, f3Kv43, 158.574) X(double, f4Kv43, 142.937)                     \
    X(double, b1Kv43, 6.7735) X(double, b2Kv43, 15.621)                       \
    X(double, b3Kv43, 28.753) X(double, b4Kv43, 524.576)                      \
    X(double, PKv14, 1.989e-7) X(double, alphaa0Kv14, 1.8931)                 \
    X(double, aaKv14, 0.006950) X(double, betaa0Kv14, 0.01179)                \
    X(double, baKv14, 0.08527) X(double, alphai0Kv14, 0.002963)               \
    X(double, aiKv14, 0.0) X(double, betai0Kv14, 1.0571e-4)                   \
    X(double, biKv14, 0.0) X(double, f1Kv14, 0.2001) X(double, f2Kv14, 0.3203)\
    X(double, f3Kv14, 13.509) X(double, f4Kv14, 1151.765)                     \
    X(double, b1Kv14, 2.230) X(double, b2Kv14, 12.0) X(double, b3Kv14, 5.370) \
    X(double, b4Kv14, 5.240) X(double, Csc, 1.0e6) X(double, GK1, 3.0)        \
    X(double, KmK1, 13.0) X(double, GKp, 0.002659) X(double, kNaCa, 0.27)     \
    X(double, KmNa, 87.5) X(double, KmCa, 1.38) X(double, ksat, 0.2)          \
    X(double, eta, 0.35) X(double, INaKmax, 0.901) X(double, KmNai, 10.0)     \
    X(double, KmKo, 1.5) X(double, IpCamax, 0.03) X(double, KmpCa, 0.0005)    \
    X(double, GCab, 0.0002536) X(double, GNab, 0.00264)                       \
    X(double, kHTRPNp, 20.0) X(double, kHTRPNm, 6.60e-5)                      \
    X(double, kLTRPNp, 40.0) X(double, kLTRPNm, 0.04)                         \
    X(double, HTRPNtot, 0.140) X(double, LTRPNtot, 0.070)                     \
    X(double, Vmaxf, 0.0002096) X(double, Vmaxr, 0.0002096)                   \
    X(double, Kmf, 0.000260) X(double, Kmr, 1.8) X(double, Hf, 0.75)          \
    X(double, Hr, 0.75)

#define GW_DECLARE_FIELD_(type, name, dflt) type name;
/* replaces GW::Parameters<double> (src/includes/GW_utils.hpp:38-160) */
typedef struct gw_params_t {
    GW_PARAMETER_FIELDS(GW_DECLARE_FIELD_)
} gw_params_t;

/* replaces GW::GlobalState<double> (src/includes/GW_utils.hpp:229-252); same field order */
typedef struct gw_globals_t {
    double V, Nai, Ki, Cai, CaNSR, CaLTRPN, CaHTRPN;
    double m, h, j, xKs;
    double Kr[5];
    double Kv43[10];
    double Kv14[10];
} gw_globals_t;

/*
 * Caller-owned views of the per-CRU state, row-major like GWCRUState
 * (src/pyGW.cpp:78-87): CaSS[n][4], CaJSR[n], LCC/LCC_inactivation/ClCh[n][4],
 * RyR[n][4][6] (counts of the 5 RyRs of a subunit in each of 6 states).
 * JLCC[n][4] is optional (NULL = all zero, as at the start of a reference run).
 */
typedef struct gw_cru_state_t {
    double  *CaSS;
    double  *CaJSR;
    int32_t *LCC;
    int32_t *LCC_inactivation;
    int32_t *RyR;
    int32_t *ClCh;
    double  *JLCC;
} gw_cru_state_t;

/*
 * Snapshot sink, the arrays of PyGWSimulation (src/pyGW.hpp:22-53), each with
 * num_records rows. Any pointer may be NULL to skip that field.
 */
typedef struct gw_record_t {
    int64_t num_records;
    double *t, *V, *m, *h, *j, *Nai, *Ki, *Cai, *CaNSR, *CaLTRPN, *CaHTRPN, *xKs;
    double *XKr;    /* [R][5]  */
    double *XKv14;  /* [R][10] */
    double *XKv43;  /* [R][10] */
    double *CaJSR;  /* [R][n]  */
    double *CaSS;   /* [R][n][4] */
    int32_t *LCC, *LCC_inactivation, *ClCh;  /* [R][n][4] */
    int32_t *RyR;   /* [R][n][4][6] */
} gw_record_t;

/*
 * The reduced encoding of the reference's sim_dict_reduced_memory
 * (src/serialization.py:11-39), formed on the device so that only it crosses
 * PCIe (40 bytes per CRU and record against 184): FP32 concentrations
 * (round to nearest), the LCC state as int8, y and ClCh as 0/1 bytes (numpy
 * bool), and per subunit the int16 RyR code of RyR_state_to_int16
 * (src/utils.py:8-22): R1 + 6 R2 + 36 R3 + 216 R4 + 1296 R5, where a lumped
 * pair (3,4) or (5,6) with a negative member counts as its sum in the first
 * member. Any pointer may be NULL to skip that field.
 */
typedef struct gw_cru_state_reduced_t {
    float   *CaSS;              /* [n][4] */
    float   *CaJSR;             /* [n]    */
    int8_t  *LCC;               /* [n][4] */
    uint8_t *LCC_inactivation;  /* [n][4] */
    uint8_t *ClCh;              /* [n][4] */
    int16_t *RyR;               /* [n][4] */
} gw_cru_state_reduced_t;

/* Reduced snapshot sink: the globals in FP32 and, as in sim_dict_reduced_memory,
 * only the first 4 of the 5 IKr states and the first 9 of the 10 Kv states
 * (the last one of each chain is 1 minus the others). */
typedef struct gw_record_reduced_t {
    int64_t num_records;
    float *t, *V, *m, *h, *j, *Nai, *Ki, *Cai, *CaNSR, *CaLTRPN, *CaHTRPN, *xKs;
    float *XKr;     /* [R][4] */
    float *XKv14;   /* [R][9] */
    float *XKv43;   /* [R][9] */
    float *CaJSR;   /* [R][n] */
    float *CaSS;    /* [R][n][4] */
    int8_t *LCC;    /* [R][n][4] */
    uint8_t *LCC_inactivation, *ClCh;  /* [R][n][4] */
    int16_t *RyR;   /* [R][n][4] */
} gw_record_reduced_t;

/* ---- stimulus / protocol of one cell ------------------------------------ */
#define GW_PROTO_TABLE   0  /* Istim[k] read from a per-step table (reference run(), src/pyGW.hpp:114) */
#define GW_PROTO_PACE    1  /* Istim = amplitude while mod(t - start, period) < duration, else 0 */
#define GW_PROTO_CLAMP   2  /* V held at a protocol voltage; membrane ODE for V is not integrated */

typedef struct gw_protocol_t {
    int32_t kind;
    int32_t _pad;
    /* GW_PROTO_PACE */
    double period, duration, amplitude, start;
    /* GW_PROTO_CLAMP: V = hold for t < t_on or t >= t_off, V = step otherwise */
    double hold, step, t_on, t_off;
} gw_protocol_t;

/* ---- one cell of a batch ------------------------------------------------ */
typedef struct gw_cell_desc_t {
    int32_t n_cru;          /* simulated CRUs of this cell (NCaRU_sim) */
    int32_t _pad;
    uint64_t seed;          /* per-cell seed of the per-CRU xoshiro streams */
    gw_params_t params;     /* per-cell parameter set */
    gw_protocol_t protocol;
} gw_cell_desc_t;

/*
 * Per-cell observables tracked on the device every step (all FP64).
 * APD90 of a beat: from the upward crossing of -40 mV to the downward crossing
 * of V_rest + 0.1 (V_peak - V_rest), both interpolated linearly within a step.
 */
typedef struct gw_cell_obs_t {
    double peak_V, min_V, peak_Cai, peak_ICaL_neg;   /* running extrema since the last reset */
    double sum_open_RyR_frac;                        /* sum over steps of the RyR open fraction */
    double int_Jrel;                                 /* integral over time of CRU_factor * sum Jrel (mM) */
    double n_steps;                                  /* steps accumulated since the last reset */
    double v_prev;                                   /* V of the previous step */
    double apd_v_rest;                               /* minimum V since the last repolarisation */
    double apd_v_peak;                               /* peak V of the beat in progress */
    double apd_t_up;                                 /* upstroke time of the beat in progress (-1: none) */
    double apd90_last, apd90_prev;                   /* APD90 of the last two completed beats (-1: none) */
    double alt_sum, alt_count;                       /* sum and count of |APD90_k - APD90_{k-1}| */
    double beats_done;
} gw_cell_obs_t;

typedef struct gw_ctx gw_ctx;
typedef struct gw_batch gw_batch;

/* Fill a gw_params_t with the reference defaults (src/includes/GW_utils.hpp:40-156). */
void gw_params_default(gw_params_t *p);
/* Fill a gw_globals_t with the reference defaults (src/includes/GW_utils.hpp:231-247). */
void gw_globals_default(gw_globals_t *g);
/* Map a reference PRNG name (src/pyGW.cpp:139-154) to a GW_PRNG_* id; -1 if unknown. */
int gw_prng_from_name(const char *name);
/* Library version (GW_CAPI_VERSION). */
int gw_version(void);
/*
 * Seeded rest state of n CRUs on the host, identical to the state a batch cell
 * with the same seed starts from. Replaces the random CRUState constructor
 * (src/includes/GW_utils.tpp:210-218, initialise_* at :5-64), whose
 * std::random_device seed is not reproducible. No GPU is used.
 */
int gw_rest_state(int32_t n_cru, uint64_t seed, int32_t prng_id, gw_cru_state_t *out);

/* Create / destroy a device context (one HIP stream + scratch) on `device`. */
int gw_ctx_create(gw_ctx **out, int device, char *err, size_t errlen);
int gw_ctx_destroy(gw_ctx *ctx);

/*
 * One simulation with the semantics of run<PRNG> (src/pyGW.hpp:102-122) as
 * dispatched by run_PRNG_arg (src/pyGW.cpp:134-157):
 *   for i in 0..num_steps-1: Istim = istim[i]; euler_step(dt); t += dt;
 *                            if i % record_every == 0: record.
 * The record sink must hold ceil(num_steps / record_every) rows.
 * seed: per-CRU stream seed (the reference seeds from std::random_device;
 * callers wanting that pass a random seed).
 */
int gw_run(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt,
           int64_t num_steps, const double *istim, int32_t record_every,
           const gw_globals_t *g0, const gw_cru_state_t *c0, uint64_t seed,
           int32_t prng_id, int32_t precision, gw_record_t *out,
           char *err, size_t errlen);
/*
 * gw_run for callers that cannot allocate device memory (code that must not
 * call hipMalloc, such as tight serving loops or harnesses that forbid it).
 * No counterpart in the reference. Same arguments, semantics and results as
 * gw_run, bit for bit, but the call allocates nothing on the device: the
 * cell's state, block partials, stimulus table and staging are carved from a
 * static 64 MiB array that each GPU the library runs on holds until the
 * process ends (a 12,500-CRU cell fits with a stimulus table of up to about
 * 7.5 million steps; a larger run is refused with GW_E_INVALID_ARG). Every
 * call on one GPU uses the same array and leaves its contents behind; each
 * call initialises what it reads. Calls that overlap on one GPU (from two
 * threads or two contexts) race on the array and must not be made.
 */
int gw_run_noalloc(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt,
                   int64_t num_steps, const double *istim, int32_t record_every,
                   const gw_globals_t *g0, const gw_cru_state_t *c0, uint64_t seed,
                   int32_t prng_id, int32_t precision, gw_record_t *out,
                   char *err, size_t errlen);
/*
 * gw_run with its records in the reduced encoding (gw_record_reduced_t),
 * replacing run() followed by sim_dict_reduced_memory (src/serialization.py:11-39)
 * without the full snapshots ever leaving the device.
 */
int gw_run_reduced(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt,
                   int64_t num_steps, const double *istim, int32_t record_every,
                   const gw_globals_t *g0, const gw_cru_state_t *c0, uint64_t seed,
                   int32_t prng_id, int32_t precision, gw_record_reduced_t *out,
                   char *err, size_t errlen);

/*
 * Batched session API: many independent cells (uneven CRU counts, per-cell
 * parameters, seeds and protocols) advanced together on one GPU.
 * Replaces repeated GWModel.simulate calls (src/model.py:940-982).
 * If g0 is NULL the reference default globals are used for every cell.
 * CRU state starts at the seeded rest-state sampler (src/includes/GW_utils.tpp:5-64);
 * gw_batch_set_cru_state overrides it per cell.
 *
 * Per-cell overrides (gw_batch_set_cru_state, gw_batch_set_globals,
 * gw_batch_set_stimulus_table) touch only the named cell. Before the first
 * gw_batch_step they set that cell's starting point (step 0, as the
 * reference's constructor and set_initial_value, GW.hpp:93-113, GW.tpp:5-10).
 * After it the cell keeps its clock, step index and observables, and its
 * current step's derivatives and SSA constants are formed again from the new
 * state; a new CRU state also re-forms the JLCC and ClCh sums, while Jxfer and
 * Jtr keep the values of the cell's last SSA (in the reference they are
 * per-CRU arrays written only by the SSA, GW.tpp:47-62). No other cell changes.
 */
int gw_batch_create(gw_ctx *ctx, int32_t ncells, const gw_cell_desc_t *cells,
                    int32_t precision, int32_t prng_id, double dt,
                    const gw_globals_t *g0, gw_batch **out, char *err, size_t errlen);
int gw_batch_destroy(gw_batch *b);
int gw_batch_set_cru_state(gw_batch *b, int32_t cell, const gw_cru_state_t *c,
                           char *err, size_t errlen);
int gw_batch_get_cru_state(gw_batch *b, int32_t cell, gw_cru_state_t *c,
                           char *err, size_t errlen);
/*
 * The post-hoc currents of GWSolution.ICaL / .Ito2 (src/model.py:414-481) at
 * each cell's current state, formed on the device: ICaL from the open LCCs
 * at the instantaneous V and CaSS, Ito2 from the open ClCh, in pA/pF. (The
 * integrator's own I_CaL, used for dV/dt and tracked in gw_cell_obs_t and the
 * trace, freezes each channel's flux at its opening, src/includes/SSA.tpp:141;
 * the two agree while CaSS and V stay at their values of the opening.)
 * ICaL or Ito2 may be NULL; each has ncells entries.
 */
int gw_batch_get_currents(gw_batch *b, double *ICaL, double *Ito2, char *err, size_t errlen);
/* One cell's CRU state in the reduced encoding of gw_cru_state_reduced_t. */
int gw_batch_get_cru_state_reduced(gw_batch *b, int32_t cell, gw_cru_state_reduced_t *c,
                                   char *err, size_t errlen);
int gw_batch_set_globals(gw_batch *b, int32_t cell, const gw_globals_t *g,
                         char *err, size_t errlen);
int gw_batch_get_globals(gw_batch *b, gw_globals_t *g_out /* [ncells] */,
                         char *err, size_t errlen);
/* Per-step table stimulus for one cell (GW_PROTO_TABLE); `istim` has n entries, copied to the device.
 * Entry k is the stimulus of step k; steps past the table get 0. A voltage-clamped cell
 * (GW_PROTO_CLAMP) is refused with GW_E_INVALID_ARG: a stimulus has no effect under a clamp. */
int gw_batch_set_stimulus_table(gw_batch *b, int32_t cell, const double *istim, int64_t n,
                                char *err, size_t errlen);
/*
 * Replay steps as captured HIP graphs of steps_per_graph launches (0 = plain
 * launches, the default). The graph is captured on first use and again after
 * any change of launch shape or mode; results are identical either way.
 * No reference counterpart.
 */
int gw_batch_set_graph(gw_batch *b, int32_t steps_per_graph, char *err, size_t errlen);
/*
 * Persistent stepping (no reference counterpart; the reference forks and joins
 * its OpenMP loop every step, src/includes/GW.tpp:223-233). When every SSA
 * block of the default launch shape (4 x 256) and one close block per cell can
 * be resident on the GPU at once -- a few cells of up to tens of thousands of
 * CRUs -- one cooperative launch runs all the steps of a gw_batch_step call
 * with the CRUs held in registers, and the cells' SSA blocks and close blocks
 * exchange each step's flux sums and constants through per-cell counters
 * instead of kernel boundaries. Results are bit-identical to one launch per
 * step. Every wait in the kernel is bounded; an expired wait makes the step
 * call return GW_E_TIMEOUT. mode: 0 = one launch per step (the default: it is
 * the faster form on MI355X, because a step of a small batch is bound by the
 * close's serial arithmetic, which the launch-per-step form overlaps with the
 * SSA and the launch queue, and the persistent form adds the exchange of the
 * step's sums and constants through memory), 1 = persistent (GW_E_INVALID_ARG
 * if the blocks cannot all be resident), 2 = persistent when they can and the
 * launch shape is the default. Frozen replay always launches per chunk.
 * gw_batch_get_persistent reports whether the next steps will run persistently.
 */
int gw_batch_set_persistent(gw_batch *b, int32_t mode, char *err, size_t errlen);
int gw_batch_get_persistent(gw_batch *b, int32_t *active);
/*
 * SSA steps per plain launch: 0 (automatic, the default), 1 or 2. No counterpart in the reference (its OpenMP loop,
 * src/includes/GW.tpp:223-233, runs one step at a time). With 2, each launch of the
 * FP32-local one-lane shapes of 64 threads (lanes_per_cru 1, block 64) advances every
 * CRU two steps with its state held in registers, so the CRU state crosses HBM once per
 * two steps; the two closes the launch's SSA steps need run just before it, in a kernel
 * of their own. Other shapes, FP64, frozen globals, replayed graphs and the persistent
 * kernel keep one step per launch. With one CRU per lane the results are bit-identical
 * to one step per launch; with refilled lanes they are deterministic, and the rounding
 * of the FP64 flux sums can differ (their order follows which CRUs each lane ran).
 * Automatic: two steps per launch when a launch's SSA blocks are more than the GPU
 * holds at once (the state's HBM traffic binds), one otherwise (latency binds, and the
 * separate close kernel would add to it). gw_batch_get_steps_per_launch reports what
 * the next plain steps use.
 */
int gw_batch_set_steps_per_launch(gw_batch *b, int32_t steps, char *err, size_t errlen);
int gw_batch_get_steps_per_launch(gw_batch *b, int32_t *steps);
/* Polls before a wait of the persistent kernel gives up (0 = the default, about 2
 * million, seconds of waiting). A diagnostic and test hook: a tiny limit makes a step
 * call end with GW_E_TIMEOUT, and every block of the kernel still exits normally. */
int gw_batch_set_spin_limit(gw_batch *b, int64_t polls, char *err, size_t errlen);
/* Advance every cell by nsteps Euler steps. */
int gw_batch_step(gw_batch *b, int64_t nsteps, char *err, size_t errlen);
/* Simulated time of the batch (accumulated t += dt, as src/pyGW.hpp:116): the clock of the
 * running cells, which they share; if every cell has stopped on an error, the latest clock. */
int gw_batch_get_time(gw_batch *b, double *t, int64_t *steps_done);
int gw_batch_get_obs(gw_batch *b, gw_cell_obs_t *obs /* [ncells] */, char *err, size_t errlen);
int gw_batch_reset_obs(gw_batch *b, char *err, size_t errlen);
/*
 * Reduced trace: every `every` steps, per cell, (V, Cai, ICaL, open RyR fraction)
 * is appended to a device ring of `capacity` rows; gw_batch_get_trace copies out
 * rows [0, *rows) as [rows][ncells][4] doubles and clears the ring. *rows is the
 * largest row count over the cells: a cell that stopped on an error wrote fewer
 * rows, and its rows after the stop read NaN.
 */
int gw_batch_enable_trace(gw_batch *b, int32_t every, int64_t capacity, char *err, size_t errlen);
int gw_batch_get_trace(gw_batch *b, double *out, int64_t max_rows, int64_t *rows,
                       char *err, size_t errlen);
/* Per-CRU replay digest: number of SSA events and a 64-bit hash of the event sequence. */
int gw_batch_get_event_digest(gw_batch *b, int32_t cell, uint64_t *count, uint64_t *hash,
                              char *err, size_t errlen);
/*
 * Replay controls (no reference counterpart; used for bit-exact replay against
 * a CPU restatement): frozen != 0 keeps every cell's globals fixed so only the
 * SSA advances; digest != 0 keeps the per-CRU event digests.
 */
int gw_batch_set_mode(gw_batch *b, int32_t frozen, int32_t digest);
/* Total SSA loop iterations executed (diagnostic for the flop count per CRU-step). */
int gw_batch_get_iterations(gw_batch *b, uint64_t *iters);
/*
 * Launch shape of the SSA kernel. lanes_per_cru: 4 (a quad of lanes per CRU,
 * one subunit per lane) or 1 (one lane per CRU, FP32-local batches only);
 * 0 = 4. block: threads per block, 64, 128, 256 or 512 with 4 lanes, 64, 128
 * or 256 with 1 lane; 0 = 256 or 64. crus_per_lane: 1, 2, 4 or 8 with 1 lane
 * (64, 128 or 256 threads); 1 with 4 lanes, or 2 (64 threads) and 4 (64, 128
 * or 256 threads); each wave then owns 64 x crus_per_lane consecutive CRUs (16 x
 * with quads) and a lane or quad whose CRU has finished its step takes the next
 * one, so waves wait less on their slowest CRU; 0 = 1. FP64 results do not depend on the shape (exact flux
 * sums). Shapes other than 4 x 256 exist for xoshiro256** batches without
 * replay digests or frozen globals. No reference counterpart (the reference
 * has one OpenMP loop over CRUs, src/includes/GW.tpp:223-233).
 */
#define GW_LAUNCH_CONFIGS 20
/* The shape of launch configuration `index` (0 .. GW_LAUNCH_CONFIGS-1), the order of autotune's timings. */
int gw_launch_config(int32_t index, int32_t *lanes_per_cru, int32_t *block, int32_t *crus_per_lane);
int gw_batch_set_launch(gw_batch *b, int32_t lanes_per_cru, int32_t block, int32_t crus_per_lane,
                        char *err, size_t errlen);
int gw_batch_get_launch(gw_batch *b, int32_t *lanes_per_cru, int32_t *block, int32_t *crus_per_lane);
/*
 * Autotuner: times trial_steps Euler steps of the batch in every launch
 * configuration available to it (the FP32-local one-lane shapes of 64 threads
 * with the batch's steps per launch), each trial starting from the same state
 * (the batch state is saved before and restored after), keeps the fastest;
 * where that shape can run one or two SSA steps per launch, times it with the
 * other choice too and keeps the faster (gw_batch_get_steps_per_launch); then
 * times the shape with its steps replayed as captured graphs of 64 steps (one
 * step per launch) and keeps graphs if they are more than 2% faster
 * (*graph_out: 64 or 0). The trials run from
 * the batch's current state: tune on a state like the one the run will spend
 * its time in. ms_per_step[GW_LAUNCH_CONFIGS + 1] (may be NULL) receives the
 * time per step of each configuration (gw_launch_config order; -1 where not
 * available) and of the graph trial.
 */
int gw_batch_autotune(gw_batch *b, int64_t trial_steps, int32_t *lanes_out, int32_t *block_out,
                      int32_t *crus_per_lane_out, int32_t *graph_out, double *ms_per_step,
                      char *err, size_t errlen);
/*
 * Time nsteps Euler steps with HIP events on the batch's stream: *ms is the
 * wall time of the whole sequence, *launches the number of SSA dispatches (one
 * per step, one per two steps with gw_batch_set_steps_per_launch(b, 2), one for
 * the persistent kernel; close-only launches are not counted). When ssa_ms is
 * not null, every SSA dispatch also records its own start/end events and *ssa_ms
 * is the sum of their durations (with two steps per launch, the SSA kernel
 * without the close kernels between its launches).
 */
int gw_batch_time_steps(gw_batch *b, int64_t nsteps, double *ms, double *ssa_ms, int64_t *launches,
                        char *err, size_t errlen);

/* One-shot convenience: create, run nsteps, copy out globals and observables, destroy. */
int gw_run_batch(gw_ctx *ctx, int32_t ncells, const gw_cell_desc_t *cells,
                 int32_t precision, double dt, int64_t nsteps,
                 gw_globals_t *g_out, gw_cell_obs_t *obs_out,
                 char *err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* GW_CAPI_H */
