// gw_ssa.h -- per-CRU Gillespie SSA over one global step on a quad of lanes
// (one lane per subunit), in FP64 or in the FP32-local variant.
//
// Restates SSA_single_CRU and its helpers (src/includes/SSA.tpp:10-310,
// SSA.hpp:67-71, GW_utils.tpp:118-255, GW_utils.hpp:378-391) for gfx950:
//   - lane j of a quad owns subunit j of one CRU: its CaSS, JLCC, Jrel and one
//     packed 64-bit word (six signed RyR counts, the LCC state, the y gate and
//     the ClCh state). CRU-level state (CaJSR, the xoshiro stream, the step's
//     time, the waiting time) is replicated in the four lanes and evolves
//     identically in each, so every lane draws the same uniforms;
//   - the 4-ring coupling Jiss (SSA.tpp:39-42) and the ordered sums over
//     subunits are DPP quad_perm moves; sums keep the reference order
//     ((x0 + x1) + x2) + x3 in every lane;
//   - the propensities of the chosen subunit are recomputed when an event is
//     sampled, from the inputs the reference's stored rate arrays were formed
//     from (CaSS before this iteration's update, RyR counts before any
//     threshold re-split), so every comparison sees identical values;
//   - all arithmetic keeps the reference's operation order (in FP64,
//     bit-exact with the CPU oracle when compiled with -ffp-contract=off);
//   - the FP32-local variant runs the same code with R = float: propensities,
//     fluxes, buffering factors and waiting times in FP32 with the hardware
//     reciprocal and log2; the integer channel states and the random streams
//     are unchanged. CaSS and CaJSR keep FP64 accumulators (each FP32
//     increment is added to an FP64 running value) because their per-step
//     increments near rest are below one FP32 ulp (CaJSR: about 0.6 ulp), and
//     the concentration differences that drive fluxes are formed in FP64.
#pragma once

#include "gw_model.h"

namespace gw {

// ---- packed subunit word -------------------------------------------------
// bytes 0..5: signed count of the subunit's 5 RyRs in states 1..6;
// bits 48..51: LCC state (1..12); bit 52: y (1 = available); bit 53: ClCh open.
GW_HD int ryr_get(uint64_t w, int k) { return static_cast<int>(static_cast<int8_t>((w >> (8 * k)) & 0xff)); }
GW_HD uint64_t ryr_set(uint64_t w, int k, int v)
{
    return (w & ~(0xffULL << (8 * k))) | (static_cast<uint64_t>(static_cast<uint8_t>(static_cast<int8_t>(v))) << (8 * k));
}
GW_HD uint64_t ryr_add(uint64_t w, int k, int d) { return ryr_set(w, k, ryr_get(w, k) + d); }
GW_HD int sub_lcc(uint64_t w) { return static_cast<int>((w >> 48) & 15u); }
GW_HD int sub_y(uint64_t w) { return static_cast<int>((w >> 52) & 1u); }
GW_HD int sub_cl(uint64_t w) { return static_cast<int>((w >> 53) & 1u); }
GW_HD uint64_t sub_set_lcc(uint64_t w, int s) { return (w & ~(15ULL << 48)) | (static_cast<uint64_t>(s) << 48); }
GW_HD uint64_t sub_flip_y(uint64_t w) { return w ^ (1ULL << 52); }
GW_HD uint64_t sub_flip_cl(uint64_t w) { return w ^ (1ULL << 53); }
GW_HD uint64_t sub_pack(int lcc, int y, int cl, const int *n6)
{
    uint64_t w = 0;
    for (int k = 0; k < 6; ++k) w = ryr_set(w, k, n6[k]);
    w = sub_set_lcc(w, lcc);
    if (y) w |= 1ULL << 52;
    if (cl) w |= 1ULL << 53;
    return w;
}

// New LCC state after transition `tr` from state s (the map of sample_LCC,
// src/includes/SSA.tpp:103-197): mode-normal states 1..6, mode-Ca 7..12.
GW_HD int lcc_next(int s, int tr)
{
    if (s == 6) return 5;
    if (s == 12) return 11;
    if (s == 1) return tr == 0 ? 2 : 7;
    if (s == 7) return tr == 0 ? 1 : 8;
    if (s <= 5) return tr == 0 ? s - 1 : (tr == 1 ? s + 1 : s + 6);
    return tr == 0 ? s - 6 : (tr == 1 ? s - 1 : s + 1);
}

// Per-cell parameters the SSA reads (a subset of gw_params_t and CellConsts),
// in the SSA's local arithmetic type.
template <class R>
struct SsaParamsT {
    R rRyR, rxfer, rtr, riss;
    R KBSR, KBSL, KCSQN, BSR_const, BSL_const, CSQN_const, VSS_VJSR;
    R k12, k21, k23, k32, k34, k43, k45, k54, k56, k65, k25, k52;
    R kfClCh, kbClCh, Cao_scaled;
    // rate sums and products of the FP32-local subunit total (formed in FP64)
    R k23_25, k45_43, k56_54, k45k34, k32k43, k65k52, k65k54;
};
using SsaParams = SsaParamsT<double>;
using SsaParamsF = SsaParamsT<float>;

GW_HD void ssa_params_init(SsaParams &s, const gw_params_t &p, const CellConsts &c)
{
    s.rRyR = p.rRyR; s.rxfer = p.rxfer; s.rtr = p.rtr; s.riss = p.riss;
    s.KBSR = p.KBSR; s.KBSL = p.KBSL; s.KCSQN = p.KCSQN;
    s.BSR_const = c.BSR_const; s.BSL_const = c.BSL_const; s.CSQN_const = c.CSQN_const; s.VSS_VJSR = c.VSS_VJSR;
    s.k12 = p.k12; s.k21 = p.k21; s.k23 = p.k23; s.k32 = p.k32; s.k34 = p.k34; s.k43 = p.k43;
    s.k45 = p.k45; s.k54 = p.k54; s.k56 = p.k56; s.k65 = p.k65; s.k25 = p.k25; s.k52 = p.k52;
    s.kfClCh = p.kfClCh; s.kbClCh = p.kbClCh; s.Cao_scaled = c.Cao_scaled;
    s.k23_25 = p.k23 + p.k25; s.k45_43 = p.k45 + p.k43; s.k56_54 = p.k56 + p.k54;
    s.k45k34 = p.k45 * p.k34; s.k32k43 = p.k32 * p.k43; s.k65k52 = p.k65 * p.k52; s.k65k54 = p.k65 * p.k54;
}

GW_HD void ssa_params_f32(SsaParamsF &f, const SsaParams &s)
{
    const double *src = &s.rRyR;
    float *dst = &f.rRyR;
    for (int i = 0; i < static_cast<int>(sizeof(SsaParams) / sizeof(double)); ++i) dst[i] = static_cast<float>(src[i]);
}

// Local arithmetic of the SSA. FP64: IEEE division and the shared log, bit-exact
// with the oracle. FP32 (the FP32-local variant): one hardware reciprocal
// (v_rcp_f32, 1 ulp) times the numerator, and the hardware log2 (v_log_f32).
__device__ __forceinline__ double ldiv(double a, double b) { return a / b; }
__device__ __forceinline__ float ldiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double lneglog(double u) { return -gw::log(u); }
__device__ __forceinline__ float lneglog(float u) { return -(__builtin_amdgcn_logf(u) * 0.693147182f); }
__device__ __forceinline__ float sq(float x) { return x * x; }

// Rapid-buffering factors of update_CaSS (SSA.tpp:90) and update_CaJSR
// (SSA.hpp:69). FP64: the reference's expressions. FP32-local: the same
// fractions over one denominator, one reciprocal each instead of three and two.
template <class R>
__device__ __forceinline__ R beta_ss(const SsaParamsT<R> &P, R ca)
{
    if constexpr (sizeof(R) == 8) {
        return ldiv(R(1), 1 + ldiv(P.BSR_const, sq(P.KBSR + ca)) + ldiv(P.BSL_const, sq(P.KBSL + ca)));
    } else {
#pragma clang fp contract(fast)
        const R x2 = sq(P.KBSR + ca), y2 = sq(P.KBSL + ca), p = x2 * y2;
        return ldiv(p, p + P.BSR_const * y2 + P.BSL_const * x2);
    }
}
template <class R>
__device__ __forceinline__ R beta_jsr(const SsaParamsT<R> &P, R jsr)
{
    if constexpr (sizeof(R) == 8) {
        return ldiv(R(1), 1 + ldiv(P.CSQN_const, sq(P.KCSQN + jsr)));
    } else {
#pragma clang fp contract(fast)
        const R x2 = sq(P.KCSQN + jsr);
        return ldiv(x2, x2 + P.CSQN_const);
    }
}

// The 12 RyR propensities of one subunit (update_RyR_rates, GW_utils.tpp:222-255),
// visited in storage order; `f(k, r)` receives each rate.
template <class R, class F>
__device__ __forceinline__ void ryr_rates(const SsaParamsT<R> &P, uint64_t w, R ca, F &&f)
{
    const R C2 = ca * ca;
    const int n1 = ryr_get(w, 0), n2 = ryr_get(w, 1), n3 = ryr_get(w, 2);
    const int n4 = ryr_get(w, 3), n5 = ryr_get(w, 4), n6 = ryr_get(w, 5);
    const bool hi34 = ca > R(3.685e-2);
    const bool hi56 = ca > R(1.15e-4);
    R tau34 = R(0), eq56 = R(0);
    if (hi34) tau34 = ldiv(R(1), P.k34 * C2 + P.k43);
    if (hi56) eq56 = ldiv(P.k65, P.k56 * C2 + P.k65);
    f(0, static_cast<R>(n1) * P.k12 * C2);
    f(1, static_cast<R>(n2) * P.k23 * C2);
    f(2, static_cast<R>(n2) * P.k25 * C2);
    f(3, hi34 ? R(0) : static_cast<R>(n3) * P.k34 * C2);
    f(4, hi34 ? static_cast<R>(n3 + n4) * P.k45 * P.k34 * C2 * tau34 : static_cast<R>(n4) * P.k45);
    f(5, hi56 ? R(0) : static_cast<R>(n5) * P.k56 * C2);
    f(6, static_cast<R>(n2) * P.k21);
    f(7, hi34 ? static_cast<R>(n3 + n4) * P.k32 * P.k43 * tau34 : static_cast<R>(n3) * P.k32);
    f(8, hi34 ? R(0) : static_cast<R>(n4) * P.k43);
    f(9, hi56 ? static_cast<R>(n5 + n6) * P.k52 * eq56 : static_cast<R>(n5) * P.k52);
    f(10, hi56 ? static_cast<R>(n5 + n6) * P.k54 * C2 * eq56 : static_cast<R>(n5) * P.k54 * C2);
    f(11, hi56 ? R(0) : static_cast<R>(n6) * P.k65);
}

// Subunit total. FP64: in the reference accumulation order (SSA.tpp:48-56),
// 0 + LCC sum, + inactivation, + RyR[0..11], + ClCh, which bit-exact replay
// needs. FP32-local: the same sum with the 12 RyR rates of ryr_rates gathered
// by count (each lumped pair's two rates in one closed form), about half the
// instructions; the event choice still walks the individual rates (ryr_pick).
template <class R, class KT>
__device__ __forceinline__ R subunit_total(const SsaParamsT<R> &P, const KT &K, const LccTableT<R> &T, uint64_t w, R ca)
{
    const int L = sub_lcc(w);
    R s = T.P[L] + T.Q[L] * ca;
    s += sub_y(w) ? K.inact1 : K.inact0;
    if constexpr (sizeof(R) == 8) {
        ryr_rates(P, w, ca, [&](int, R r) { s += r; });
        s += sub_cl(w) ? P.kbClCh : P.kfClCh * ca;
        return s;
    } else {
#pragma clang fp contract(fast)
        s += sub_cl(w) ? P.kbClCh : P.kfClCh * ca;
        const R C2 = ca * ca;
        const R n1 = static_cast<R>(ryr_get(w, 0)), n2 = static_cast<R>(ryr_get(w, 1));
        const R n3 = static_cast<R>(ryr_get(w, 2)), n4 = static_cast<R>(ryr_get(w, 3));
        const R n5 = static_cast<R>(ryr_get(w, 4)), n6 = static_cast<R>(ryr_get(w, 5));
        // 1->2, 2->3, 2->5, 2->1
        R r = C2 * (n1 * P.k12 + n2 * P.k23_25) + n2 * P.k21;
        // 3<->4 with 4->5 and 3->2: lumped, (n3 + n4)(k45 k34 C2 + k32 k43) / (k34 C2 + k43)
        if (ca > R(3.685e-2)) r += (n3 + n4) * ldiv(P.k45k34 * C2 + P.k32k43, P.k34 * C2 + P.k43);
        else r += n3 * (P.k34 * C2 + P.k32) + n4 * P.k45_43;
        // 5<->6 with 5->2 and 5->4: lumped, (n5 + n6) k65 (k52 + k54 C2) / (k56 C2 + k65)
        if (ca > R(1.15e-4)) r += (n5 + n6) * ldiv(P.k65k52 + P.k65k54 * C2, P.k56 * C2 + P.k65);
        else r += n5 * (P.k56_54 * C2 + P.k52) + n6 * P.k65;
        return s + r;
    }
}

// Rest-state sampler of one CRU: initialise_LCC / _LCC_i / _RyR / _ClCh
// (src/includes/GW_utils.tpp:5-64) and the CRUState constants (GW_utils.tpp:212-213),
// drawing from a stream keyed by the complement of the CRU's SSA stream key.
// Per CRU the draws are LCC[0..3], y[0..3], RyR[0..3][0..4], ClCh[0..3].
template <int PRNG>
GW_HD void sample_rest_cru(uint64_t seed, int64_t i, double ca[4], double &jsr, uint64_t sub[4])
{
    uint64_t s[4];
    rng_seed(PRNG, s, ~stream_key(seed, i));
    auto u = [&]() { return u01(rng_next<PRNG>(s[0], s[1], s[2], s[3])); };
    int lcc[4], y[4], cl[4], n[4][6];
    for (int j = 0; j < 4; ++j) {
        const double v = u();  // sample_weights over {0.958, 0.038, 0.004} -> LCC 1, 2, 7
        int idx = 0;
        double cum = 0.958;
        if (cum < v) { idx = 1; cum += 0.038; if (cum < v) idx = 2; }
        lcc[j] = idx == 0 ? 1 : (idx == 1 ? 2 : 7);
    }
    for (int j = 0; j < 4; ++j) {
        const double v = u();  // {0.9425, 0.0575} -> y = 1, 0
        y[j] = (0.9425 < v) ? 0 : 1;
    }
    for (int j = 0; j < 4; ++j) {
        for (int k = 0; k < 6; ++k) n[j][k] = 0;
        for (int k = 0; k < 5; ++k) {
            const double v = u();  // {0.609, 0.1955, 0.1955} -> RyR states 1, 5, 6
            int idx = 0;
            double cum = 0.609;
            if (cum < v) { idx = 1; cum += 0.5 * 0.391; if (cum < v) idx = 2; }
            n[j][idx == 0 ? 0 : (idx == 1 ? 4 : 5)] += 1;
        }
    }
    for (int j = 0; j < 4; ++j) {
        const double v = u();  // {0.998, 0.002} -> ClCh 0, 1
        cl[j] = (0.998 < v) ? 1 : 0;
    }
    for (int j = 0; j < 4; ++j) {
        ca[j] = 1.45370e-4;
        sub[j] = sub_pack(lcc[j], y[j], cl[j], n[j]);
    }
    jsr = 0.908408;
}

GW_HD void sample_rest_cru_any(int prng, uint64_t seed, int64_t i, double ca[4], double &jsr, uint64_t sub[4])
{
    switch (prng) {
    case 0: sample_rest_cru<0>(seed, i, ca, jsr, sub); break;
    case 1: sample_rest_cru<1>(seed, i, ca, jsr, sub); break;
    case 2: sample_rest_cru<2>(seed, i, ca, jsr, sub); break;
    case 3: sample_rest_cru<3>(seed, i, ca, jsr, sub); break;
    case 4: sample_rest_cru<4>(seed, i, ca, jsr, sub); break;
    default: sample_rest_cru<5>(seed, i, ca, jsr, sub); break;
    }
}

struct Digest {
    uint64_t count, hash;
};

GW_HD void digest_event(Digest &d, uint32_t step_lo, int j, int kind, int detail)
{
    const uint32_t code = static_cast<uint32_t>(j) | (static_cast<uint32_t>(kind) << 2) |
                          ((static_cast<uint32_t>(detail) & 15u) << 5) | (step_lo << 16);
    d.count++;
    d.hash = (d.hash ^ code) * 0x100000001b3ULL;
}

constexpr uint32_t kMaxItersPerStep = 1u << 22;

// ---------------------------------------------------------------------------
// Quad helpers: DPP quad_perm moves within each group of 4 lanes. Every call
// sits in quad-uniform control flow, so the source lanes are active.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ uint32_t qmov32(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ uint64_t qmov64(uint64_t u)
{
    return (static_cast<uint64_t>(qmov32<CTRL>(static_cast<uint32_t>(u >> 32))) << 32) |
           qmov32<CTRL>(static_cast<uint32_t>(u));
}
template <int CTRL>
__device__ __forceinline__ double qmov(double v) { return double_of(qmov64<CTRL>(bits_of(v))); }
template <int CTRL>
__device__ __forceinline__ float qmov(float v) { return __builtin_bit_cast(float, qmov32<CTRL>(__builtin_bit_cast(uint32_t, v))); }
template <int CTRL>
__device__ __forceinline__ int qmovi(int v) { return static_cast<int>(qmov32<CTRL>(static_cast<uint32_t>(v))); }

template <int K, class R>
__device__ __forceinline__ R qget(R v) { return qmov<K * 0x55>(v); }          // lane K of the quad
template <int K>
__device__ __forceinline__ int qgeti(int v) { return qmovi<K * 0x55>(v); }
template <class R>
__device__ __forceinline__ R qnext(R v) { return qmov<0x39>(v); }             // lane (j + 1) & 3
template <class R>
__device__ __forceinline__ R qprev(R v) { return qmov<0x93>(v); }             // lane (j + 3) & 3
__device__ __forceinline__ int qseli(int v, int k)                             // lane k (quad-uniform k)
{
    const int a = qgeti<0>(v), b = qgeti<1>(v), c = qgeti<2>(v), d = qgeti<3>(v);
    return k == 0 ? a : k == 1 ? b : k == 2 ? c : d;
}
__device__ __forceinline__ uint64_t qselu(uint64_t v, int k)
{
    const uint64_t a = qmov64<0x00>(v), b = qmov64<0x55>(v), c = qmov64<0xAA>(v), d = qmov64<0xFF>(v);
    return k == 0 ? a : k == 1 ? b : k == 2 ? c : d;
}
__device__ __forceinline__ double qsel(double v, int k) { return double_of(qselu(bits_of(v), k)); }
__device__ __forceinline__ float qsel(float v, int k)
{
    return __builtin_bit_cast(float, static_cast<uint32_t>(qseli(static_cast<int>(__builtin_bit_cast(uint32_t, v)), k)));
}
template <class R>
__device__ __forceinline__ R qsum4(R v) { return qget<0>(v) + qget<1>(v) + qget<2>(v) + qget<3>(v); }
__device__ __forceinline__ int qsum4i(int v) { return qgeti<0>(v) + qgeti<1>(v) + qgeti<2>(v) + qgeti<3>(v); }

// Replicated CRU-level state of a quad (CaJSR accumulates in FP64 in both forms).
// sample_RyR's choice (SSA.tpp:200-259): the first of the 12 transitions
// whose cumulative rate reaches uu. In FP64 this is the reference's rule as
// it stands. In FP32 the scaled uniform can exceed the cumulative sum by
// rounding, which would take the last transition whatever its rate; there
// only transitions with a positive rate are taken (the last of them on such
// an excess, none (-1) if no rate is positive), so no RyR leaves an empty
// state.
template <class R>
__device__ __forceinline__ int ryr_pick(const SsaParamsT<R> &P, uint64_t w, R ca, R uu)
{
    int tr = sizeof(R) == 8 ? 0 : -1;
    R cum = R(0);
    bool done = false;
    ryr_rates(P, w, ca, [&](int k, R r) {
        if (done) return;
        if (sizeof(R) == 8) {
            if (k == 0) cum = r;
            else if (cum < uu) { tr = k; cum += r; }
            else done = true;
        } else {
            if (k == 0) {
                cum = r;
                if (r > R(0)) tr = 0;
            } else if (cum < uu) {
                if (r > R(0)) tr = k;
                cum += r;
            } else {
                done = true;
            }
        }
    });
    return tr;
}

// Source and target states (0..5) of RyR transition tr (0..11, in
// sample_RyR's order, SSA.tpp:200-259).
GW_HD int ryr_from(int tr)
{
    return (tr == 0) ? 0 : (tr <= 2 || tr == 6) ? 1 : (tr == 3 || tr == 7) ? 2
         : (tr == 4 || tr == 8) ? 3 : (tr == 5 || tr == 9 || tr == 10) ? 4 : 5;
}
GW_HD int ryr_to(int tr)
{
    return (tr == 0 || tr == 7 || tr == 9) ? 1 : (tr == 1 || tr == 8) ? 2 : (tr == 3 || tr == 10) ? 3
         : (tr == 2 || tr == 4 || tr == 11) ? 4 : (tr == 5) ? 5 : 0;
}

// FP32-local only: whether transition tr may be applied to word w at the
// subspace Ca the step has just reached. The transition is chosen from the
// propensities before the Ca update (as in the reference); when that update
// re-split a lumped pair, a transition chosen at a pair-sum rate can take a
// member below zero outside lumped mode, and negative counts make negative
// propensities (a negative waiting time runs the Euler updates backwards).
// Such a transition is not applied: a member leaves only while it holds a
// RyR, or while its pair is lumped and the pair holds one.
GW_HD bool ryr_allowed(uint64_t w, int tr, double ca)
{
    const int from = ryr_from(tr);
    if ((from == 2 || from == 3) && ca > 0.03685) return ryr_get(w, 2) + ryr_get(w, 3) > 0;
    if ((from == 4 || from == 5) && ca > 1.15e-4) return ryr_get(w, 4) + ryr_get(w, 5) > 0;
    return ryr_get(w, from) > 0;
}

// Re-splits due in update_CaSS (SSA.tpp:93-96) as a mask: 1 for the (5,6)
// pair, 2 for the (3,4) pair. The reference re-splits at most one pair per
// update (the (5,6) pair when Ca falls through both thresholds at once); the
// FP32-local form re-splits both then, so no lumped pair is left unsplit
// below its threshold.
template <class R>
GW_HD int split_mask(double oldA, double tmp)
{
    const int m56 = (oldA > 1.15e-4 && tmp <= 1.15e-4) ? 1 : 0;
    const int m34 = (oldA > 0.03685 && tmp <= 0.03685) ? 2 : 0;
    if (sizeof(R) == 8) return m56 ? m56 : m34;
    return m56 | m34;
}

// Whether transition tr keeps both counts it changes within the signed byte
// a subunit word holds them in. Lumped-pair members drift (a transition at a
// pair-sum rate moves one named member), and nothing bounds the drift; the
// reference counts in int, so a count leaving [-128, 127] is reported
// (GW_E_RANGE) instead of wrapping.
GW_HD bool ryr_in_range(uint64_t w, int tr)
{
    return ryr_get(w, ryr_from(tr)) > -128 && ryr_get(w, ryr_to(tr)) < 127;
}

// Apply RyR transition tr (0..11, in sample_RyR's order) to a subunit word.
GW_HD uint64_t ryr_apply(uint64_t w, int tr)
{
    return ryr_add(ryr_add(w, ryr_from(tr), -1), ryr_to(tr), +1);
}

struct CruQ {
    double jsr;
    uint64_t r0, r1, r2, r3;
};

// Cai and CaNSR of a step at the accumulator precision.
GW_HD double kcai(const StepConsts &k) { return k.Cai; }
GW_HD double kcansr(const StepConsts &k) { return k.CaNSR; }
GW_HD double kcai(const StepConstsF &k) { return k.Cai_d; }
GW_HD double kcansr(const StepConstsF &k) { return k.CaNSR_d; }

// Per-CRU outputs of one step for the reductions (identical in the 4 lanes).
struct QuadOut {
    double Jxfer;    // ((Jxfer0 + Jxfer1) + Jxfer2) + Jxfer3 of the last iteration
    double Jtr;      // Jtr of the last iteration
    double JLCC;     // ((JLCC0 + JLCC1) + JLCC2) + JLCC3 after the step
    double Jrel;     // sum of Jrel after the step
    int open;        // open RyRs (states 3 + 4) after the step
    int clch;        // open ClCh after the step
    uint32_t iters;  // SSA loop iterations
    int err;         // nonzero if the iteration cap was hit
    int changed;     // nonzero if this lane's subunit word or JLCC changed (an event or a re-split)
};

// Uniform draws: FP64 uses the reference's 64-bit map; FP32 the top 24 bits.
template <int PRNG, class R>
__device__ __forceinline__ R qdraw(CruQ &q)
{
    if constexpr (sizeof(R) == 8) {
        return u01(rng_next<PRNG>(q.r0, q.r1, q.r2, q.r3));
    } else {
        return static_cast<float>(static_cast<uint32_t>(rng_next<PRNG>(q.r0, q.r1, q.r2, q.r3) >> 40)) * 0x1p-24f;
    }
}

// Two uniforms from one 64-bit draw, bits 40..63 and 16..39 (the one-lane
// FP32-local layout, built for xoshiro256** only, whose scrambled output has
// no weak bit): an event iteration then advances the engine twice, not four
// times.
template <int PRNG>
__device__ __forceinline__ void draw2(CruQ &q, float &hi, float &mid)
{
    const uint64_t x = rng_next<PRNG>(q.r0, q.r1, q.r2, q.r3);
    hi = static_cast<float>(static_cast<uint32_t>(x >> 40)) * 0x1p-24f;
    mid = static_cast<float>(static_cast<uint32_t>(x >> 16) & 0xffffffu) * 0x1p-24f;
}

// One global step of SSA_single_CRU (SSA.tpp:286-310); lane j holds subunit j.
// R is the local arithmetic type (double: the reference path; float: the
// FP32-local variant); KT the step constants of that type. CaSS (`ca`) and
// CaJSR accumulate in FP64 in both; with R = double every expression below is
// the reference's, in its order.
template <class R>
struct QuadState {
    double ca;      // this lane's subunit CaSS (FP64 accumulator in both forms)
    R jl;           // this lane's subunit JLCC
    uint64_t w;     // this lane's subunit word
    CruQ q;         // CaJSR and the CRU's engine (the same in the four lanes)
    Digest dig;     // replay digest (the same in the four lanes)
    int open;
    R Jrel, t, Jtr, Jxfer;
    uint32_t iters;
    int err, changed;
};

// copy_from_CRUState (SSA.tpp:10-33): the step's starting fluxes.
template <class R>
__device__ __forceinline__ void quad_begin(QuadState<R> &S, const SsaParamsT<R> &P)
{
    S.iters = 0;
    S.err = 0;
    S.changed = 0;
    S.open = ryr_get(S.w, 2) + ryr_get(S.w, 3);
    S.Jrel = P.rRyR * static_cast<R>(S.open) * static_cast<R>(S.q.jsr - S.ca);
    S.t = R(0);
    S.Jtr = R(0);
    S.Jxfer = R(0);
}

// One iteration of SSA_single_CRU's loop (SSA.tpp:290-309) for a quad; true
// when the step is finished (the same in the four lanes).
template <int PRNG, bool DIGEST, class R, class KT>
__device__ __forceinline__ bool quad_iterate(QuadState<R> &S, const SsaParamsT<R> &P, const KT &K,
                                             const LccTableT<R> &T, R dt, uint32_t step_lo, int j)
{
    const double Cai = kcai(K), CaNSR = kcansr(K);
    S.iters++;
    // update_fluxes + update_rates (SSA.tpp:37-56)
    const double oldA = S.ca;
    const R old = static_cast<R>(oldA);
    const double o_next = qnext(oldA);
    const double o_prev = qprev(oldA);
    S.Jtr = P.rtr * static_cast<R>(CaNSR - S.q.jsr);
    const R s_me = subunit_total(P, K, T, S.w, old);
    const R s0 = qget<0>(s_me), s1 = qget<1>(s_me), s2 = qget<2>(s_me), s3 = qget<3>(s_me);
    const R total = s0 + s1 + s2 + s3;
    const R jump = ldiv(lneglog(qdraw<PRNG, R>(S.q)), total);
    const bool last = !(S.t + jump < dt);
    const R h = last ? dt - S.t : jump;
    const uint64_t w_pre = S.w;
    // update_CaSS (SSA.tpp:87-99) for this lane's subunit
    const R Jiss = P.riss * static_cast<R>(o_next + o_prev - 2 * oldA);
    S.Jxfer = P.rxfer * static_cast<R>(oldA - Cai);
    const R betaSS = beta_ss(P, old);
    const R dCaSS = h * (S.jl + S.Jrel - S.Jxfer + Jiss) * betaSS;
    const double tmp = oldA + static_cast<double>(dCaSS);
    const int split = split_mask<R>(oldA, tmp);
    const int any_split = qgeti<0>(split) | qgeti<1>(split) | qgeti<2>(split) | qgeti<3>(split);
    if (any_split) {
        // sample_RyR56 / sample_RyR34 (SSA.tpp:59-84), in subunit order; every
        // lane draws the same Bernoulli uniforms and the owner applies them
        for (int jj = 0; jj < 4; ++jj) {
            const int sj = qseli(split, jj);
            for (int bit = 1; bit <= 2; bit <<= 1) {
                if (!(sj & bit)) continue;
                const int a_ = (bit == 1) ? 4 : 2;
                const uint64_t wq = qselu(S.w, jj);
                const R oq = qsel(old, jj);
                const R pr = (bit == 1) ? ldiv(P.k65, P.k65 + P.k56 * sq(oq)) : ldiv(P.k43, P.k43 + P.k34 * sq(oq));
                const int n = ryr_get(wq, a_) + ryr_get(wq, a_ + 1);
                int X = 0;
                for (int i = 0; i < n; ++i)
                    if (qdraw<PRNG, R>(S.q) < pr) X++;
                const int na = (n > 0) ? X : 0;
                const int nb = (n > 0) ? n - X : 0;
                if (j == jj) {
                    S.w = ryr_set(ryr_set(S.w, a_, na), a_ + 1, nb);
                    S.changed = 1;
                }
                if (DIGEST) digest_event(S.dig, step_lo, jj, bit == 1 ? 4 : 5, na);
            }
        }
    }
    S.ca = tmp;
    // update_CaJSR (SSA.hpp:67-71), replicated
    {
        const R Jsum = qsum4(S.Jrel);
        const R betaJSR = beta_jsr(P, static_cast<R>(S.q.jsr));
        S.q.jsr += static_cast<double>(h * betaJSR * (S.Jtr - P.VSS_VJSR * Jsum));
    }
    if (last) return true;
    if (S.iters >= kMaxItersPerStep) {   // no physical state gets here; stop instead of spinning
        S.err = 1;
        return true;
    }
    S.t += jump;
    // sample_weights over the subunits (common.hpp:49-59)
    int js;
    {
        const R u = qdraw<PRNG, R>(S.q) * total;
        R cum = s0;
        js = 0;
        if (cum < u) { js = 1; cum += s1;
            if (cum < u) { js = 2; cum += s2;
                if (cum < u) { js = 3; } } }
    }
    // sample_new_state (SSA.tpp:262-283): every lane classifies its own
    // subunit; the chosen lane's class is broadcast to the quad
    const R u3 = qdraw<PRNG, R>(S.q);
    const int L = sub_lcc(S.w);
    const int y = sub_y(S.w);
    const int cl = sub_cl(S.w);
    const R sum_LCC = T.P[L] + T.Q[L] * old;
    const R inact = y ? K.inact1 : K.inact0;
    const R clr = cl ? P.kbClCh : P.kfClCh * old;
    const R sum_RyR = s_me - (sum_LCC + inact + clr);
    const R u = u3 * s_me;
    const int cls = (u < sum_LCC) ? 0 : (u < (sum_LCC + inact)) ? 1 : (u < (sum_LCC + inact + sum_RyR)) ? 2 : 3;
    const int cls_q = qseli(cls, js);
    const R u4 = (cls_q == 0 || cls_q == 2) ? qdraw<PRNG, R>(S.q) : R(0);
    int detail = 0;
    if (j == js) {
        S.changed = 1;
        const R caR = static_cast<R>(S.ca);
        if (cls == 0) {
            // sample_LCC (SSA.tpp:103-197); the scan never compares after adding r2
            const R r0 = T.R0[L];
            const R r1 = (L == 1) ? T.Q[L] * old : T.R1[L];
            const R uu = u4 * sum_LCC;
            R cum = r0;
            int tr = 0;
            if (cum < uu) { tr = 1; cum += r1;
                if (cum < uu) { tr = 2; } }
            const int Ln = lcc_next(L, tr);
            if ((L == 5 && Ln == 6) || (L == 11 && Ln == 12))
                S.jl = (y == 1) ? K.JLCC_mult * (P.Cao_scaled - K.JLCC_exp * caR) : R(0);
            else if (L == 6 || L == 12)
                S.jl = R(0);
            S.w = sub_set_lcc(S.w, Ln);
            detail = Ln;
        } else if (cls == 1) {
            const int yn = 1 - y;
            S.w = sub_flip_y(S.w);
            if (yn == 0) S.jl = R(0);
            else if (L == 6 || L == 12) S.jl = K.JLCC_mult * (P.Cao_scaled - K.JLCC_exp * caR);
            detail = yn;
        } else if (cls == 2) {
            // sample_RyR (SSA.tpp:200-259) over the propensities before this iteration's update
            int tr = ryr_pick(P, w_pre, old, u4 * sum_RyR);
            if (sizeof(R) == 4 && tr >= 0 && !ryr_allowed(S.w, tr, S.ca)) tr = -1;
            if (tr >= 0) {
                if (!ryr_in_range(S.w, tr)) S.err = 1;
                S.w = ryr_apply(S.w, tr);
            }
            if (tr == 1 || tr == 10) S.open += 1;
            else if (tr == 4 || tr == 7) S.open -= 1;
            S.Jrel = P.rRyR * static_cast<R>(S.open) * static_cast<R>(S.q.jsr - S.ca);
            detail = tr;
        } else {
            S.w = sub_flip_cl(S.w);
            detail = 1 - cl;
        }
    }
    if (DIGEST) digest_event(S.dig, step_lo, js, cls_q, qseli(detail, js));
    return false;
}

// outputs of update_CRUstate_from_temp (GW.tpp:47-62) for the reductions;
// FP32 values enter the FP64 sums exactly
template <class R>
__device__ __forceinline__ QuadOut quad_outputs(const QuadState<R> &S)
{
    QuadOut out;
    out.Jxfer = qsum4(static_cast<double>(S.Jxfer));
    out.Jtr = static_cast<double>(S.Jtr);
    out.JLCC = qsum4(static_cast<double>(S.jl));
    out.Jrel = qsum4(static_cast<double>(S.Jrel));
    out.open = qsum4i(ryr_get(S.w, 2) + ryr_get(S.w, 3));
    out.clch = qsum4i(sub_cl(S.w));
    out.iters = S.iters;
    out.err = S.err;
    out.changed = S.changed;
    return out;
}

template <int PRNG, bool DIGEST, class R, class KT>
__device__ QuadOut ssa_step_quad(double &ca, R &jl, uint64_t &w, CruQ &q, const SsaParamsT<R> &P,
                                 const KT &K, const LccTableT<R> &T, R dt, Digest &dig, uint32_t step_lo, int j)
{
    QuadState<R> S;
    S.ca = ca;
    S.jl = jl;
    S.w = w;
    S.q = q;
    S.dig = dig;
    quad_begin(S, P);
    while (!quad_iterate<PRNG, DIGEST>(S, P, K, T, dt, step_lo, j)) {
    }
    ca = S.ca;
    jl = S.jl;
    w = S.w;
    q = S.q;
    dig = S.dig;
    return quad_outputs(S);
}

// ---------------------------------------------------------------------------
// One lane per CRU (FP32-local variant). The lane visits its four subunits in
// order, as SSA_single_CRU itself does (SSA.tpp:37-99, 262-283); the sampled
// subunit's values are read and written through selects, so no register array
// is indexed with a run-time index.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T sel4(const T (&v)[4], int k) { return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : v[3]; }
template <class T>
__device__ __forceinline__ void put4(T (&v)[4], int k, T x)
{
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (k == j) v[j] = x;
}

struct LaneOut {
    double Jxfer, Jtr, JLCC, Jrel;   // per-CRU sums in subunit order
    int open, clch;
    uint32_t iters;
    int err;
};

// The working state of one CRU during one global step.
struct LaneState {
    double ca[4];
    uint64_t w[4];
    CruQ q;
    float jl[4];
    float Jrel[4];
    float Jxfer[4];
    int open[4];
    float t, Jtr;
    uint32_t iters;
    int err;
    int changed;   // the subunit words or JLCC changed this step (an event or a re-split)
};

// copy_from_CRUState (SSA.tpp:10-33): the step's starting fluxes.
__device__ __forceinline__ void lane_begin(LaneState &S, const SsaParamsF &P)
{
#pragma clang fp contract(fast)   // FP32-local only: fused multiply-adds allowed
    using R = float;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        S.open[j] = ryr_get(S.w[j], 2) + ryr_get(S.w[j], 3);
        S.Jrel[j] = P.rRyR * static_cast<R>(S.open[j]) * static_cast<R>(S.q.jsr - S.ca[j]);
        S.Jxfer[j] = R(0);
    }
    S.t = R(0);
    S.Jtr = R(0);
    S.iters = 0;
    S.err = 0;
    S.changed = 0;
}

// One iteration of SSA_single_CRU's loop (SSA.tpp:290-309); true when the
// step is finished. The event (subunit, class, transition) is chosen from the
// propensities before this iteration's Ca update, as in the reference, but
// its uniforms are drawn before the re-split Bernoulli draws of that update
// rather than after them: the draws are independent, so the joint law of the
// choice and the re-splits is the same, and only the choice (not the
// pre-update channel words and totals) has to be kept across the update.
// The four uniforms of an event iteration (waiting time, subunit, class,
// transition) and the Bernoulli trials of a re-split come two per engine draw
// (draw2).
template <int PRNG>
__device__ __forceinline__ bool lane_iterate(LaneState &S, const SsaParamsF &P, const StepConstsF &K,
                                             const LccTableF &T, float dt)
{
#pragma clang fp contract(fast)   // FP32-local only: fused multiply-adds allowed
    using R = float;
    const double Cai = K.Cai_d, CaNSR = K.CaNSR_d;
    S.iters++;
    // update_fluxes + update_rates (SSA.tpp:37-56)
    double oldA[4];
    R old[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        oldA[j] = S.ca[j];
        old[j] = static_cast<R>(oldA[j]);
    }
    S.Jtr = P.rtr * static_cast<R>(CaNSR - S.q.jsr);
    int js = 0, cls = 0, tr = 0;
    R jump, h;
    bool last;
    {
        R s[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) s[j] = subunit_total(P, K, T, S.w[j], old[j]);
        const R total = s[0] + s[1] + s[2] + s[3];
        R u_jump, u_sub;
        draw2<PRNG>(S.q, u_jump, u_sub);
        jump = ldiv(lneglog(u_jump), total);
        last = !(S.t + jump < dt);
        h = last ? dt - S.t : jump;
        if (!last) {
            // sample_weights over the subunits (common.hpp:49-59)
            const R u = u_sub * total;
            R cum = s[0];
            if (cum < u) { js = 1; cum += s[1];
                if (cum < u) { js = 2; cum += s[2];
                    if (cum < u) { js = 3; } } }
            // the class and transition of sample_new_state (SSA.tpp:262-283) on subunit js
            const uint64_t wj = sel4(S.w, js);
            const R oldj = sel4(old, js);
            const R sj = sel4(s, js);
            const int L = sub_lcc(wj);
            const R sum_LCC = T.P[L] + T.Q[L] * oldj;
            const R inact = sub_y(wj) ? K.inact1 : K.inact0;
            const R clr = sub_cl(wj) ? P.kbClCh : P.kfClCh * oldj;
            const R sum_RyR = sj - (sum_LCC + inact + clr);
            R u_cls, u_tr;
            draw2<PRNG>(S.q, u_cls, u_tr);
            const R u3 = u_cls * sj;
            cls = (u3 < sum_LCC) ? 0 : (u3 < (sum_LCC + inact)) ? 1 : (u3 < (sum_LCC + inact + sum_RyR)) ? 2 : 3;
            if (cls == 0) {
                // sample_LCC (SSA.tpp:103-197)
                const R r0 = T.R0[L];
                const R r1 = (L == 1) ? T.Q[L] * oldj : T.R1[L];
                const R uu = u_tr * sum_LCC;
                R c = r0;
                if (c < uu) { tr = 1; c += r1;
                    if (c < uu) { tr = 2; } }
            } else if (cls == 2) {
                tr = ryr_pick(P, wj, oldj, u_tr * sum_RyR);
            }
        }
    }
    // update_CaSS (SSA.tpp:87-99): subunits in order, re-splits drawn as they come
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const R Jiss = P.riss * static_cast<R>(oldA[(j + 1) & 3] + oldA[(j + 3) & 3] - 2 * oldA[j]);
        S.Jxfer[j] = P.rxfer * static_cast<R>(oldA[j] - Cai);
        const R betaSS = beta_ss(P, old[j]);
        const R dCaSS = h * (S.jl[j] + S.Jrel[j] - S.Jxfer[j] + Jiss) * betaSS;
        const double tmp = oldA[j] + static_cast<double>(dCaSS);
        int split = split_mask<R>(oldA[j], tmp);
        while (split) {
            // sample_RyR56 / sample_RyR34 (SSA.tpp:59-84), the (5,6) pair first
            const int bit = split & -split;
            split ^= bit;
            const int a_ = (bit == 1) ? 4 : 2;
            const R pr = (bit == 1) ? ldiv(P.k65, P.k65 + P.k56 * sq(old[j])) : ldiv(P.k43, P.k43 + P.k34 * sq(old[j]));
            const int n = ryr_get(S.w[j], a_) + ryr_get(S.w[j], a_ + 1);
            int X = 0;
            for (int i = 0; i < n; i += 2) {   // two Bernoulli trials per engine draw
                R b0, b1;
                draw2<PRNG>(S.q, b0, b1);
                X += (b0 < pr) ? 1 : 0;
                X += (i + 1 < n && b1 < pr) ? 1 : 0;
            }
            S.w[j] = ryr_set(ryr_set(S.w[j], a_, (n > 0) ? X : 0), a_ + 1, (n > 0) ? n - X : 0);
            S.changed = 1;
        }
        S.ca[j] = tmp;
    }
    // update_CaJSR (SSA.hpp:67-71)
    {
        const R Jsum = S.Jrel[0] + S.Jrel[1] + S.Jrel[2] + S.Jrel[3];
        const R betaJSR = beta_jsr(P, static_cast<R>(S.q.jsr));
        S.q.jsr += static_cast<double>(h * betaJSR * (S.Jtr - P.VSS_VJSR * Jsum));
    }
    if (last) return true;
    if (S.iters >= kMaxItersPerStep) {   // no physical state gets here; stop instead of spinning
        S.err = 1;
        return true;
    }
    S.t += jump;
    S.changed = 1;
    // apply the chosen event to subunit js after the update (SSA.tpp:262-283)
    uint64_t wj = sel4(S.w, js);
    const double caj_d = sel4(S.ca, js);
    const R caj = static_cast<R>(caj_d);
    R jlj = sel4(S.jl, js);
    R Jrelj = sel4(S.Jrel, js);
    int openj = sel4(S.open, js);
    const int L = sub_lcc(wj);
    const int y = sub_y(wj);
    if (cls == 0) {
        const int Ln = lcc_next(L, tr);
        if ((L == 5 && Ln == 6) || (L == 11 && Ln == 12))
            jlj = (y == 1) ? K.JLCC_mult * (P.Cao_scaled - K.JLCC_exp * caj) : R(0);
        else if (L == 6 || L == 12)
            jlj = R(0);
        wj = sub_set_lcc(wj, Ln);
    } else if (cls == 1) {
        const int yn = 1 - y;
        wj = sub_flip_y(wj);
        if (yn == 0) jlj = R(0);
        else if (L == 6 || L == 12) jlj = K.JLCC_mult * (P.Cao_scaled - K.JLCC_exp * caj);
    } else if (cls == 2) {
        // sample_RyR (SSA.tpp:200-259): the transition chosen above, unless it
        // would take a RyR out of an empty state (ryr_allowed)
        if (tr >= 0 && ryr_allowed(wj, tr, caj_d)) {
            if (!ryr_in_range(wj, tr)) S.err = 1;
            wj = ryr_apply(wj, tr);
            if (tr == 1 || tr == 10) openj += 1;
            else if (tr == 4 || tr == 7) openj -= 1;
        }
        Jrelj = P.rRyR * static_cast<R>(openj) * static_cast<R>(S.q.jsr - caj_d);
    } else {
        wj = sub_flip_cl(wj);
    }
    put4(S.w, js, wj);
    put4(S.jl, js, jlj);
    put4(S.Jrel, js, Jrelj);
    put4(S.open, js, openj);
    return false;
}

// outputs of update_CRUstate_from_temp (GW.tpp:47-62), summed in subunit order in FP64
__device__ __forceinline__ LaneOut lane_outputs(const LaneState &S)
{
    LaneOut out;
    out.Jxfer = ((static_cast<double>(S.Jxfer[0]) + S.Jxfer[1]) + S.Jxfer[2]) + S.Jxfer[3];
    out.Jtr = static_cast<double>(S.Jtr);
    out.JLCC = ((static_cast<double>(S.jl[0]) + S.jl[1]) + S.jl[2]) + S.jl[3];
    out.Jrel = ((static_cast<double>(S.Jrel[0]) + S.Jrel[1]) + S.Jrel[2]) + S.Jrel[3];
    out.open = 0;
    out.clch = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        out.open += ryr_get(S.w[j], 2) + ryr_get(S.w[j], 3);
        out.clch += sub_cl(S.w[j]);
    }
    out.iters = S.iters;
    out.err = S.err;
    return out;
}

}  // namespace gw
