// gw_ssa.h -- per-CRU Gillespie SSA over one global step, FP64, on one lane.
//
// Restates SSA_single_CRU and its helpers (src/includes/SSA.tpp:10-310,
// SSA.hpp:67-71, GW_utils.tpp:118-255, GW_utils.hpp:378-391) for a GPU lane:
//   - the CRU lives in registers: CaSS[4], CaJSR, JLCC[4], packed channel
//     states and four packed RyR count words (six signed bytes per subunit);
//   - only the four subunit totals are kept per iteration; the individual
//     propensities of the chosen subunit are recomputed when an event is
//     sampled, from the same inputs the reference's stored rate arrays were
//     formed from (CaSS before this iteration's update, RyR counts before any
//     threshold re-split), so every comparison sees identical values;
//   - all arithmetic keeps the reference's operation order (bit-exact with
//     the CPU oracle when compiled with -ffp-contract=off).
#pragma once

#include "gw_model.h"

namespace gw {

// ---- packed channel state ----------------------------------------------
// chan: bits 4j..4j+3 = LCC state (1..12) of subunit j; bit 16+j = y (LCC
// voltage-inactivation gate, 1 = available); bit 20+j = ClCh open.
GW_HD int chan_lcc(uint32_t c, int j) { return static_cast<int>((c >> (4 * j)) & 15u); }
GW_HD int chan_y(uint32_t c, int j) { return static_cast<int>((c >> (16 + j)) & 1u); }
GW_HD int chan_cl(uint32_t c, int j) { return static_cast<int>((c >> (20 + j)) & 1u); }
GW_HD uint32_t chan_set_lcc(uint32_t c, int j, int s) { return (c & ~(15u << (4 * j))) | (static_cast<uint32_t>(s) << (4 * j)); }
GW_HD uint32_t chan_flip_y(uint32_t c, int j) { return c ^ (1u << (16 + j)); }
GW_HD uint32_t chan_flip_cl(uint32_t c, int j) { return c ^ (1u << (20 + j)); }

// ---- packed RyR counts: byte k (k = 0..5) = signed count in state k+1 ---
GW_HD int ryr_get(uint64_t w, int k) { return static_cast<int>(static_cast<int8_t>((w >> (8 * k)) & 0xff)); }
GW_HD uint64_t ryr_set(uint64_t w, int k, int v)
{
    return (w & ~(0xffULL << (8 * k))) | (static_cast<uint64_t>(static_cast<uint8_t>(static_cast<int8_t>(v))) << (8 * k));
}
GW_HD uint64_t ryr_add(uint64_t w, int k, int d) { return ryr_set(w, k, ryr_get(w, k) + d); }

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

// Per-cell parameters the SSA reads (a subset of gw_params_t and CellConsts).
struct SsaParams {
    double rRyR, rxfer, rtr, riss;
    double KBSR, KBSL, KCSQN, BSR_const, BSL_const, CSQN_const, VSS_VJSR;
    double k12, k21, k23, k32, k34, k43, k45, k54, k56, k65, k25, k52;
    double kfClCh, kbClCh, Cao_scaled;
};

GW_HD void ssa_params_init(SsaParams &s, const gw_params_t &p, const CellConsts &c)
{
    s.rRyR = p.rRyR; s.rxfer = p.rxfer; s.rtr = p.rtr; s.riss = p.riss;
    s.KBSR = p.KBSR; s.KBSL = p.KBSL; s.KCSQN = p.KCSQN;
    s.BSR_const = c.BSR_const; s.BSL_const = c.BSL_const; s.CSQN_const = c.CSQN_const; s.VSS_VJSR = c.VSS_VJSR;
    s.k12 = p.k12; s.k21 = p.k21; s.k23 = p.k23; s.k32 = p.k32; s.k34 = p.k34; s.k43 = p.k43;
    s.k45 = p.k45; s.k54 = p.k54; s.k56 = p.k56; s.k65 = p.k65; s.k25 = p.k25; s.k52 = p.k52;
    s.kfClCh = p.kfClCh; s.kbClCh = p.kbClCh; s.Cao_scaled = c.Cao_scaled;
}

// The 12 RyR propensities of one subunit (update_RyR_rates, GW_utils.tpp:222-255),
// visited in storage order; `f(k, r)` receives each rate.
template <class F>
GW_HD void ryr_rates(const SsaParams &P, uint64_t w, double ca, F &&f)
{
    const double C2 = ca * ca;
    const int n1 = ryr_get(w, 0), n2 = ryr_get(w, 1), n3 = ryr_get(w, 2);
    const int n4 = ryr_get(w, 3), n5 = ryr_get(w, 4), n6 = ryr_get(w, 5);
    const bool hi34 = ca > 3.685e-2;
    const bool hi56 = ca > 1.15e-4;
    double tau34 = 0.0, eq56 = 0.0;
    if (hi34) tau34 = 1.0 / (P.k34 * C2 + P.k43);
    if (hi56) eq56 = P.k65 / (P.k56 * C2 + P.k65);
    f(0, static_cast<double>(n1) * P.k12 * C2);
    f(1, static_cast<double>(n2) * P.k23 * C2);
    f(2, static_cast<double>(n2) * P.k25 * C2);
    f(3, hi34 ? 0.0 : static_cast<double>(n3) * P.k34 * C2);
    f(4, hi34 ? static_cast<double>(n3 + n4) * P.k45 * P.k34 * C2 * tau34 : static_cast<double>(n4) * P.k45);
    f(5, hi56 ? 0.0 : static_cast<double>(n5) * P.k56 * C2);
    f(6, static_cast<double>(n2) * P.k21);
    f(7, hi34 ? static_cast<double>(n3 + n4) * P.k32 * P.k43 * tau34 : static_cast<double>(n3) * P.k32);
    f(8, hi34 ? 0.0 : static_cast<double>(n4) * P.k43);
    f(9, hi56 ? static_cast<double>(n5 + n6) * P.k52 * eq56 : static_cast<double>(n5) * P.k52);
    f(10, hi56 ? static_cast<double>(n5 + n6) * P.k54 * C2 * eq56 : static_cast<double>(n5) * P.k54 * C2);
    f(11, hi56 ? 0.0 : static_cast<double>(n6) * P.k65);
}

// Subunit total in the reference accumulation order (SSA.tpp:48-56):
// 0 + LCC sum, + inactivation, + RyR[0..11], + ClCh.
GW_HD double subunit_total(const SsaParams &P, const StepConsts &K, const LccTable &T, uint32_t chan, uint64_t w,
                           int j, double ca)
{
    const int L = chan_lcc(chan, j);
    double s = T.P[L] + T.Q[L] * ca;
    s += chan_y(chan, j) ? K.inact1 : K.inact0;
    ryr_rates(P, w, ca, [&](int, double r) { s += r; });
    s += chan_cl(chan, j) ? P.kbClCh : P.kfClCh * ca;
    return s;
}

// Rest-state sampler of one CRU: initialise_LCC / _LCC_i / _RyR / _ClCh
// (src/includes/GW_utils.tpp:5-64) and the CRUState constants (GW_utils.tpp:212-213),
// drawing from a stream keyed by the complement of the CRU's SSA stream key.
// Per CRU the draws are LCC[0..3], y[0..3], RyR[0..3][0..4], ClCh[0..3].
template <int PRNG>
GW_HD void sample_rest_cru(uint64_t seed, int64_t i, double ca[4], double &jsr, uint32_t &chan, uint64_t ryr[4])
{
    uint64_t s[4];
    rng_seed(PRNG, s, ~stream_key(seed, i));
    auto u = [&]() { return u01(rng_next<PRNG>(s[0], s[1], s[2], s[3])); };
    chan = 0;
    for (int j = 0; j < 4; ++j) {
        const double v = u();  // sample_weights over {0.958, 0.038, 0.004} -> LCC 1, 2, 7
        int idx = 0;
        double cum = 0.958;
        if (cum < v) { idx = 1; cum += 0.038; if (cum < v) idx = 2; }
        chan = chan_set_lcc(chan, j, idx == 0 ? 1 : (idx == 1 ? 2 : 7));
    }
    for (int j = 0; j < 4; ++j) {
        const double v = u();  // {0.9425, 0.0575} -> y = 1, 0
        if (!(0.9425 < v)) chan |= 1u << (16 + j);
    }
    for (int j = 0; j < 4; ++j) {
        uint64_t w = 0;
        for (int k = 0; k < 5; ++k) {
            const double v = u();  // {0.609, 0.1955, 0.1955} -> RyR states 1, 5, 6
            int idx = 0;
            double cum = 0.609;
            if (cum < v) { idx = 1; cum += 0.5 * 0.391; if (cum < v) idx = 2; }
            w = ryr_add(w, idx == 0 ? 0 : (idx == 1 ? 4 : 5), 1);
        }
        ryr[j] = w;
    }
    for (int j = 0; j < 4; ++j) {
        const double v = u();  // {0.998, 0.002} -> ClCh 0, 1
        if (0.998 < v) chan |= 1u << (20 + j);
    }
    for (int j = 0; j < 4; ++j) ca[j] = 1.45370e-4;
    jsr = 0.908408;
}

GW_HD void sample_rest_cru_any(int prng, uint64_t seed, int64_t i, double ca[4], double &jsr, uint32_t &chan,
                               uint64_t ryr[4])
{
    switch (prng) {
    case 0: sample_rest_cru<0>(seed, i, ca, jsr, chan, ryr); break;
    case 1: sample_rest_cru<1>(seed, i, ca, jsr, chan, ryr); break;
    case 2: sample_rest_cru<2>(seed, i, ca, jsr, chan, ryr); break;
    case 3: sample_rest_cru<3>(seed, i, ca, jsr, chan, ryr); break;
    case 4: sample_rest_cru<4>(seed, i, ca, jsr, chan, ryr); break;
    default: sample_rest_cru<5>(seed, i, ca, jsr, chan, ryr); break;
    }
}

// Register-resident CRU.
struct Cru64 {
    double ca[4];
    double jsr;
    double jl[4];
    uint32_t chan;
    uint64_t ryr[4];
    uint64_t r0, r1, r2, r3;   // engine state
};

// Per-CRU outputs of one step for the global reductions.
struct CruOut {
    double Jxfer;   // ((Jxfer0 + Jxfer1) + Jxfer2) + Jxfer3 of the last iteration
    double Jtr;     // Jtr of the last iteration
    double JLCC;    // ((JLCC0 + JLCC1) + JLCC2) + JLCC3 after the step
    double Jrel;    // sum of Jrel after the step (observable)
    int open;       // open RyRs (states 3 + 4) after the step
    int clch;       // open ClCh after the step
    uint32_t iters; // SSA loop iterations
    int err;        // nonzero if the iteration cap was hit
};

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

template <int PRNG>
GW_HD double draw(Cru64 &c) { return u01(rng_next<PRNG>(c.r0, c.r1, c.r2, c.r3)); }

// 4-way register select helpers (no dynamically indexed private arrays).
GW_HD double sel4(const double *a, int j) { return j == 0 ? a[0] : j == 1 ? a[1] : j == 2 ? a[2] : a[3]; }
GW_HD uint64_t sel4(const uint64_t *a, int j) { return j == 0 ? a[0] : j == 1 ? a[1] : j == 2 ? a[2] : a[3]; }

constexpr uint32_t kMaxItersPerStep = 1u << 22;

// One global step of SSA_single_CRU (SSA.tpp:286-310) on a register-resident CRU.
template <int PRNG, bool DIGEST>
GW_HD CruOut ssa_step(Cru64 &c, const SsaParams &P, const StepConsts &K, const LccTable &T, double dt, Digest &dig,
                      uint32_t step_lo)
{
    CruOut out;
    out.iters = 0;
    out.err = 0;
    // copy_from_CRUState (SSA.tpp:10-33): open counts and Jrel at the step start
    double Jrel[4];
    int open[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        open[j] = ryr_get(c.ryr[j], 2) + ryr_get(c.ryr[j], 3);
        Jrel[j] = P.rRyR * static_cast<double>(open[j]) * (c.jsr - c.ca[j]);
    }
    double t = 0.0;
    double old[4];
    double Jtr = 0.0;
    while (true) {
        out.iters++;
        // update_fluxes + update_rates (SSA.tpp:37-56)
#pragma unroll
        for (int j = 0; j < 4; ++j) old[j] = c.ca[j];
        Jtr = P.rtr * (K.CaNSR - c.jsr);
        double s[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) s[j] = subunit_total(P, K, T, c.chan, c.ryr[j], j, old[j]);
        const double total = s[0] + s[1] + s[2] + s[3];
        const double jump = -gw::log(draw<PRNG>(c)) / total;
        const bool last = !(t + jump < dt);
        const double h = last ? dt - t : jump;
        const uint64_t ryr_pre[4] = {c.ryr[0], c.ryr[1], c.ryr[2], c.ryr[3]};
        // update_CaSS (SSA.tpp:87-99)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double Jiss = P.riss * (old[(j + 1) & 3] + old[(j + 3) & 3] - 2 * old[j]);
            const double Jxfer = P.rxfer * (old[j] - K.Cai);
            const double betaSS = 1.0 / (1 + (P.BSR_const / sq(P.KBSR + old[j])) + (P.BSL_const / sq(P.KBSL + old[j])));
            const double dCaSS = h * (c.jl[j] + Jrel[j] - Jxfer + Jiss) * betaSS;
            const double tmp = old[j] + dCaSS;
            if (old[j] > 1.15e-4 && tmp <= 1.15e-4) {
                // sample_RyR56 (SSA.tpp:59-71)
                uint64_t w = c.ryr[j];
                const int n56 = ryr_get(w, 4) + ryr_get(w, 5);
                if (n56 > 0) {
                    const double pr = P.k65 / (P.k65 + P.k56 * sq(old[j]));
                    int X = 0;
                    for (int i = 0; i < n56; ++i)
                        if (draw<PRNG>(c) < pr) X++;
                    w = ryr_set(ryr_set(w, 4, X), 5, n56 - X);
                } else {
                    w = ryr_set(ryr_set(w, 4, 0), 5, 0);
                }
                c.ryr[j] = w;
                if (DIGEST) digest_event(dig, step_lo, j, 4, ryr_get(w, 4));
            } else if (old[j] > 0.03685 && tmp <= 0.03685) {
                // sample_RyR34 (SSA.tpp:73-84)
                uint64_t w = c.ryr[j];
                const int n34 = ryr_get(w, 2) + ryr_get(w, 3);
                if (n34 > 0) {
                    const double pr = P.k43 / (P.k43 + P.k34 * sq(old[j]));
                    int X = 0;
                    for (int i = 0; i < n34; ++i)
                        if (draw<PRNG>(c) < pr) X++;
                    w = ryr_set(ryr_set(w, 2, X), 3, n34 - X);
                } else {
                    w = ryr_set(ryr_set(w, 2, 0), 3, 0);
                }
                c.ryr[j] = w;
                if (DIGEST) digest_event(dig, step_lo, j, 5, ryr_get(w, 2));
            }
            c.ca[j] = tmp;
        }
        // update_CaJSR (SSA.hpp:67-71)
        {
            const double betaJSR = 1.0 / (1 + P.CSQN_const / sq(P.KCSQN + c.jsr));
            c.jsr += (h * betaJSR * (Jtr - P.VSS_VJSR * (Jrel[0] + Jrel[1] + Jrel[2] + Jrel[3])));
        }
        if (last) break;
        if (out.iters >= kMaxItersPerStep) {   // no physical state gets here; stop instead of spinning
            out.err = 1;
            break;
        }
        t += jump;
        // sample_weights over the subunits (common.hpp:49-59)
        int j;
        {
            const double u = draw<PRNG>(c) * total;
            double cum = s[0];
            j = 0;
            if (cum < u) { j = 1; cum += s[1];
                if (cum < u) { j = 2; cum += s[2];
                    if (cum < u) { j = 3; } } }
        }
        // sample_new_state (SSA.tpp:262-283) with the propensities as stored
        // before this iteration's CaSS update
        const double caj = sel4(old, j);
        const double sj = sel4(s, j);
        const int L = chan_lcc(c.chan, j);
        const int y = chan_y(c.chan, j);
        const int cl = chan_cl(c.chan, j);
        const double sum_LCC = T.P[L] + T.Q[L] * caj;
        const double inact = y ? K.inact1 : K.inact0;
        const double clr = cl ? P.kbClCh : P.kfClCh * caj;
        const double sum_RyR = sj - (sum_LCC + inact + clr);
        const double u = draw<PRNG>(c) * sj;
        const double ca_new = sel4(c.ca, j);
        double jl_new = sel4(c.jl, j);
        if (u < sum_LCC) {
            // sample_LCC (SSA.tpp:103-197)
            // the scan over (r0, r1, r2) never compares after adding r2, so r2 is not needed
            const double r0 = T.R0[L];
            const double r1 = (L == 1) ? T.Q[L] * caj : T.R1[L];
            const double uu = draw<PRNG>(c) * sum_LCC;
            double cum = r0;
            int tr = 0;
            if (cum < uu) { tr = 1; cum += r1;
                if (cum < uu) { tr = 2; } }
            const int Ln = lcc_next(L, tr);
            if ((L == 5 && Ln == 6) || (L == 11 && Ln == 12))
                jl_new = (y == 1) ? K.JLCC_mult * (P.Cao_scaled - K.JLCC_exp * ca_new) : 0.0;
            else if (L == 6 || L == 12)
                jl_new = 0.0;
            c.chan = chan_set_lcc(c.chan, j, Ln);
            if (DIGEST) digest_event(dig, step_lo, j, 0, Ln);
        } else if (u < (sum_LCC + inact)) {
            const int yn = 1 - y;
            c.chan = chan_flip_y(c.chan, j);
            if (yn == 0) jl_new = 0.0;
            else if (L == 6 || L == 12) jl_new = K.JLCC_mult * (P.Cao_scaled - K.JLCC_exp * ca_new);
            if (DIGEST) digest_event(dig, step_lo, j, 1, yn);
        } else if (u < (sum_LCC + inact + sum_RyR)) {
            // sample_RyR (SSA.tpp:200-259)
            const double uu = draw<PRNG>(c) * sum_RyR;
            int tr = 0;
            double cum = 0.0;
            bool done = false;
            ryr_rates(P, sel4(ryr_pre, j), caj, [&](int k, double r) {
                if (done) return;
                if (k == 0) { cum = r; }
                else { if (cum < uu) { tr = k; cum += r; } else done = true; }
            });
            // transition (from, to) in storage order, applied to the current counts
            const int from = (tr == 0) ? 0 : (tr <= 2 || tr == 6) ? 1 : (tr == 3 || tr == 7) ? 2
                           : (tr == 4 || tr == 8) ? 3 : (tr == 5 || tr == 9 || tr == 10) ? 4 : 5;
            const int to = (tr == 0 || tr == 7 || tr == 9) ? 1 : (tr == 1 || tr == 8) ? 2 : (tr == 3 || tr == 10) ? 3
                         : (tr == 2 || tr == 4 || tr == 11) ? 4 : (tr == 5) ? 5 : 0;
            uint64_t w = sel4(c.ryr, j);
            w = ryr_add(ryr_add(w, from, -1), to, +1);
            int dopen = 0;
            if (tr == 1 || tr == 10) dopen = 1;
            else if (tr == 4 || tr == 7) dopen = -1;
            const int oj = (j == 0 ? open[0] : j == 1 ? open[1] : j == 2 ? open[2] : open[3]) + dopen;
            const double jr = P.rRyR * static_cast<double>(oj) * (c.jsr - ca_new);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q == j) { c.ryr[q] = w; open[q] = oj; Jrel[q] = jr; }
            if (DIGEST) digest_event(dig, step_lo, j, 2, tr);
        } else {
            c.chan = chan_flip_cl(c.chan, j);
            if (DIGEST) digest_event(dig, step_lo, j, 3, 1 - cl);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q == j) c.jl[q] = jl_new;
    }
    // outputs of update_CRUstate_from_temp (GW.tpp:47-62) for the reductions
    double jx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) jx[j] = P.rxfer * (old[j] - K.Cai);
    out.Jxfer = jx[0] + jx[1] + jx[2] + jx[3];
    out.Jtr = Jtr;
    out.JLCC = c.jl[0] + c.jl[1] + c.jl[2] + c.jl[3];
    out.Jrel = Jrel[0] + Jrel[1] + Jrel[2] + Jrel[3];
    int op = 0, cl = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        op += ryr_get(c.ryr[j], 2) + ryr_get(c.ryr[j], 3);
        cl += chan_cl(c.chan, j);
    }
    out.open = op;
    out.clch = cl;
    return out;
}

}  // namespace gw
