// gw_capi.hip -- kernels of the MI355X Greenstein-Winslow integrator and the
// C-ABI of include/gw_capi.h.
//
// Device layout: every CRU of every cell owns one slot. Per-subunit words are
// stored [slot][4] (CaSS, JLCC, the packed subunit word holding the six RyR
// counts, LCC state, y gate and ClCh); CaJSR is [slot]; the engine state is
// [slot][4] words and the replay digest [slot][2]. A CRU is advanced by a quad
// of lanes (lane j owns subunit j), so one 256-thread block holds 64 CRUs; each
// cell's slots are padded to whole blocks, so a block belongs to one cell.
//
// One Euler step (GW_model::euler_step, src/includes/GW.tpp:236-272) is one
// launch of k_step. Its SSA blocks advance every CRU through SSA(dt) with the
// cell's frozen constants and reduce the flux sums exactly (128-bit fixed
// point) into one partial per block. Its first blocks, one per cell, close
// the previous step at the same time: they sum that step's partials, apply
// globals += dGlobals, form the derivatives of the new step and the SSA
// constants of the step after it. This works because SSA(n+1) needs only the
// globals of step n+1 = globals(n) + dGlobals(n), and dGlobals(n) was formed
// from the sums of step n-1; the sums of step n enter dGlobals(n+1) only.
// Step constants and block partials are double-buffered by step parity.
#include "gw_kernels.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

using namespace gw;

namespace {

// ===========================================================================
// device helpers
// ===========================================================================
__device__ double protocol_istim(const DevCell &C, double t, int64_t step)
{
    switch (C.proto.kind) {
    case GW_PROTO_TABLE:
        return (C.istim && step < C.istim_len) ? C.istim[step] : 0.0;
    case GW_PROTO_PACE: {
        const double x = t - C.proto.start;
        if (x < 0.0) return 0.0;
        const double ph = (C.proto.period > 0.0 && C.proto.period < 1e300) ? fmod(x, C.proto.period) : x;
        return ph < C.proto.duration ? C.proto.amplitude : 0.0;
    }
    default:
        return 0.0;
    }
}

__device__ double protocol_vclamp(const DevCell &C, double t)
{
    return (t >= C.proto.t_on && t < C.proto.t_off) ? C.proto.step : C.proto.hold;
}

__device__ void obs_update(gw_cell_obs_t &o, double V_old, double V, double t_old, double dt, double Cai, double ICaL,
                           double open_frac, double jrel_scaled)
{
    o.peak_V = fmax(o.peak_V, V);
    o.min_V = fmin(o.min_V, V);
    o.peak_Cai = fmax(o.peak_Cai, Cai);
    o.peak_ICaL_neg = fmin(o.peak_ICaL_neg, ICaL);
    o.sum_open_RyR_frac += open_frac;
    o.int_Jrel += jrel_scaled * dt;
    o.n_steps += 1.0;
    // APD90: upstroke = upward crossing of -40 mV; end = downward crossing of
    // V_rest + 0.1 (V_peak - V_rest), both linearly interpolated between steps.
    if (o.apd_t_up < 0.0) {
        o.apd_v_rest = fmin(o.apd_v_rest, V);
        if (V_old < -40.0 && V >= -40.0) {
            o.apd_t_up = t_old + dt * (-40.0 - V_old) / (V - V_old);
            o.apd_v_peak = V;
        }
    } else {
        o.apd_v_peak = fmax(o.apd_v_peak, V);
        const double thr = o.apd_v_rest + 0.1 * (o.apd_v_peak - o.apd_v_rest);
        if (V_old > thr && V <= thr && V < o.apd_v_peak - 10.0) {
            const double tx = t_old + dt * (V_old - thr) / (V_old - V);
            const double apd = tx - o.apd_t_up;
            o.apd90_prev = o.apd90_last;
            o.apd90_last = apd;
            if (o.beats_done >= 1.0) {
                o.alt_sum += fabs(apd - o.apd90_prev);
                o.alt_count += 1.0;
            }
            o.beats_done += 1.0;
            o.apd_t_up = -1.0;
            o.apd_v_rest = V;
        }
    }
    o.v_prev = V;
}

// Derivatives of the step d.step at state d.g (the part of euler_step before
// SSA(dt), GW.tpp:238-247) and, when kslot >= 0, that step's SSA constants
// (GW.tpp:216-221) into d.k[kslot]. Works on register copies of the cell's
// globals and constants; the results are written back to the cell's state once.
__device__ void prepare_step(const DevCell &C, CellDyn &d, const FluxSums &S, double dt, Currents &cur, int kslot)
{
    const gw_params_t &p = C.p;
    const CellConsts &c = C.c;
    gw_globals_t g = d.g;
    gw_globals_t dg;
    const double Istim = protocol_istim(C, d.t, d.step);
    if (d.clamp) g.V = protocol_vclamp(C, d.t);
    const double VF_RT = g.V * c.F_RT;
    const double expm = gw::exp(-VF_RT);
    const int rc = global_derivatives(p, c, g, VF_RT, dt, Istim, S, dg, cur);
    if (rc != GW_OK && d.status == 0) d.status = rc;
    d.Istim = Istim;
    d.g = g;
    d.dg = dg;
    if (kslot >= 0) {
        step_consts(d.k[kslot], p, c, g.V, expm, g.Cai, g.CaNSR, static_cast<uint32_t>(d.step & 0xffff));
        step_consts_f32(d.kf[kslot], d.k[kslot]);
    }
}

// SSA constants of the step after d.step, into d.k[(d.step + 1) & 1]: the
// potential and concentrations that step will start from (globals + dGlobals
// unless frozen; the clamp potential at t + dt under a clamp protocol), formed
// with the same operations prepare_step will apply at that step.
__device__ void lookahead_consts(const DevCell &C, CellDyn &d, double dt, bool frozen)
{
    double V = d.g.V, Cai = d.g.Cai, CaNSR = d.g.CaNSR;
    if (!frozen) {
        if (!d.clamp) V += d.dg.V;
        Cai += d.dg.Cai;
        CaNSR += d.dg.CaNSR;
    }
    if (d.clamp) V = protocol_vclamp(C, d.t + dt);
    const double VF_RT = V * C.c.F_RT;
    const double expm = gw::exp(-VF_RT);
    const int slot = static_cast<int>((d.step + 1) & 1);
    step_consts(d.k[slot], C.p, C.c, V, expm, Cai, CaNSR, static_cast<uint32_t>((d.step + 1) & 0xffff));
    step_consts_f32(d.kf[slot], d.k[slot]);
}

__device__ void sum_block_partials(const Partial *parts, int64_t first, int n, Partial &tot)
{
    fx_zero(tot.jlcc); fx_zero(tot.jxfer); fx_zero(tot.jtr); fx_zero(tot.jrel);
    tot.clch = 0; tot.open = 0; tot.iters = 0; tot.err = 0;
    for (int b = 0; b < n; ++b) {
        const Partial &q = parts[first + b];
        fx_add(tot.jlcc, q.jlcc);
        fx_add(tot.jxfer, q.jxfer);
        fx_add(tot.jtr, q.jtr);
        fx_add(tot.jrel, q.jrel);
        tot.clch += q.clch;
        tot.open += q.open;
        tot.iters += q.iters;
        tot.err |= q.err;
    }
}

__device__ __forceinline__ void partial_zero(Partial &m)
{
    fx_zero(m.jlcc); fx_zero(m.jxfer); fx_zero(m.jtr); fx_zero(m.jrel);
    m.clch = 0; m.open = 0; m.iters = 0; m.err = 0;
}

__device__ void partial_add(Partial &a, const Partial &b)
{
    fx_add(a.jlcc, b.jlcc);
    fx_add(a.jxfer, b.jxfer);
    fx_add(a.jtr, b.jtr);
    fx_add(a.jrel, b.jrel);
    a.clch += b.clch;
    a.open += b.open;
    a.iters += b.iters;
    a.err |= b.err;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m)
{
    return static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(v), m));
}

// Exact sum of a partial over the 64 lanes of the calling wave, left in every
// lane (a butterfly of lane exchanges). Every lane of the wave must call it.
__device__ void partial_wave_sum(Partial &p)
{
    for (int m = kWave / 2; m > 0; m >>= 1) {
        Partial o;
        o.jlcc.lo = shfl_xor_u64(p.jlcc.lo, m);   o.jlcc.hi = shfl_xor_u64(p.jlcc.hi, m);
        o.jxfer.lo = shfl_xor_u64(p.jxfer.lo, m); o.jxfer.hi = shfl_xor_u64(p.jxfer.hi, m);
        o.jtr.lo = shfl_xor_u64(p.jtr.lo, m);     o.jtr.hi = shfl_xor_u64(p.jtr.hi, m);
        o.jrel.lo = shfl_xor_u64(p.jrel.lo, m);   o.jrel.hi = shfl_xor_u64(p.jrel.hi, m);
        o.clch = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(p.clch), m));
        o.open = static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(p.open), m));
        o.iters = shfl_xor_u64(p.iters, m);
        o.err = __shfl_xor(p.err, m);
        partial_add(p, o);
    }
}

// Exact per-block sums of the SSA outputs. Each output's 128-bit fixed-point
// value is split into three 43-bit limbs (the top one signed) that are added
// into LDS with 64-bit atomics; 64 CRUs per block cannot overflow a limb sum,
// and the limbs are recombined with carries into the same 128-bit total.
// The integer outputs (iterations, open RyRs, open ClCh, error flags) are
// added the same way, lane j of each quad adding output j, so that no atomic
// address is shared by the whole wave (which would make the compiler scan
// the active lanes one by one).
struct BlockAcc {
    unsigned long long limb[4][3];   // Jxfer, Jtr, JLCC, Jrel
    unsigned long long ints[4];      // iterations, open RyRs, open ClCh, error flags
};

__device__ __forceinline__ void acc_zero(BlockAcc &A)
{
    const int t = threadIdx.x;
    if (t < 16) (&A.limb[0][0])[t] = 0;
}

__device__ __forceinline__ void acc_add_fx(BlockAcc &A, int q, const Fx &f)
{
    const uint64_t m43 = (1ULL << 43) - 1;
    atomicAdd(&A.limb[q][0], static_cast<unsigned long long>(f.lo & m43));
    atomicAdd(&A.limb[q][1], static_cast<unsigned long long>((f.lo >> 43) | ((f.hi & ((1ULL << 22) - 1)) << 21)));
    atomicAdd(&A.limb[q][2], static_cast<unsigned long long>(static_cast<int64_t>(f.hi) >> 22));
}

__device__ __forceinline__ Fx acc_fx(const BlockAcc &A, int q)
{
    const uint64_t S0 = A.limb[q][0], S1 = A.limb[q][1], S2 = A.limb[q][2];
    Fx f;
    f.lo = S0;
    f.hi = 0;
    fx_add_raw(f, S1 << 43, S1 >> 21);
    fx_add_raw(f, 0, S2 << 22);
    return f;
}

template <int NT>
__device__ void block_reduce(Partial *red, Partial mine)
{
    const int tid = threadIdx.x;
    red[tid] = mine;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) partial_add(red[tid], red[tid + s]);
        __syncthreads();
    }
}

// ===========================================================================
// kernels
// ===========================================================================
// Thread t of an SSA-shaped block handles subunit (t & 3) of CRU (t >> 2).

// Seed the per-CRU engines and replay digests; give padding slots an inert state.
__global__ void k_seed(KArgs a)
{
    const int b = blockIdx.x;
    const int cell = a.block_cell[b];
    const DevCell &C = a.cells[cell];
    const int j = threadIdx.x & 3;
    const int64_t i = (b - C.blk0[kShapeDefault]) * kCruPerBlock + (threadIdx.x >> 2);
    const int64_t slot = C.slot0 + i;
    if (j == 0) {
        uint64_t s[4];
        rng_seed(a.prng, s, stream_key(C.seed, i));
        for (int k = 0; k < 4; ++k) a.rng[slot * 4 + k] = s[k];
        if (a.dig) {
            a.dig[slot * 2] = 0;
            a.dig[slot * 2 + 1] = 0xcbf29ce484222325ULL;
        }
        if (i >= C.n_cru) st_cajsr(a, slot, 0.0);
    }
    if (i >= C.n_cru) {
        st_cass(a, slot * 4 + j, 0.0);
        st_jlcc(a, slot * 4 + j, 0.0);
        a.sub[slot * 4 + j] = sub_set_lcc(0, 1);
    }
}

// Rest-state sampler (src/includes/GW_utils.tpp:5-64) of every CRU.
template <int PRNG>
__global__ void k_sample_rest(KArgs a)
{
    const int b = blockIdx.x;
    const int cell = a.block_cell[b];
    const DevCell &C = a.cells[cell];
    if ((threadIdx.x & 3) != 0) return;
    const int64_t i = (b - C.blk0[kShapeDefault]) * kCruPerBlock + (threadIdx.x >> 2);
    if (i >= C.n_cru) return;
    const int64_t slot = C.slot0 + i;
    double ca[4], jsr;
    uint64_t sub[4];
    sample_rest_cru<PRNG>(C.seed, i, ca, jsr, sub);
    for (int j = 0; j < 4; ++j) {
        st_cass(a, slot * 4 + j, ca[j]);
        st_jlcc(a, slot * 4 + j, 0.0);
        a.sub[slot * 4 + j] = sub[j];
    }
    st_cajsr(a, slot, jsr);
}

// Step-0 preparation of every cell (one block per cell): the flux sums the
// reference's first euler_step reads -- JLCC of the given state, Jxfer and Jtr
// of the constructor's default CRU state and default globals (GW.hpp:104-113),
// ClCh of the given state -- then derivatives and SSA constants.
__global__ void k_init_cells(KArgs a)
{
    __shared__ Partial red[kBlock];
    const int cell = blockIdx.x;
    const DevCell &C = a.cells[cell];
    CellDyn &D = a.dyn[cell];
    Partial mine;
    partial_zero(mine);
    const double X = C.p.rxfer * (1.45370e-4 - 1.45273e-4);
    const double X4 = X + X + X + X;
    const double Y = C.p.rtr * (0.908882 - 0.908408);
    for (int64_t i = threadIdx.x; i < C.n_cru; i += kBlock) {
        const int64_t slot = C.slot0 + i;
        double jl[4];
        for (int j = 0; j < 4; ++j) jl[j] = ld_jlcc(a, slot * 4 + j);
        mine.err |= fx_add(mine.jlcc, jl[0] + jl[1] + jl[2] + jl[3]);
        mine.err |= fx_add(mine.jxfer, X4);
        mine.err |= fx_add(mine.jtr, Y);
        for (int j = 0; j < 4; ++j) {
            const uint64_t w = a.sub[slot * 4 + j];
            mine.clch += sub_cl(w);
            mine.open += ryr_get(w, 2) + ryr_get(w, 3);
        }
    }
    block_reduce<kBlock>(red, mine);
    if (threadIdx.x == 0) {
        const Partial &t = red[0];
        FluxSums S{fx_to_double(t.jlcc), fx_to_double(t.jxfer), fx_to_double(t.jtr), static_cast<double>(t.clch)};
        D.sums[0] = S.JLCC; D.sums[1] = S.Jxfer; D.sums[2] = S.Jtr; D.sums[3] = S.ClCh;
        D.sums[4] = static_cast<double>(t.open); D.sums[5] = 0.0;
        D.t = 0.0;
        D.step = 0;
        D.status = t.err ? GW_E_RANGE : 0;
        D.stop_step = kNoStop;
        Currents cur;
        prepare_step(C, D, S, a.dt, cur, 0);
        lookahead_consts(C, D, a.dt, a.frozen != 0);
        if (D.status != 0) D.stop_step = 0;
        D.obs.v_prev = D.g.V;
    }
}

// Copy a step's LCC rate table into LDS (it is read with a per-lane index;
// the cell's other constants are wave-uniform and are read as scalars).
template <class KT, class TT>
__device__ __forceinline__ void load_lcc_table(const KT &K, TT &T)
{
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&K.lcc);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&T);
    for (int w = threadIdx.x; w < static_cast<int>(sizeof(TT) / 4); w += blockDim.x) dst[w] = src[w];
}


// SSA(dt) of `nsteps` steps (GW_model::SSA, src/includes/GW.tpp:214-234) for
// the 64 CRUs of SSA block b, with the constants of step step0 (nsteps > 1
// only in frozen-globals replay, where they do not change). The block's flux
// partial of the last step goes to partials[step0 & 1][b]; only the quad
// leaders contribute. R is the CRU-local arithmetic (double, or float for the
// FP32-local variant); CaSS and CaJSR are FP64 in both.
template <class R> struct LocalState;
template <> struct LocalState<double> {
    using Consts = StepConsts;
    static __device__ __forceinline__ double *jlcc(const KArgs &a) { return a.JLCC; }
    static __device__ __forceinline__ const SsaParams &params(const DevCell &C) { return C.sp; }
    static __device__ __forceinline__ const StepConsts &consts(const CellDyn &D, int ks) { return D.k[ks]; }
};
template <> struct LocalState<float> {
    using Consts = StepConstsF;
    static __device__ __forceinline__ float *jlcc(const KArgs &a) { return a.JLCCf; }
    static __device__ __forceinline__ const SsaParamsF &params(const DevCell &C) { return C.spf; }
    static __device__ __forceinline__ const StepConstsF &consts(const CellDyn &D, int ks) { return D.kf[ks]; }
};

template <int PRNG, bool DIGEST, class R, int BLOCK>
__device__ __forceinline__ void ssa_block(const KArgs &a, int b, int64_t step0, int64_t nsteps, LccTableT<R> &T, BlockAcc &acc)
{
    constexpr int kShape = shape_of_block(BLOCK);
    using LS = LocalState<R>;
    const int tid = threadIdx.x;
    const int j = tid & 3;
    const int cell = __builtin_amdgcn_readfirstlane(a.block_cell[b]);
    const int ks = static_cast<int>(step0 & 1);
    const DevCell &C = a.cells[cell];
    const CellDyn &D = a.dyn[cell];
    const auto &sp = LS::params(C);
    const auto &kk = LS::consts(D, ks);
    load_lcc_table(kk, T);
    acc_zero(acc);
    __syncthreads();
    const int64_t i = (b - C.blk0[kShape]) * (BLOCK / 4) + (tid >> 2);
    const bool active = (i < C.n_cru) && (step0 < D.stop_step);
    if (active) {
        R *const JLCC = LS::jlcc(a);
        const int64_t slot = C.slot0 + i;
        double ca = a.CaSS[slot * 4 + j];
        R jl = JLCC[slot * 4 + j];
        uint64_t w = a.sub[slot * 4 + j];
        CruQ q;
        q.jsr = a.CaJSR[slot];
        q.r0 = a.rng[slot * 4];
        q.r1 = a.rng[slot * 4 + 1];
        q.r2 = a.rng[slot * 4 + 2];
        q.r3 = a.rng[slot * 4 + 3];
        Digest dg{0, 0};
        if (DIGEST) {
            dg.count = a.dig[slot * 2];
            dg.hash = a.dig[slot * 2 + 1];
        }
        QuadOut o{};
        uint64_t iters = 0;
        int changed = 0;
        const R dt = static_cast<R>(a.dt);
        for (int64_t s = 0; s < nsteps; ++s) {
            o = ssa_step_quad<PRNG, DIGEST>(ca, jl, w, q, sp, kk, T, dt, dg, static_cast<uint32_t>((step0 + s) & 0xffff), j);
            iters += o.iters;
            changed |= o.changed;
            if (o.err) break;
        }
        a.CaSS[slot * 4 + j] = ca;
        if (changed) {   // the channel word and JLCC change only at events and re-splits
            JLCC[slot * 4 + j] = jl;
            a.sub[slot * 4 + j] = w;
        }
        if (j == 0) {
            a.CaJSR[slot] = q.jsr;
            a.rng[slot * 4] = q.r0;
            a.rng[slot * 4 + 1] = q.r1;
            a.rng[slot * 4 + 2] = q.r2;
            a.rng[slot * 4 + 3] = q.r3;
            if (DIGEST) {
                a.dig[slot * 2] = dg.count;
                a.dig[slot * 2 + 1] = dg.hash;
            }
        }
        // lane j adds output j of its CRU (update_CRUstate_from_temp, GW.tpp:47-62)
        const double v = j == 0 ? o.Jxfer : j == 1 ? o.Jtr : j == 2 ? o.JLCC : o.Jrel;
        Fx f;
        fx_zero(f);
        const int fl = fx_add(f, v);
        acc_add_fx(acc, j, f);
        // error flags of the three summed fluxes (the range of Jrel is not checked)
        const int e = (qgeti<0>(fl) | qgeti<1>(fl) | qgeti<2>(fl)) | (o.err ? 4 : 0);
        const int64_t iv = j == 0 ? static_cast<int64_t>(iters) : j == 1 ? o.open : j == 2 ? o.clch : e;
        atomicAdd(&acc.ints[j], static_cast<unsigned long long>(iv));
    }
    __syncthreads();
    if (tid == 0) {
        Partial p;
        p.jxfer = acc_fx(acc, 0);
        p.jtr = acc_fx(acc, 1);
        p.jlcc = acc_fx(acc, 2);
        p.jrel = acc_fx(acc, 3);
        p.iters = acc.ints[0];
        p.open = static_cast<int64_t>(acc.ints[1]);
        p.clch = static_cast<int64_t>(acc.ints[2]);
        p.err = acc.ints[3] ? 1 : 0;
        p.tag = 0;
        a.partials[ks * a.nblocks + b] = p;
    }
}

constexpr int kGd0 = 7;   // V, Nai, Ki, Cai, CaNSR, CaLTRPN, CaHTRPN: the globals gd_finish_main derives

// LDS of a close block. The working set of the close reuses the slots of the
// partial-sum reduction once its total has been read.
struct CloseLds {
    double g[36];            // the globals the derivatives start from (gw_globals_t order)
    double x[GD_N];          // arguments, then values, of the batched transcendentals
    double py[GD_N - GD_POW0];
    double sx[SC_N];         // arguments, then values, of the next step's constants
    double lookahead[3];     // V, Cai, CaNSR of the next step
    double ch[29];           // the derivatives of the gates m, h, j, xKs and of the IKr, Kv4.3 and
                             // Kv1.4 chains (the last 29 globals, gw_globals_t order)
    double dg0[kGd0];        // the derivatives of V and the concentrations (the first 7 globals)
    double Istim;            // the stimulus of the new time
    double ICaL;             // I_CaL of the close, for the observables
    gw_params_t p;           // the cell's parameters, derived constants and observables,
    CellConsts c;            // staged for the serial part (which would otherwise wait on
    gw_cell_obs_t obs;       // them load by load)
};
union CloseShared {
    Partial red[kWave];
    CloseLds L;
};
static_assert(sizeof(gw_globals_t) == 36 * sizeof(double), "globals are 36 doubles");
static_assert(sizeof(CloseLds) <= sizeof(Partial) * kWave, "the close's working set fits the reduction slots");
static_assert(sizeof(gw_params_t) % 8 == 0 && sizeof(CellConsts) % 8 == 0, "staged as 8-byte words");
static_assert(offsetof(gw_globals_t, h) == offsetof(gw_globals_t, m) + sizeof(double) &&
                  offsetof(gw_globals_t, xKs) == offsetof(gw_globals_t, m) + 3 * sizeof(double) &&
                  offsetof(gw_globals_t, Kr) == offsetof(gw_globals_t, m) + 4 * sizeof(double) &&
                  offsetof(gw_globals_t, Kv43) == offsetof(gw_globals_t, Kr) + 5 * sizeof(double) &&
                  offsetof(gw_globals_t, Kv14) == offsetof(gw_globals_t, Kr) + 15 * sizeof(double) &&
                  sizeof(gw_globals_t) == offsetof(gw_globals_t, m) + 29 * sizeof(double) &&
                  offsetof(gw_globals_t, m) == kGd0 * sizeof(double),
              "the gates and the three chains are the last 29 globals: m, h, j, xKs, Kr, Kv4.3, Kv1.4");

// Close step n = D.step of one cell: the exact sum of the cell's block
// partials of step n, globals += dGlobals (GW.tpp:252-271), t += dt, the
// derivatives of step n+1, observables and trace, and the SSA constants of
// step n+2 (those of step n+1 were formed one close earlier and are in use by
// the SSA blocks running beside this one). The step's 42 exponentials,
// logarithms and powers that depend on the state alone are evaluated one per
// lane, and the 5 exponentials of the next step's constants likewise; the
// rest is the serial arithmetic of prepare_step and lookahead_consts. Every
// thread of the block takes part (the block synchronises).
// D is the cell's dynamic state: a.dyn[cell], or in the persistent kernel a
// copy the close block keeps in LDS for the whole launch. With COH the block
// partials are read with device-coherent atomic loads (the persistent
// kernel's SSA blocks write them that way within the launch).
template <int NS, bool COH = false>
__device__ __forceinline__ void close_cell(const KArgs &a, int cell, CellDyn &D, CloseShared &sh)
{
    const DevCell &C = a.cells[cell];
    const int tid = threadIdx.x;
    const int nw = static_cast<int>(blockDim.x) / kWave;
    // the cell's status, step, time and potential, read before the reduction's barriers
    const int status = D.status;
    const bool clamp = D.clamp != 0;
    const int64_t step_new = D.step + 1;
    const int ks = static_cast<int>((step_new - 1) & 1);
    const double t_old = D.t;
    const double V_old = D.g.V;
    const double t_new = t_old + a.dt;
    // the stimulus of the new time, on the last wave (the first leads the
    // serial part below): formed below while that wave's first block partial
    // is loading, or here where the partials are read coherently
    const int istim_tid = (nw - 1) * kWave;
    double Istim = 0.0;
    if constexpr (COH) {
        if (tid == istim_tid && status == 0) Istim = protocol_istim(C, t_new, step_new);
    }
    // the parameters, derived constants and observables, loaded now and stored
    // to LDS once the reduction slots are free
    constexpr int kPW = sizeof(gw_params_t) / 8, kCW = sizeof(CellConsts) / 8, kOW = sizeof(gw_cell_obs_t) / 8;
    constexpr int kStage = (kPW + kCW + kOW + NS - 1) / NS;
    unsigned long long stage[kStage];
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
        const int w = tid + r * NS;
        stage[r] = 0;
        if (tid < NS) {
            if (w < kPW) stage[r] = reinterpret_cast<const unsigned long long *>(&C.p)[w];
            else if (w < kPW + kCW) stage[r] = reinterpret_cast<const unsigned long long *>(&C.c)[w - kPW];
            else if (w < kPW + kCW + kOW)
                stage[r] = reinterpret_cast<const unsigned long long *>(&D.obs)[w - kPW - kCW];
        }
    }
    // this thread's global and its increment, read now for globals_apply below
    double g_in = 0.0, dg_in = 0.0;
    if (tid < 36) {
        g_in = reinterpret_cast<const double *>(&D.g)[tid];
        dg_in = reinterpret_cast<const double *>(&D.dg)[tid];
    }
    // the block partials: read by every wave (by the first alone where they
    // are read coherently, after its wait for them), summed within each wave,
    // then across the waves; the sums are exact, so the order does not matter
    Partial mine;
    partial_zero(mine);
    {
        const Partial *src = a.partials + ks * a.nblocks + C.blk0[a.shape];
        const int nb = C.nblk[a.shape];
        if constexpr (COH) {
            if (tid < NS) {
                for (int q = tid; q < nb; q += NS) {
                    Partial p;
                    unsigned long long *pw = reinterpret_cast<unsigned long long *>(&p);
                    const unsigned long long *sw = reinterpret_cast<const unsigned long long *>(src + q);
                    for (int t = 0; t < static_cast<int>(sizeof(Partial) / 8); ++t)
                        pw[t] = __hip_atomic_load(sw + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    partial_add(mine, p);
                }
            }
        } else {
            // the thread's first partial (the only one where the cell has no
            // more blocks than the close has threads) is in flight while the
            // stimulus is formed; the rest two at a time
            const int nt = static_cast<int>(blockDim.x);
            int q = tid;
            Partial first;
            const bool has_first = q < nb;
            if (has_first) first = src[q];
            if (tid == istim_tid && status == 0) Istim = protocol_istim(C, t_new, step_new);
            if (has_first) partial_add(mine, first);
            for (q += nt; q + nt < nb; q += 2 * nt) {
                const Partial p0 = src[q], p1 = src[q + nt];
                partial_add(mine, p0);
                partial_add(mine, p1);
            }
            if (q < nb) partial_add(mine, src[q]);
        }
    }
    partial_wave_sum(mine);
    if ((tid & (kWave - 1)) == 0) sh.red[tid / kWave] = mine;
    __syncthreads();
    Partial tot = sh.red[0];
    for (int w = 1; w < nw; ++w) partial_add(tot, sh.red[w]);
    __syncthreads();   // the reduction slots are reused below
    if (tid == 0 && a.iters_total) atomicAdd(a.iters_total, static_cast<unsigned long long>(tot.iters));
    if (status != 0) return;
    if (tot.err) {
        if (tid == 0) {
            D.status = GW_E_RANGE;
            D.stop_step = step_new + 1;   // SSA(step_new) runs beside this close; none after it
        }
        return;
    }
    CloseLds &L = sh.L;
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
        const int w = tid + r * NS;
        if (tid < NS) {
            if (w < kPW) reinterpret_cast<unsigned long long *>(&L.p)[w] = stage[r];
            else if (w < kPW + kCW) reinterpret_cast<unsigned long long *>(&L.c)[w - kPW] = stage[r];
            else if (w < kPW + kCW + kOW) reinterpret_cast<unsigned long long *>(&L.obs)[w - kPW - kCW] = stage[r];
        }
    }
    if (nw > 1 && tid == istim_tid) L.Istim = Istim;
    const FluxSums S{fx_to_double(tot.jlcc), fx_to_double(tot.jxfer), fx_to_double(tot.jtr), static_cast<double>(tot.clch)};
    const double jrel = fx_to_double(tot.jrel);
    // globals_apply (GW.tpp:252-271), one global per thread; V is held under a
    // clamp and then set to the clamp potential of the new time (prepare_step)
    if (tid < 36) {
        double v = g_in;
        if (!a.frozen && !(tid == 0 && clamp)) v += dg_in;
        if (tid == 0 && clamp) v = protocol_vclamp(C, t_new);
        L.g[tid] = v;
    }
    __syncthreads();
    const gw_globals_t &g = *reinterpret_cast<const gw_globals_t *>(L.g);
    const double VF_RT = g.V * L.c.F_RT;
    // the arguments of the state-dependent transcendentals: the two halves on
    // two waves when the block has them
    if (tid == 0) {
        gd_args_lo(L.p, g, VF_RT, L.x);
        if (blockDim.x < 2 * kWave) gd_args_hi(L.p, g, L.x, L.py);
    } else if (tid == kWave) {
        gd_args_hi(L.p, g, L.x, L.py);
    }
    __syncthreads();
    if (tid < GD_N) {
        const double v = gd_eval(tid, L.x[tid], tid >= GD_POW0 ? L.py[tid - GD_POW0] : 0.0);
        L.x[tid] = v;   // each thread reads and then overwrites its own slot
    }
    __syncthreads();
    // gd_finish: the currents, concentrations and V on lane 0 of wave 0; the
    // gates and the three K-channel chains on the block's second and third
    // waves when it has them (they read only the parameters, the globals and
    // the evaluated exponentials), on lane 0 otherwise
    if (tid == 0) {
        Currents cur;
        gw_globals_t dg;
        if (nw > 1) Istim = L.Istim;
        const int rc = gd_finish_main(L.p, L.c, g, VF_RT, L.x, a.dt, Istim, S, dg, cur);
        if (rc != GW_OK) D.status = rc;
        L.ICaL = cur.ICaL;
        if (nw == 1) {
            double gates[4];
            gd_gates(L.x, g, a.dt, gates);
            dg.m = gates[0];
            dg.h = gates[1];
            dg.j = gates[2];
            dg.xKs = gates[3];
            gd_chain_kr(L.p, L.x, g, a.dt, dg.Kr);
            gd_chain_kv(L.p, L.x, g, a.dt, 0, dg.Kv14);
            gd_chain_kv(L.p, L.x, g, a.dt, 1, dg.Kv43);
        }
        D.Istim = Istim;
        if (nw == 1) D.dg = dg;
        else for (int k = 0; k < kGd0; ++k) L.dg0[k] = reinterpret_cast<const double *>(&dg)[k];
        // lookahead_consts: the potential and concentrations the step after starts from
        double V = g.V, Cai = g.Cai, CaNSR = g.CaNSR;
        if (!a.frozen) {
            if (!clamp) V += dg.V;
            Cai += dg.Cai;
            CaNSR += dg.CaNSR;
        }
        if (clamp) V = protocol_vclamp(C, t_new + a.dt);
        sc_args(V, L.c.F_RT, L.sx);
        L.lookahead[0] = V;
        L.lookahead[1] = Cai;
        L.lookahead[2] = CaNSR;
    } else if (tid == kWave) {
        gd_chain_kv(L.p, L.x, g, a.dt, 0, L.ch + 19);   // Kv1.4
        gd_gates(L.x, g, a.dt, L.ch);                   // m, h, j, xKs
        if (nw == 2) {
            gd_chain_kv(L.p, L.x, g, a.dt, 1, L.ch + 9);
            gd_chain_kr(L.p, L.x, g, a.dt, L.ch + 4);
        }
    } else if (tid == 2 * kWave) {
        gd_chain_kv(L.p, L.x, g, a.dt, 1, L.ch + 9);    // Kv4.3
        gd_chain_kr(L.p, L.x, g, a.dt, L.ch + 4);       // IKr
    }
    __syncthreads();
    if (nw > 1 && tid < 36) reinterpret_cast<double *>(&D.dg)[tid] = tid < kGd0 ? L.dg0[tid] : L.ch[tid - kGd0];
    if (tid < SC_N) {
        const double v = gw::exp(L.sx[tid]);
        L.sx[tid] = v;
    }
    __syncthreads();
    if (tid < 36) reinterpret_cast<double *>(&D.g)[tid] = L.g[tid];
    // the SSA constants of step_new + 1: the LCC rate table one state per lane
    const int slot = static_cast<int>((step_new + 1) & 1);
    if (tid < 13) {
        const LccRow w = lcc_row(tid, L.p, L.c, L.sx);
        lcc_row_store(D.k[slot].lcc, tid, w);
        lcc_row_store(D.kf[slot].lcc, tid, w);
    }
    // the rest of the constants of step_new + 1, formed in registers and
    // stored in both precisions: on the third wave when the block has one
    if (tid == (nw > 2 ? 2 * kWave : 0)) {
        StepConsts k;
        step_consts_scalars(k, L.c, L.lookahead[0], L.sx, L.lookahead[1], L.lookahead[2],
                            static_cast<uint32_t>((step_new + 1) & 0xffff));
        StepConsts &K = D.k[slot];
        K.Cai = k.Cai;
        K.CaNSR = k.CaNSR;
        K.inact0 = k.inact0;
        K.inact1 = k.inact1;
        K.JLCC_mult = k.JLCC_mult;
        K.JLCC_exp = k.JLCC_exp;
        K.step_lo = k.step_lo;
        K._pad = 0;
        step_consts_f32_scalars(D.kf[slot], k);
    }
    // observables and trace, on the second wave when the block has one
    if (tid == (nw > 1 ? kWave : 0)) {
        const double open_frac = static_cast<double>(tot.open) / (20.0 * C.n_cru);
        gw_cell_obs_t o = L.obs;
        obs_update(o, V_old, g.V, t_old, a.dt, g.Cai, L.ICaL, open_frac, jrel * L.c.CRU_factor);
        D.obs = o;
        if (a.trace_every > 0 && (step_new % a.trace_every) == 0 && D.trace_rows < a.trace_cap) {
            double *row = a.trace + (D.trace_rows * a.ncells + cell) * 4;
            row[0] = g.V;
            row[1] = g.Cai;
            row[2] = L.ICaL;
            row[3] = open_frac;
            D.trace_rows += 1;
        }
    }
    if (tid != 0) return;
    D.sums[0] = S.JLCC; D.sums[1] = S.Jxfer; D.sums[2] = S.Jtr; D.sums[3] = S.ClCh;
    D.sums[4] = static_cast<double>(tot.open); D.sums[5] = jrel;
    D.t = t_new;
    D.step = step_new;
    if (!(g.V == g.V) || !(g.Cai == g.Cai)) {
        if (D.status == 0) D.status = GW_E_NAN;
    }
    // A cell that fails here stops after the SSA of step_new, which runs beside
    // this close in the same launch: the SSA blocks read stop_step, which only
    // changes from a value above step_new to step_new + 1, so every block of
    // that launch runs the step and none runs a later one.
    if (D.status != 0) D.stop_step = step_new + 1;
}

// One launch per Euler step: the first `nclose` blocks close step ssa_step-1
// of every cell while the remaining blocks run SSA(ssa_step). The two halves
// share no data: the SSA reads the constants of its step (formed by the close
// one launch earlier) and writes the other partials buffer. ssa_step < 0
// launches the closes alone (the last close of a sequence).
// Resident waves per SIMD the SSA kernels are compiled for: FP64 needs about
// 128 VGPRs; the FP32-local SSA fits 80 (the single-lane close path then
// spills to scratch, which it can afford).
template <class R> constexpr int kWavesPerSimd = sizeof(R) == 4 ? 6 : 4;

// ssa_step is relative to *base when base is given (replayed graphs of steps).
template <int PRNG, bool DIGEST, class R, int BLOCK>
__global__ __launch_bounds__(BLOCK, kWavesPerSimd<R>) void k_step(KArgs a, int64_t ssa_step, int32_t nclose,
                                                                   const int64_t *base)
{
    __shared__ LccTableT<R> T;
    __shared__ CloseShared csh;
    __shared__ BlockAcc acc;
    const int bid = blockIdx.x;
    if (bid < nclose) {
        close_cell<kWave>(a, bid, a.dyn[bid], csh);
        return;
    }
    ssa_block<PRNG, DIGEST, R, BLOCK>(a, bid - nclose, base ? *base + ssa_step : ssa_step, 1, T, acc);
}

// The SSA of block b with a quad of lanes per CRU and CPT CRUs per quad: wave
// w of the block owns 16 * CPT consecutive CRUs of the block's cell; a quad
// whose CRU has finished the step takes the next CRU not yet taken (the
// finishing quads in quad order, by a ballot of their first lanes). Lane j
// adds output j of each CRU it finishes to the exact block sums, so the
// results do not depend on which quad ran which CRU. No replay digests.
template <int PRNG, class R, int BLOCK, int CPT>
__device__ __forceinline__ void ssa_block_refill(const KArgs &a, int b, int64_t step0, LccTableT<R> &T, BlockAcc &acc)
{
    constexpr int kQuads = kWave / 4;
    constexpr int kMap = shape_of_crus(BLOCK / 4 * CPT);
    static_assert(crus_per_block(kMap) == BLOCK / 4 * CPT, "CRUs per block must be 16 << map");
    using LS = LocalState<R>;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int j = tid & 3;
    const int qd = lane >> 2;
    const int cell = __builtin_amdgcn_readfirstlane(a.block_cell[b]);
    const int ks = static_cast<int>(step0 & 1);
    const DevCell &C = a.cells[cell];
    const CellDyn &D = a.dyn[cell];
    const auto &sp = LS::params(C);
    const auto &kk = LS::consts(D, ks);
    load_lcc_table(kk, T);
    acc_zero(acc);
    __syncthreads();
    const int64_t w0 = static_cast<int64_t>(b - C.blk0[kMap]) * (BLOCK / 4 * CPT) + wave * (kQuads * CPT);
    const int64_t w1 = min(w0 + kQuads * CPT, static_cast<int64_t>(C.n_cru));
    int64_t i = w0 + qd;
    int64_t next = w0 + kQuads;   // first CRU of the range not yet taken (wave-uniform)
    bool have = (i < w1) && (step0 < D.stop_step);
    R *const JLCC = LS::jlcc(a);
    const R dt = static_cast<R>(a.dt);
    const uint32_t step_lo = static_cast<uint32_t>(step0 & 0xffff);
    QuadState<R> S;
    auto load = [&](int64_t slot) {
        S.ca = a.CaSS[slot * 4 + j];
        S.jl = JLCC[slot * 4 + j];
        S.w = a.sub[slot * 4 + j];
        S.q.jsr = a.CaJSR[slot];
        S.q.r0 = a.rng[slot * 4];
        S.q.r1 = a.rng[slot * 4 + 1];
        S.q.r2 = a.rng[slot * 4 + 2];
        S.q.r3 = a.rng[slot * 4 + 3];
        S.dig = Digest{0, 0};
        quad_begin(S, sp);
    };
    if (have) load(C.slot0 + i);
    while (__ballot(have)) {
        bool fin = false;
        if (have) {
            fin = quad_iterate<PRNG, false>(S, sp, kk, T, dt, step_lo, j);
            if (fin) {
                const int64_t slot = C.slot0 + i;
                a.CaSS[slot * 4 + j] = S.ca;
                if (S.changed) {   // the channel word and JLCC change only at events and re-splits
                    JLCC[slot * 4 + j] = S.jl;
                    a.sub[slot * 4 + j] = S.w;
                }
                if (j == 0) {
                    a.CaJSR[slot] = S.q.jsr;
                    a.rng[slot * 4] = S.q.r0;
                    a.rng[slot * 4 + 1] = S.q.r1;
                    a.rng[slot * 4 + 2] = S.q.r2;
                    a.rng[slot * 4 + 3] = S.q.r3;
                }
                const QuadOut o = quad_outputs(S);
                // lane j adds output j of its CRU (update_CRUstate_from_temp, GW.tpp:47-62)
                const double v = j == 0 ? o.Jxfer : j == 1 ? o.Jtr : j == 2 ? o.JLCC : o.Jrel;
                Fx f;
                fx_zero(f);
                const int fl = fx_add(f, v);
                acc_add_fx(acc, j, f);
                const int e = (qgeti<0>(fl) | qgeti<1>(fl) | qgeti<2>(fl)) | (o.err ? 4 : 0);
                const int64_t iv = j == 0 ? static_cast<int64_t>(o.iters) : j == 1 ? o.open : j == 2 ? o.clch : e;
                atomicAdd(&acc.ints[j], static_cast<unsigned long long>(iv));
            }
        }
        const uint64_t m = __ballot(fin && j == 0);
        if (fin) {
            i = next + static_cast<int64_t>(__popcll(m & ((1ULL << (4 * qd)) - 1)));
            have = i < w1;
            if (have) load(C.slot0 + i);
        }
        next += __popcll(m);
    }
    __syncthreads();
    if (tid == 0) {
        Partial p;
        p.jxfer = acc_fx(acc, 0);
        p.jtr = acc_fx(acc, 1);
        p.jlcc = acc_fx(acc, 2);
        p.jrel = acc_fx(acc, 3);
        p.iters = acc.ints[0];
        p.open = static_cast<int64_t>(acc.ints[1]);
        p.clch = static_cast<int64_t>(acc.ints[2]);
        p.err = acc.ints[3] ? 1 : 0;
        p.tag = 0;
        a.partials[ks * a.nblocks + b] = p;
    }
}

// The refill loop keeps a CRU's addresses and the next CRU's index live beside
// the quad state, beyond the 128 VGPRs of 4 waves per SIMD; 3 waves per SIMD
// keep it free of spills.
template <int PRNG, class R, int BLOCK, int CPT>
__global__ __launch_bounds__(BLOCK, 3) void k_step_refill(KArgs a, int64_t ssa_step, int32_t nclose,
                                                           const int64_t *base)
{
    __shared__ LccTableT<R> T;
    __shared__ CloseShared csh;
    __shared__ BlockAcc acc;
    const int bid = blockIdx.x;
    if (bid < nclose) {
        close_cell<kWave>(a, bid, a.dyn[bid], csh);
        return;
    }
    ssa_block_refill<PRNG, R, BLOCK, CPT>(a, bid - nclose, base ? *base + ssa_step : ssa_step, T, acc);
}

// Sum over the 64 lanes of a wave by a fixed xor butterfly (the same value,
// bit for bit, in every lane).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

__device__ __forceinline__ void lane_load(const KArgs &a, int64_t slot, LaneState &S)
{
    const double2 *pc = reinterpret_cast<const double2 *>(a.CaSS + slot * 4);
    const double2 c01 = pc[0], c23 = pc[1];
    S.ca[0] = c01.x; S.ca[1] = c01.y; S.ca[2] = c23.x; S.ca[3] = c23.y;
    const float4 j4 = *reinterpret_cast<const float4 *>(a.JLCCf + slot * 4);
    S.jl[0] = j4.x; S.jl[1] = j4.y; S.jl[2] = j4.z; S.jl[3] = j4.w;
    const ulonglong2 *pw = reinterpret_cast<const ulonglong2 *>(a.sub + slot * 4);
    const ulonglong2 w01 = pw[0], w23 = pw[1];
    S.w[0] = w01.x; S.w[1] = w01.y; S.w[2] = w23.x; S.w[3] = w23.y;
    const ulonglong2 *pr = reinterpret_cast<const ulonglong2 *>(a.rng + slot * 4);
    const ulonglong2 r01 = pr[0], r23 = pr[1];
    S.q.r0 = r01.x; S.q.r1 = r01.y; S.q.r2 = r23.x; S.q.r3 = r23.y;
    S.q.jsr = a.CaJSR[slot];
}

__device__ __forceinline__ void lane_store(const KArgs &a, int64_t slot, const LaneState &S)
{
    double2 *qc = reinterpret_cast<double2 *>(a.CaSS + slot * 4);
    qc[0] = make_double2(S.ca[0], S.ca[1]);
    qc[1] = make_double2(S.ca[2], S.ca[3]);
    if (S.changed) {   // the channel words and JLCC change only at events and re-splits
        *reinterpret_cast<float4 *>(a.JLCCf + slot * 4) = make_float4(S.jl[0], S.jl[1], S.jl[2], S.jl[3]);
        ulonglong2 *qw = reinterpret_cast<ulonglong2 *>(a.sub + slot * 4);
        qw[0] = make_ulonglong2(S.w[0], S.w[1]);
        qw[1] = make_ulonglong2(S.w[2], S.w[3]);
    }
    ulonglong2 *qr = reinterpret_cast<ulonglong2 *>(a.rng + slot * 4);
    qr[0] = make_ulonglong2(S.q.r0, S.q.r1);
    qr[1] = make_ulonglong2(S.q.r2, S.q.r3);
    a.CaJSR[slot] = S.q.jsr;
}

// Running sums of the outputs of the CRUs each lane finishes, in finishing
// order, one LDS slot per thread (kept out of the registers the SSA needs).
template <int BLOCK>
struct LaneSums {
    double Jxfer[BLOCK], Jtr[BLOCK], JLCC[BLOCK], Jrel[BLOCK];
    int open[BLOCK], clch[BLOCK], iters[BLOCK], err[BLOCK];
    __device__ __forceinline__ void zero(int t)
    {
        Jxfer[t] = Jtr[t] = JLCC[t] = Jrel[t] = 0.0;
        open[t] = clch[t] = iters[t] = err[t] = 0;
    }
    __device__ __forceinline__ void add(int t, const LaneOut &o)
    {
        Jxfer[t] += o.Jxfer;
        Jtr[t] += o.Jtr;
        JLCC[t] += o.JLCC;
        Jrel[t] += o.Jrel;
        open[t] += o.open;
        clch[t] += o.clch;
        iters[t] += static_cast<int>(o.iters);
        err[t] |= o.err;
    }
    __device__ __forceinline__ LaneOut get(int t) const
    {
        LaneOut o;
        o.Jxfer = Jxfer[t];
        o.Jtr = Jtr[t];
        o.JLCC = JLCC[t];
        o.Jrel = Jrel[t];
        o.open = open[t];
        o.clch = clch[t];
        o.iters = static_cast<uint32_t>(iters[t]);
        o.err = err[t];
        return o;
    }
};

// Per-wave sums of the lanes' output sums (a fixed FP64 xor butterfly, the
// same in every lane), added exactly into the block accumulators by lanes 0..3.
__device__ __forceinline__ void lane_accumulate(BlockAcc &acc, const LaneOut &o)
{
    const int lane = threadIdx.x & (kWave - 1);
    const double sx = wave_sum(o.Jxfer), st = wave_sum(o.Jtr), sl = wave_sum(o.JLCC), sr = wave_sum(o.Jrel);
    const int n_it = wave_sum(static_cast<int>(o.iters)), n_open = wave_sum(o.open), n_cl = wave_sum(o.clch);
    const int k = lane & 3;
    const double v = k == 0 ? sx : k == 1 ? st : k == 2 ? sl : sr;
    Fx f;
    fx_zero(f);
    const int fl = fx_add(f, v);
    const bool bad = (__ballot(o.err != 0) | __ballot(lane < 3 && fl != 0)) != 0;
    if (lane < 4) {
        acc_add_fx(acc, lane, f);
        const int64_t iv = lane == 0 ? n_it : lane == 1 ? n_open : lane == 2 ? n_cl : (bad ? 1 : 0);
        atomicAdd(&acc.ints[lane], static_cast<unsigned long long>(iv));
    }
}

// The SSA of block b with one lane per CRU (FP32-local variant). Wave w of the
// block owns kWave * CPT consecutive CRUs of the block's cell. Its lanes start
// on the first kWave of them; a lane whose CRU has finished the step takes the
// next CRU of the range not yet taken (the finishing lanes in lane order, by
// a ballot), so the wave makes about (iterations of its range) / kWave trips
// rather than CPT times the most iterations of any one CRU. Which lane runs
// which CRU depends only on the CRUs' own iteration counts, the same on every
// run. Each lane sums the outputs of its CRUs in FP64 as it finishes them; the
// per-wave sums are a fixed xor butterfly, added exactly across waves and
// blocks.
template <int PRNG, int BLOCK, int CPT>
__device__ __forceinline__ void ssa_block_lane(const KArgs &a, int b, int64_t step0, LccTableF &T, BlockAcc &acc,
                                               LaneSums<BLOCK> &Ls)
{
    constexpr int kMap = shape_of_crus(BLOCK * CPT);
    static_assert(crus_per_block(kMap) == BLOCK * CPT, "CRUs per block must be 16 << map");
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int cell = __builtin_amdgcn_readfirstlane(a.block_cell[b]);
    const int ks = static_cast<int>(step0 & 1);
    const DevCell &C = a.cells[cell];
    const CellDyn &D = a.dyn[cell];
    const SsaParamsF &sp = C.spf;
    const StepConstsF &kk = D.kf[ks];
    const float dt = static_cast<float>(a.dt);
    load_lcc_table(kk, T);
    acc_zero(acc);
    Ls.zero(tid);
    __syncthreads();
    const int64_t w0 = static_cast<int64_t>(b - C.blk0[kMap]) * (BLOCK * CPT) + wave * (kWave * CPT);
    const int64_t w1 = min(w0 + kWave * CPT, static_cast<int64_t>(C.n_cru));
    int64_t i = w0 + lane;
    int64_t next = w0 + kWave;   // first CRU of the range not yet taken (wave-uniform)
    bool have = (i < w1) && (step0 < D.stop_step);
    LaneState S;
    if (have) {
        lane_load(a, C.slot0 + i, S);
        lane_begin(S, sp);
    }
    while (__ballot(have)) {
        bool fin = false;
        if (have) {
            fin = lane_iterate<PRNG>(S, sp, kk, T, dt);
            if (fin) {
                lane_store(a, C.slot0 + i, S);
                Ls.add(tid, lane_outputs(S));
            }
        }
        if constexpr (CPT > 1) {
            const uint64_t m = __ballot(fin);
            if (fin) {
                const int rank = static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                  __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
                i = next + rank;
                have = i < w1;
                if (have) {
                    lane_load(a, C.slot0 + i, S);
                    lane_begin(S, sp);
                }
            }
            next += __popcll(m);
        } else {
            if (fin) have = false;
        }
    }
    lane_accumulate(acc, Ls.get(tid));
    __syncthreads();
    if (tid == 0) {
        Partial p;
        p.jxfer = acc_fx(acc, 0);
        p.jtr = acc_fx(acc, 1);
        p.jlcc = acc_fx(acc, 2);
        p.jrel = acc_fx(acc, 3);
        p.iters = acc.ints[0];
        p.open = static_cast<int64_t>(acc.ints[1]);
        p.clch = static_cast<int64_t>(acc.ints[2]);
        p.err = acc.ints[3] ? 1 : 0;
        p.tag = 0;
        a.partials[ks * a.nblocks + b] = p;
    }
}

// One lane per CRU needs about 100-128 VGPRs; 4 waves per SIMD keep it free
// of spills and still hold 256 lanes per SIMD.
template <int PRNG, int BLOCK, int CPT>
__global__ __launch_bounds__(BLOCK, 4) void k_step_lane(KArgs a, int64_t ssa_step, int32_t nclose, const int64_t *base)
{
    __shared__ LccTableF T;
    __shared__ CloseShared csh;
    __shared__ BlockAcc acc;
    __shared__ LaneSums<BLOCK> Ls;
    const int bid = blockIdx.x;
    if (bid < nclose) {
        close_cell<kWave>(a, bid, a.dyn[bid], csh);
        return;
    }
    ssa_block_lane<PRNG, BLOCK, CPT>(a, bid - nclose, base ? *base + ssa_step : ssa_step, T, acc, Ls);
}

// ---------------------------------------------------------------------------
// Persistent stepping of a batch whose SSA blocks can all be resident at once
// (a few cells of up to tens of thousands of CRUs, where a step is short and a
// launch per step leaves the chip waiting on launch latency). One launch runs
// nsteps Euler steps with the cells' CRUs held in registers. The order of the
// work is that of the launch-per-step form: SSA(n) of a cell runs with the
// constants formed by the close of step n - 2, beside the close of step n - 1,
// and the close of step n sums the block partials of SSA(n). Within the
// launch the kernel boundaries are replaced by tags and a counter: each SSA
// block writes its partial and then the partial's tag (the step index + 1);
// the cell's close block waits until every partial of the cell carries the
// tag of the step it closes, and after forming the constants of step n + 2
// it publishes them and raises the cell's `kready` to n + 3, which the SSA
// blocks wait for before step n + 2. (A counter that every SSA block
// incremented would serialise hundreds of same-address atomics per step.)

// Everything that crosses blocks (partials, constants, counters) is written
// and read with device-coherent atomic accesses, and each writer waits for its
// data stores to complete before it moves a counter, so no cache-wide
// writeback or invalidation is issued. The last close runs in the same launch,
// so every step is closed when it returns. The results are bit-identical to
// the launch-per-step form (the same arithmetic, exact flux sums). Every wait
// is bounded: after a.spin_limit polls the waiter sets `timeout`, which every
// other waiter also sees, and all blocks leave.
// ---------------------------------------------------------------------------
constexpr int64_t kSpinLimit = int64_t(1) << 21;   // polls of about 1 us: a wait gives up after seconds

__device__ __forceinline__ unsigned long long ld_coh(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until this thread's stores have completed (the counters move after).
__device__ __forceinline__ void stores_done() { __builtin_amdgcn_s_waitcnt(0); }

__device__ bool wait_at_least(unsigned long long *p, unsigned long long v, int32_t *timeout, int64_t limit)
{
    for (int64_t it = 0;; ++it) {
        if (ld_coh(p) >= v) return true;
        if (__hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if (it >= limit) {
            __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Counter, published constants and partial tags of every cell before a
// persistent launch from the fully closed state at step step0, whose
// constants of steps step0 and step0 + 1 are in the cell's parity slots. One
// block per cell.
template <class R>
__global__ void k_sync_init(KArgs a, int64_t step0)
{
    using KT = typename LocalState<R>::Consts;
    constexpr int kWords = sizeof(KT) / 8;
    const int c = blockIdx.x;
    if (c == 0 && threadIdx.x == 0) *a.timeout = 0;
    if (c >= a.ncells) return;
    CellSync &X = a.sync[c];
    const CellDyn &D = a.dyn[c];
    for (int t = threadIdx.x; t < 2 * kXchgWords; t += blockDim.x) {
        const int par = t / kXchgWords, w = t % kXchgWords;
        unsigned long long v = 0;
        if (w < kWords) v = reinterpret_cast<const unsigned long long *>(&LocalState<R>::consts(D, par))[w];
        else if (w == kXchgWords - 1) v = static_cast<unsigned long long>(D.stop_step);
        X.k[par][w] = v;
    }
    if (threadIdx.x == 0) X.kready = static_cast<unsigned long long>(step0 + 2);
    const DevCell &C = a.cells[c];
    for (int q = threadIdx.x; q < 2 * C.nblk[kShapeDefault]; q += blockDim.x) {
        const int par = q & 1, blk = q >> 1;
        a.partials[par * a.nblocks + C.blk0[kShapeDefault] + blk].tag = 0;
    }
}

template <int PRNG, bool DIGEST, class R>
__device__ __forceinline__ void ssa_block_persist(const KArgs &a, int b, int64_t step0, int64_t nsteps,
                                                  LccTableT<R> &T, BlockAcc &acc,
                                                  typename LocalState<R>::Consts &K, int64_t &stop, int &leave)
{
    using LS = LocalState<R>;
    using KT = typename LS::Consts;
    constexpr int kWords = sizeof(KT) / 8;
    static_assert(sizeof(KT) % 8 == 0 && kWords < kXchgWords, "step constants fit the exchange slot");
    const int tid = threadIdx.x;
    const int j = tid & 3;
    const int cell = __builtin_amdgcn_readfirstlane(a.block_cell[b]);
    const DevCell &C = a.cells[cell];
    CellSync &X = a.sync[cell];
    const auto &sp = LS::params(C);
    const int64_t i = (b - C.blk0[kShapeDefault]) * (kBlock / 4) + (tid >> 2);
    const bool mine = i < C.n_cru;
    const int64_t slot = C.slot0 + (mine ? i : 0);
    R *const JLCC = LS::jlcc(a);
    double ca = 0.0;
    R jl = R(0);
    uint64_t w = 0;
    CruQ q{};
    Digest dg{0, 0};
    if (mine) {
        ca = a.CaSS[slot * 4 + j];
        jl = JLCC[slot * 4 + j];
        w = a.sub[slot * 4 + j];
        q.jsr = a.CaJSR[slot];
        q.r0 = a.rng[slot * 4];
        q.r1 = a.rng[slot * 4 + 1];
        q.r2 = a.rng[slot * 4 + 2];
        q.r3 = a.rng[slot * 4 + 3];
        if (DIGEST) {
            dg.count = a.dig[slot * 2];
            dg.hash = a.dig[slot * 2 + 1];
        }
    }
    int changed = 0;
    const R dt = static_cast<R>(a.dt);
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t n = step0 + k;
        const int ks = static_cast<int>(n & 1);
        if (tid == 0)
            leave = wait_at_least(&X.kready, static_cast<unsigned long long>(n + 1), a.timeout, a.spin_limit) ? 0 : 1;
        __syncthreads();
        if (leave) break;
        // the step's constants and the cell's stop step, published by its close
        for (int t = tid; t < kWords; t += kBlock) reinterpret_cast<unsigned long long *>(&K)[t] = ld_coh(&X.k[ks][t]);
        if (tid == 0) stop = static_cast<int64_t>(ld_coh(&X.k[ks][kXchgWords - 1]));
        acc_zero(acc);
        __syncthreads();
        load_lcc_table(K, T);
        __syncthreads();
        if (mine && n < stop) {
            const QuadOut o = ssa_step_quad<PRNG, DIGEST>(ca, jl, w, q, sp, K, T, dt, dg, static_cast<uint32_t>(n & 0xffff), j);
            changed |= o.changed;
            // lane j adds output j of its CRU (update_CRUstate_from_temp, GW.tpp:47-62)
            const double v = j == 0 ? o.Jxfer : j == 1 ? o.Jtr : j == 2 ? o.JLCC : o.Jrel;
            Fx f;
            fx_zero(f);
            const int fl = fx_add(f, v);
            acc_add_fx(acc, j, f);
            const int e = (qgeti<0>(fl) | qgeti<1>(fl) | qgeti<2>(fl)) | (o.err ? 4 : 0);
            const int64_t iv = j == 0 ? static_cast<int64_t>(o.iters) : j == 1 ? o.open : j == 2 ? o.clch : e;
            atomicAdd(&acc.ints[j], static_cast<unsigned long long>(iv));
        }
        __syncthreads();
        if (tid == 0) {
            Partial p;
            p.jxfer = acc_fx(acc, 0);
            p.jtr = acc_fx(acc, 1);
            p.jlcc = acc_fx(acc, 2);
            p.jrel = acc_fx(acc, 3);
            p.iters = acc.ints[0];
            p.open = static_cast<int64_t>(acc.ints[1]);
            p.clch = static_cast<int64_t>(acc.ints[2]);
            p.err = acc.ints[3] ? 1 : 0;
            p.tag = static_cast<int32_t>(n + 1);
            const unsigned long long *pw = reinterpret_cast<const unsigned long long *>(&p);
            unsigned long long *dst = reinterpret_cast<unsigned long long *>(a.partials + ks * a.nblocks + b);
            constexpr int kLast = static_cast<int>(sizeof(Partial) / 8) - 1;   // the word holding err and tag
            for (int t = 0; t < kLast; ++t) st_coh(dst + t, pw[t]);
            stores_done();
            st_coh(dst + kLast, pw[kLast]);
        }
    }
    if (mine) {
        a.CaSS[slot * 4 + j] = ca;
        if (changed) {
            JLCC[slot * 4 + j] = jl;
            a.sub[slot * 4 + j] = w;
        }
        if (j == 0) {
            a.CaJSR[slot] = q.jsr;
            a.rng[slot * 4] = q.r0;
            a.rng[slot * 4 + 1] = q.r1;
            a.rng[slot * 4 + 2] = q.r2;
            a.rng[slot * 4 + 3] = q.r3;
            if (DIGEST) {
                a.dig[slot * 2] = dg.count;
                a.dig[slot * 2 + 1] = dg.hash;
            }
        }
    }
}

// After the close of step n: publish the constants of step n + 2 it formed
// (the cell's parity slot (n + 2) & 1) and the cell's stop step, then raise
// the cell's counter to n + 3. The counter moves after the other stores have
// completed. A cell that had already stopped publishes its unchanged slot,
// so that every waiter goes on (and finds the stop step). Every thread of the
// block takes part.
template <class R>
__device__ __forceinline__ void publish_consts(CellSync &X, const CellDyn &D, int64_t n)
{
    using KT = typename LocalState<R>::Consts;
    constexpr int kWords = sizeof(KT) / 8;
    const int par = static_cast<int>((n + 2) & 1);
    for (int t = threadIdx.x; t < kWords; t += blockDim.x)
        st_coh(&X.k[par][t], reinterpret_cast<const unsigned long long *>(&LocalState<R>::consts(D, par))[t]);
    if (threadIdx.x == 0) st_coh(&X.k[par][kXchgWords - 1], static_cast<unsigned long long>(D.stop_step));
    stores_done();
    __syncthreads();
    if (threadIdx.x == 0) st_coh(&X.kready, static_cast<unsigned long long>(n + 3));
}

// Wave 0 of a close block waits until every partial of the cell for step n
// carries the tag n + 1 (each lane polls its share, the wave votes). The
// partials of a parity are cleared before the launch (k_sync_init), and
// SSA(n + 2), which writes the same slots, waits for the close of step n.
__device__ bool wait_partials(const KArgs &a, const DevCell &C, int64_t n, int32_t *timeout, int64_t limit)
{
    constexpr int kTagWord = static_cast<int>(sizeof(Partial) / 8) - 1;
    const int nblk = C.nblk[kShapeDefault];
    const unsigned long long *base =
        reinterpret_cast<const unsigned long long *>(a.partials + (n & 1) * a.nblocks + C.blk0[kShapeDefault]);
    const int32_t want = static_cast<int32_t>(n + 1);
    for (int64_t it = 0;; ++it) {
        bool ok = true;
        for (int q = threadIdx.x; q < nblk; q += kWave)
            ok = ok && static_cast<int32_t>(ld_coh(base + q * (sizeof(Partial) / 8) + kTagWord) >> 32) == want;
        if (__all(ok)) return true;
        if (__hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if (it >= limit) {
            if (threadIdx.x == 0) __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// The close block of one cell in the persistent kernel: keeps the cell's
// dynamic state in LDS for the launch (it is its only writer), closes steps
// step0 .. step0 + nsteps - 1 in turn, each once every SSA block of the cell
// has written its partial of that step, and publishes the constants it
// formed.
template <class R>
__device__ __forceinline__ void close_persist(const KArgs &a, int cell, int64_t step0, int64_t nsteps, CellDyn &Dl,
                                              CloseShared &sh, int &leave)
{
    static_assert(sizeof(CellDyn) % 8 == 0, "the dynamic state is whole 8-byte words");
    constexpr int kDynWords = sizeof(CellDyn) / 8;
    const DevCell &C = a.cells[cell];
    CellSync &X = a.sync[cell];
    unsigned long long *g = reinterpret_cast<unsigned long long *>(&a.dyn[cell]);
    unsigned long long *l = reinterpret_cast<unsigned long long *>(&Dl);
    for (int t = threadIdx.x; t < kDynWords; t += blockDim.x) l[t] = g[t];
    __syncthreads();
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t n = step0 + k;
        if (threadIdx.x < kWave) {
            const bool ok = wait_partials(a, C, n, a.timeout, a.spin_limit);
            if (threadIdx.x == 0) leave = ok ? 0 : 1;
        }
        __syncthreads();
        if (leave) break;
        close_cell<kWave, true>(a, cell, Dl, sh);
        __syncthreads();
        publish_consts<R>(X, Dl, n);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kDynWords; t += blockDim.x) g[t] = l[t];
}

// The first ncells blocks close (their first wave does the arithmetic), the
// others are the SSA blocks of the default shape (64 CRUs, a quad of lanes
// per CRU).
template <int PRNG, bool DIGEST, class R>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<R>) void k_persist(KArgs a, int64_t step0, int64_t nsteps)
{
    __shared__ LccTableT<R> T;
    __shared__ CloseShared csh;
    __shared__ BlockAcc acc;
    __shared__ typename LocalState<R>::Consts K;
    __shared__ CellDyn Dl;
    __shared__ int64_t stop;
    __shared__ int leave;
    const int bid = blockIdx.x;
    if (bid < a.ncells) {
        close_persist<R>(a, bid, step0, nsteps, Dl, csh, leave);
        return;
    }
    ssa_block_persist<PRNG, DIGEST, R>(a, bid - a.ncells, step0, nsteps, T, acc, K, stop, leave);
}

// ---------------------------------------------------------------------------
// Two Euler steps per launch, one lane per CRU (FP32-local). A launch runs
// SSA(m) and SSA(m + 1) for every CRU with its state held in registers
// between them, so the state crosses HBM once per two steps. SSA(m) needs the
// constants formed by the close of step m - 2 and SSA(m + 1) those of the
// close of step m - 1; neither close can run before every SSA block of the
// previous launch has written its partial, so both run just before the
// launch, in their own kernel (k_close_cells, one block per cell, compiled
// with the whole register file so that its serial arithmetic does not spill).
// A lane that has finished step m of its CRU goes on to step m + 1 of the
// same CRU (reading that step's constants from its LDS slot), and only then
// stores the CRU and takes the next one of its wave's range. Each lane sums
// the outputs of each step separately in finishing order; the per-wave sums
// are added exactly into one block partial per step. The launch sequence of a
// call (host, launch_steps_pair): SSA(step0), SSA(step0 + 1) with no closes
// before them (the call starts with every step closed); then the closes of
// the two steps before each later pair; an odd last step runs alone after one
// close; a last close kernel closes the rest. A cell that fails in either
// close stops as with one step per launch: SSA(s) runs while s < stop_step,
// which the closes before the launch have set. With one CRU per lane (CPT =
// 1) a lane's sums are its CRU's own outputs, so the results are bit-identical
// to one step per launch; with refilled lanes the order of a lane's FP64 sums
// depends on which CRUs it ran, which can change their rounding.
// ---------------------------------------------------------------------------
template <int BLOCK>
struct PairLds {
    StepConstsF K[2];       // constants of steps m and m + 1
    LccTableF T[2];
    BlockAcc acc[2];
    LaneSums<BLOCK> Ls[2];
};

// The block's exact partial of one step, from its accumulators.
__device__ __forceinline__ void write_partial(const KArgs &a, const BlockAcc &acc, int64_t index)
{
    Partial p;
    p.jxfer = acc_fx(acc, 0);
    p.jtr = acc_fx(acc, 1);
    p.jlcc = acc_fx(acc, 2);
    p.jrel = acc_fx(acc, 3);
    p.iters = acc.ints[0];
    p.open = static_cast<int64_t>(acc.ints[1]);
    p.clch = static_cast<int64_t>(acc.ints[2]);
    p.err = acc.ints[3] ? 1 : 0;
    p.tag = 0;
    a.partials[index] = p;
}

template <int PRNG, int BLOCK, int CPT>
__global__ __launch_bounds__(BLOCK, 4) void k_step_lane2(KArgs a, int64_t m, int32_t nss)
{
    constexpr int kMap = shape_of_crus(BLOCK * CPT);
    static_assert(crus_per_block(kMap) == BLOCK * CPT, "CRUs per block must be 16 << map");
    constexpr int kWords = sizeof(StepConstsF) / 4;
    __shared__ PairLds<BLOCK> P;
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int cell = __builtin_amdgcn_readfirstlane(a.block_cell[b]);
    const DevCell &C = a.cells[cell];
    const CellDyn &D = a.dyn[cell];
    const SsaParamsF &sp = C.spf;
    for (int t = tid; t < 2 * kWords; t += BLOCK) {
        const int s = t / kWords, w = t % kWords;
        if (s < nss) reinterpret_cast<uint32_t *>(&P.K[s])[w] = reinterpret_cast<const uint32_t *>(&D.kf[(m + s) & 1])[w];
    }
    acc_zero(P.acc[0]);
    acc_zero(P.acc[1]);
    P.Ls[0].zero(tid);
    P.Ls[1].zero(tid);
    __syncthreads();
    for (int s = 0; s < nss; ++s) load_lcc_table(P.K[s], P.T[s]);
    __syncthreads();
    const float dt = static_cast<float>(a.dt);
    const int64_t stop = D.stop_step;   // set by the closes before this launch
    const bool run0 = m < stop;
    const bool run1 = nss > 1 && m + 1 < stop;
    const int64_t w0 = static_cast<int64_t>(b - C.blk0[kMap]) * (BLOCK * CPT) + wave * (kWave * CPT);
    const int64_t w1 = min(w0 + kWave * CPT, static_cast<int64_t>(C.n_cru));
    int64_t i = w0 + lane;
    int64_t next = w0 + kWave;   // first CRU of the range not yet taken (wave-uniform)
    bool have = (i < w1) && run0;
    int sidx = 0;                // the step (0: m, 1: m + 1) this lane's CRU is in
    LaneState S;
    if (have) {
        lane_load(a, C.slot0 + i, S);
        lane_begin(S, sp);
    }
    while (__ballot(have)) {
        bool fin = false;
        if (have) {
            fin = lane_iterate<PRNG>(S, sp, P.K[sidx], P.T[sidx], dt);
            if (fin) {
                P.Ls[sidx].add(tid, lane_outputs(S));
                if (sidx == 0 && run1) {   // the same CRU, next step
                    const int changed = S.changed;   // the store after step m + 1 keeps step m's changes
                    sidx = 1;
                    lane_begin(S, sp);
                    S.changed = changed;
                    fin = false;
                } else {
                    sidx = 0;
                    lane_store(a, C.slot0 + i, S);
                }
            }
        }
        if constexpr (CPT > 1) {
            const uint64_t msk = __ballot(fin);
            if (fin) {
                const int rank = static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(msk >> 32),
                                                  __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(msk), 0u)));
                i = next + rank;
                have = i < w1;
                if (have) {
                    lane_load(a, C.slot0 + i, S);
                    lane_begin(S, sp);
                }
            }
            next += __popcll(msk);
        } else {
            if (fin) have = false;
        }
    }
    for (int s = 0; s < nss; ++s) lane_accumulate(P.acc[s], P.Ls[s].get(tid));
    __syncthreads();
    if (tid < nss) write_partial(a, P.acc[tid], ((m + tid) & 1) * a.nblocks + b);
}

// ncl consecutive closes of every cell, one block of two waves per cell (the
// two-steps-per-launch sequence): the second wave forms the K-channel chains
// beside the first. No occupancy target, so the close's serial arithmetic
// keeps its values in registers (up to 256 VGPRs: two waves per SIMD, so the
// blocks of 1,024 cells are resident at once; a third wave per block would
// not be).
__global__ __launch_bounds__(2 * kWave, 1) void k_close_cells(KArgs a, int32_t ncl)
{
    __shared__ CloseShared sh;
    const int cell = blockIdx.x;
    for (int c = 0; c < ncl; ++c) {
        close_cell<kWave>(a, cell, a.dyn[cell], sh);
        __syncthreads();
    }
}

// Graph bookkeeping: the step index a replayed chunk starts at.
__global__ void k_set_base(int64_t *base, int64_t v) { *base = v; }
__global__ void k_advance_base(int64_t *base, int64_t k) { *base += k; }

// Frozen-globals replay: nsteps SSA steps in one launch.
template <int PRNG, bool DIGEST, class R>
__global__ __launch_bounds__(kBlock, kWavesPerSimd<R>) void k_ssa(KArgs a, int64_t step0, int64_t nsteps)
{
    __shared__ LccTableT<R> T;
    __shared__ BlockAcc acc;
    ssa_block<PRNG, DIGEST, R, kBlock>(a, blockIdx.x, step0, nsteps, T, acc);
}

// Close a frozen replay of nsteps: sums of the last step, clock advanced as
// t += dt per step (src/pyGW.hpp:116).
__global__ void k_frozen_close(KArgs a, int64_t step0, int64_t nsteps)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= a.ncells) return;
    const DevCell &C = a.cells[cell];
    CellDyn &D = a.dyn[cell];
    Partial tot;
    sum_block_partials(a.partials + (step0 & 1) * a.nblocks, C.blk0[a.shape], C.nblk[a.shape], tot);
    if (a.iters_total) atomicAdd(a.iters_total, static_cast<unsigned long long>(tot.iters));
    if (tot.err && D.status == 0) D.status = GW_E_RANGE;
    D.sums[0] = fx_to_double(tot.jlcc); D.sums[1] = fx_to_double(tot.jxfer); D.sums[2] = fx_to_double(tot.jtr);
    D.sums[3] = static_cast<double>(tot.clch); D.sums[4] = static_cast<double>(tot.open); D.sums[5] = fx_to_double(tot.jrel);
    for (int64_t s = 0; s < nsteps; ++s) D.t += a.dt;
    D.step += nsteps;
    if (D.status != 0 && D.stop_step > D.step) D.stop_step = D.step;
}

// Packed device state <-> row-major host layout of one cell (GWCRUState, src/pyGW.cpp:78-87).
struct StateView {
    double *CaSS, *CaJSR, *JLCC;
    int32_t *LCC, *LCCi, *RyR, *ClCh;
};

__global__ void k_put_state(KArgs a, int cell, StateView v, int has_jlcc)
{
    const DevCell &C = a.cells[cell];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= C.n_cru) return;
    const int64_t slot = C.slot0 + i;
    for (int j = 0; j < 4; ++j) {
        st_cass(a, slot * 4 + j, v.CaSS[i * 4 + j]);
        st_jlcc(a, slot * 4 + j, has_jlcc ? v.JLCC[i * 4 + j] : 0.0);
        int n6[6];
        for (int k = 0; k < 6; ++k) n6[k] = v.RyR[(i * 4 + j) * 6 + k];
        a.sub[slot * 4 + j] = sub_pack(v.LCC[i * 4 + j], v.LCCi[i * 4 + j], v.ClCh[i * 4 + j], n6);
    }
    st_cajsr(a, slot, v.CaJSR[i]);
}

__global__ void k_get_state(KArgs a, int cell, StateView v)
{
    const DevCell &C = a.cells[cell];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= C.n_cru) return;
    const int64_t slot = C.slot0 + i;
    for (int j = 0; j < 4; ++j) {
        v.CaSS[i * 4 + j] = ld_cass(a, slot * 4 + j);
        v.JLCC[i * 4 + j] = ld_jlcc(a, slot * 4 + j);
        const uint64_t w = a.sub[slot * 4 + j];
        v.LCC[i * 4 + j] = sub_lcc(w);
        v.LCCi[i * 4 + j] = sub_y(w);
        v.ClCh[i * 4 + j] = sub_cl(w);
        for (int k = 0; k < 6; ++k) v.RyR[(i * 4 + j) * 6 + k] = ryr_get(w, k);
    }
    v.CaJSR[i] = ld_cajsr(a, slot);
}

// One cell's CRU state in the reduced record encoding (gw_cru_state_reduced_t).
struct ReducedView {
    float *CaSS, *CaJSR;
    int16_t *RyR;
    int8_t *LCC;
    uint8_t *LCCi, *ClCh;
};

// RyR_state_to_int16 (src/utils.py:8-22) of one subunit word: a lumped pair
// with a negative member counts as its sum in the pair's first member.
__device__ __forceinline__ int16_t ryr_code(uint64_t w)
{
    const int r1 = ryr_get(w, 0), r2 = ryr_get(w, 1), r3 = ryr_get(w, 2), r4 = ryr_get(w, 3);
    const int r5 = ryr_get(w, 4), r6 = ryr_get(w, 5);
    const bool neg34 = r3 < 0 || r4 < 0, neg56 = r5 < 0 || r6 < 0;
    const int s3 = neg34 ? r3 + r4 : r3, s4 = neg34 ? 0 : r4, s5 = neg56 ? r5 + r6 : r5;
    return static_cast<int16_t>(r1 + 6 * r2 + 36 * s3 + 216 * s4 + 1296 * s5);
}

__global__ void k_get_state_reduced(KArgs a, int cell, ReducedView v)
{
    const DevCell &C = a.cells[cell];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= C.n_cru) return;
    const int64_t slot = C.slot0 + i;
    for (int j = 0; j < 4; ++j) {
        v.CaSS[i * 4 + j] = static_cast<float>(ld_cass(a, slot * 4 + j));
        const uint64_t w = a.sub[slot * 4 + j];
        v.LCC[i * 4 + j] = static_cast<int8_t>(sub_lcc(w));
        v.LCCi[i * 4 + j] = static_cast<uint8_t>(sub_y(w));
        v.ClCh[i * 4 + j] = static_cast<uint8_t>(sub_cl(w));
        v.RyR[i * 4 + j] = ryr_code(w);
    }
    v.CaJSR[i] = static_cast<float>(ld_cajsr(a, slot));
}

__global__ void k_reset_obs(CellDyn *dyn, int ncells)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    gw_cell_obs_t &o = dyn[c].obs;
    const double V = dyn[c].g.V;
    o.peak_V = -1e300; o.min_V = 1e300; o.peak_Cai = -1e300; o.peak_ICaL_neg = 1e300;
    o.sum_open_RyR_frac = 0; o.int_Jrel = 0; o.n_steps = 0;
    o.v_prev = V; o.apd_v_rest = V; o.apd_v_peak = V; o.apd_t_up = -1.0;
    o.apd90_last = -1.0; o.apd90_prev = -1.0; o.alt_sum = 0; o.alt_count = 0; o.beats_done = 0;
}

// The post-hoc L-type Ca and Ca-activated Cl currents of src/model.py:414-481
// at each cell's current state (one block per cell): ICaL from the open LCCs
// (state 6 or 12 with the y gate open) at the instantaneous V and CaSS,
//   ICaL = CRU_factor PCaL/CSA 4F (VF/RT)/(e^{2VF/RT} - 1) sum_open (CaSS e^{2VF/RT} - 0.341 Cao) 1e9,
// and Ito2 from the open ClCh,
//   Ito2 = CRU_factor Pto2/CSA F (VF/RT)/(e^{-VF/RT} - 1) n_open (Clcyto e^{-VF/RT} - Clo) 1e9.
// The sum of CaSS over open channels is exact (fixed point), so the result
// does not depend on the launch. At V = 0 the factors take their limits,
// 1/2 and -1, where the reference divides 0/0 (deviation D4). out[2c] = ICaL,
// out[2c + 1] = Ito2 of cell c.
__global__ void k_posthoc_currents(KArgs a, double *out)
{
    __shared__ Partial red[kBlock];
    const int cell = blockIdx.x;
    const DevCell &C = a.cells[cell];
    const CellDyn &D = a.dyn[cell];
    Partial mine;
    partial_zero(mine);
    for (int64_t i = threadIdx.x; i < C.n_cru; i += kBlock) {
        const int64_t slot = C.slot0 + i;
        double s = 0.0;
        for (int j = 0; j < 4; ++j) {
            const uint64_t w = a.sub[slot * 4 + j];
            const int L = sub_lcc(w);
            if ((L == 6 || L == 12) && sub_y(w)) {
                s += ld_cass(a, slot * 4 + j);
                mine.open += 1;
            }
            mine.clch += sub_cl(w);
        }
        mine.err |= fx_add(mine.jlcc, s);
    }
    block_reduce<kBlock>(red, mine);
    if (threadIdx.x != 0) return;
    const Partial &t = red[0];
    const gw_params_t &p = C.p;
    const double V = D.g.V;
    const double x = V * C.c.F_RT;
    const double e2 = gw::exp(2.0 * x), em = gw::exp(-x);
    const double f2 = (V == 0.0) ? 0.5 : x / (e2 - 1.0);
    const double fm = (V == 0.0) ? -1.0 : x / (em - 1.0);
    const double n_open = static_cast<double>(t.open);
    out[2 * cell] = C.c.CRU_factor * (p.PCaL / p.CSA) * (4.0 * kFaraday) * f2 *
                    (fx_to_double(t.jlcc) * e2 - 0.341 * p.Cao * n_open) * 1e9;
    out[2 * cell + 1] = C.c.CRU_factor * (p.Pto2 / p.CSA) * kFaraday * fm *
                        static_cast<double>(t.clch) * (p.Clcyto * em - p.Clo) * 1e9;
}

// Status and trace row count of every cell into out[0..n) and out[n..2n),
// and the persistent kernel's timeout flag into out[2n], so that the host
// reads them with one copy.
__global__ void k_gather_dyn(const CellDyn *dyn, int ncells, const int32_t *timeout, int64_t *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) out[2 * ncells] = *timeout;
    if (c >= ncells) return;
    out[c] = dyn[c].status;
    out[ncells + c] = dyn[c].trace_rows;
}

__global__ void k_zero_trace_rows(CellDyn *dyn, int ncells)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncells) dyn[c].trace_rows = 0;
}

// One cell whose globals, CRU state or stimulus changed after it has been
// stepped: the derivatives and SSA constants of its current step are formed
// again from its state (prepare_step + lookahead_consts, as the constructor
// does at step 0), keeping its clock, step index and observables; no other
// cell is touched. With resum != 0 (a new CRU state) the sums read from the
// CRU state -- JLCC, ClCh, open RyRs -- are formed again from it, while Jxfer
// and Jtr keep the values of the cell's last SSA: in the reference those are
// per-CRU arrays only the SSA writes (GW.tpp:47-62), not part of GWCRUState.
__global__ void k_refresh_cell(KArgs a, int cell, int resum)
{
    __shared__ Partial red[kBlock];
    const DevCell &C = a.cells[cell];
    CellDyn &D = a.dyn[cell];
    if (resum) {
        Partial mine;
        partial_zero(mine);
        for (int64_t i = threadIdx.x; i < C.n_cru; i += kBlock) {
            const int64_t slot = C.slot0 + i;
            double jl[4];
            for (int j = 0; j < 4; ++j) jl[j] = ld_jlcc(a, slot * 4 + j);
            mine.err |= fx_add(mine.jlcc, jl[0] + jl[1] + jl[2] + jl[3]);
            for (int j = 0; j < 4; ++j) {
                const uint64_t w = a.sub[slot * 4 + j];
                mine.clch += sub_cl(w);
                mine.open += ryr_get(w, 2) + ryr_get(w, 3);
            }
        }
        block_reduce<kBlock>(red, mine);
    }
    if (threadIdx.x != 0) return;
    if (resum) {
        D.sums[0] = fx_to_double(red[0].jlcc);
        D.sums[3] = static_cast<double>(red[0].clch);
        D.sums[4] = static_cast<double>(red[0].open);
        if (red[0].err && D.status == 0) D.status = GW_E_RANGE;
    }
    const FluxSums S{D.sums[0], D.sums[1], D.sums[2], D.sums[3]};
    Currents cur;
    prepare_step(C, D, S, a.dt, cur, static_cast<int>(D.step & 1));
    lookahead_consts(C, D, a.dt, a.frozen != 0);
    if (D.status != 0 && D.stop_step > D.step) D.stop_step = D.step;
}
// ===========================================================================
// host side
// ===========================================================================
void set_err(char *err, size_t errlen, const char *fmt, ...)
{
    if (!err || errlen == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

#define HIP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            set_err(err, errlen, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return GW_E_HIP;                                                                        \
        }                                                                                           \
    } while (0)

gw_params_t default_params()
{
    gw_params_t p;
#define GW_SET_DEFAULT_(type, name, dflt) p.name = static_cast<type>(dflt);
    GW_PARAMETER_FIELDS(GW_SET_DEFAULT_)
#undef GW_SET_DEFAULT_
    return p;
}

gw_globals_t default_globals()
{
    gw_globals_t g;
    g.V = -91.382; g.Nai = 10.0; g.Ki = 131.84; g.Cai = 1.45273e-4; g.CaNSR = 0.908882;
    g.CaLTRPN = 8.98282e-3; g.CaHTRPN = 0.137617; g.m = 5.33837e-4; g.h = 0.996345; g.j = 0.997315;
    g.xKs = 2.04171e-4;
    const double Kr[5] = {0.999503, 4.13720e-4, 7.27568e-5, 8.73984e-6, 1.36159e-6};
    const double Kv43[10] = {0.953060, 0.0253906, 2.53848e-4, 1.12796e-6, 1.87950e-9,
                             0.0151370, 0.00517622, 8.96600e-4, 8.17569e-5, 2.24032e-6};
    const double Kv14[10] = {0.722328, 0.101971, 0.00539932, 1.27081e-4, 1.82742e-6,
                             0.152769, 0.00962328, 0.00439043, 0.00195348, 0.00143629};
    for (int i = 0; i < 5; ++i) g.Kr[i] = Kr[i];
    for (int i = 0; i < 10; ++i) { g.Kv43[i] = Kv43[i]; g.Kv14[i] = Kv14[i]; }
    return g;
}

}  // namespace

// Device memory of the allocation-free entry point (gw_run_noalloc): one
// static array, which every GPU the library runs on holds until the process
// ends. Host code never passes the array's name to a kernel or a copy: it
// uses the address hipGetSymbolAddress gives on each device (gw_ctx_create).
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

struct gw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void *scratch = nullptr;          // this device's address of `scratch`
    std::vector<gw_batch *> batches;  // live batches; released by gw_ctx_destroy
};

struct gw_batch {
    gw_ctx *ctx = nullptr;
    int32_t ncells = 0, precision = 0, prng = 0;
    double dt = 1e-3;
    int64_t ntot = 0, nblocks = 0;
    bool digest = true;
    bool frozen = false;
    bool prepared = false;
    int64_t steps_done = 0;
    int32_t trace_every = 0;
    int64_t trace_cap = 0;
    int32_t shape = kShapeDefault;   // block -> cell map in use (16 << shape CRUs per block)
    int32_t persist_mode = 0;        // persistent step kernel: 0 off (default), 1 on, 2 when every block fits at once
    int32_t persist_fit = -1;        // whether every block of the persistent kernel fits at once (-1: not known)
    int32_t steps_per_launch = 0;    // SSA steps per plain launch (FP32-local one-lane shapes of 64 threads): 1, 2, 0 automatic
    int64_t pair_capacity[GW_LAUNCH_CONFIGS] = {};   // blocks of k_step_lane2 resident at once, per configuration (0: not known)
    int32_t cfg = 2;                 // launch configuration (kConfigs), 4 x 256 by default
    // replayed graphs of step chunks (0 = plain launches)
    int32_t graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int32_t graph_steps_built = 0, graph_cfg = -1;
    bool graph_digest = false;
    KArgs graph_args{};
    int64_t *d_step_base = nullptr;
    int64_t nblocks_shape[kNumShapes] = {};
    int64_t bc_off[kNumShapes] = {};   // offset of each map's block -> cell table
    int64_t max_nblocks = 0;
    std::vector<DevCell> cells_h;
    std::vector<int32_t> block_cell_h;
    std::vector<gw_globals_t> g0;
    std::vector<double *> istim_dev;
    KArgs a{};
    DevCell *d_cells = nullptr;
    CellDyn *d_dyn = nullptr;
    int32_t *d_block_cell = nullptr;
    void *d_staging = nullptr;
    size_t staging_bytes = 0;
    int64_t *d_gather = nullptr;   // [2][ncells]: status, trace rows (k_gather_dyn)
    // allocation-free batches carve their device memory from the scratch array
    // instead of allocating it, and free nothing
    char *arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
    double *arena_table = nullptr;   // the stimulus table of cell 0
    int64_t arena_table_cap = 0;
};

namespace {

// Device memory of a batch: hipMalloc, or for an allocation-free batch the
// next 256-byte aligned piece of its arena.
int carve(gw_batch *B, void **p, size_t bytes, char *err, size_t errlen)
{
    const size_t at = (B->arena_used + 255) & ~size_t(255);
    if (at + bytes > B->arena_size) {
        set_err(err, errlen, "the allocation-free path holds at most %zu bytes of device state (%zu MiB of static "
                "scratch per GPU); this run needs more: use fewer CRUs or fewer steps, or gw_run",
                B->arena_size, B->arena_size >> 20);
        return GW_E_INVALID_ARG;
    }
    *p = B->arena + at;
    B->arena_used = at + bytes;
    return GW_OK;
}

template <class T>
int dmalloc(gw_batch *B, T **p, size_t n, char *err, size_t errlen)
{
    *p = nullptr;
    if (n == 0) return GW_OK;
    if (B->arena) return carve(B, reinterpret_cast<void **>(p), n * sizeof(T), err, errlen);
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
    return GW_OK;
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_i32(int32_t *p, int64_t n)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int batch_create(gw_ctx *ctx, int32_t ncells, const gw_cell_desc_t *cells, int32_t precision, int32_t prng_id,
                 double dt, const gw_globals_t *g0, void *arena, size_t arena_size, int64_t table_cap,
                 gw_batch **out, char *err, size_t errlen);

int validate_params(const gw_params_t &p, char *err, size_t errlen)
{
    if (!(p.T > 0 && p.CSA > 0 && p.Vcyto > 0 && p.VNSR > 0 && p.VJSR > 0 && p.VSS > 0 && p.NCaRU > 0 && p.Ko > 0 &&
          p.Nao > 0 && p.Cao > 0 && p.b > 0)) {
        set_err(err, errlen, "parameters: temperatures, volumes, NCaRU, b and external concentrations must be > 0");
        return GW_E_INVALID_ARG;
    }
    return GW_OK;
}

int launch_cfg_check(gw_batch *B, char *err, size_t errlen)
{
    if (B->nblocks <= 0 || B->nblocks > (1LL << 31) - 1) {
        set_err(err, errlen, "invalid number of SSA blocks %lld", static_cast<long long>(B->nblocks));
        return GW_E_INVALID_ARG;
    }
    return GW_OK;
}

// The launch configurations: lanes per CRU (4: a quad per CRU, FP64 or
// FP32-local; 1: one lane per CRU, FP32-local), threads per block, and CRUs
// per lane (lane kernels only: each wave owns 64 x that many CRUs and refills
// its lanes as CRUs finish their step).
struct LaunchCfg {
    int lanes, threads, cpl;
};
constexpr int kNumConfigs = GW_LAUNCH_CONFIGS;
constexpr int kConfigDefault = 2;
constexpr LaunchCfg kConfigs[kNumConfigs] = {
    {4, 64, 1},  {4, 128, 1}, {4, 256, 1}, {4, 512, 1},                 // quads: 16..128 CRUs per block
    {1, 64, 1},  {1, 128, 1}, {1, 256, 1},                               // one lane per CRU
    {1, 64, 2},  {1, 128, 2}, {1, 256, 2},                               // two CRUs per lane, refilled
    {1, 64, 4},  {1, 128, 4}, {1, 256, 4},                               // four
    {1, 64, 8},  {1, 128, 8}, {1, 256, 8},                               // eight (up to 2048 CRUs per block)
    {4, 64, 2},  {4, 64, 4},  {4, 128, 4}, {4, 256, 4}};                 // quads refilled from 2 or 4 CRUs each
constexpr int config_shape(int c)
{
    return shape_of_crus(kConfigs[c].threads / kConfigs[c].lanes * kConfigs[c].cpl);
}
static_assert(config_shape(kConfigDefault) == kShapeDefault, "the default configuration uses the setup map");
static_assert(config_shape(kNumConfigs - 1) < kNumShapes, "block -> cell maps cover every configuration");

void set_shape(gw_batch *B, int c)
{
    B->cfg = c;
    B->shape = config_shape(c);
    B->nblocks = B->nblocks_shape[B->shape];
    B->a.shape = B->shape;
    B->a.nblocks = B->nblocks;
    B->a.block_cell = B->d_block_cell + B->bc_off[B->shape];
}

// Configurations other than the default are built for the fast path only:
// FP64 without replay digests or FP32-local, xoshiro256**, free-running
// globals; one lane per CRU is FP32-local only.
bool shape_allowed(const gw_batch *B, int c)
{
    const bool fast = B->prng == GW_PRNG_XOSHIRO256SS && !B->frozen;
    if (kConfigs[c].lanes == 1) return fast && B->precision == GW_PREC_FP32_LOCAL;
    if (kConfigs[c].cpl > 1) return fast && B->precision == GW_PREC_FP64 && !B->digest;   // refilled quads
    return c == kConfigDefault || (fast && (B->precision == GW_PREC_FP32_LOCAL || !B->digest));
}

constexpr int kAutotuneGraphSteps = 64;   // steps per captured graph the autotuner tries
constexpr int kAutotuneWarm = kAutotuneGraphSteps + 2;   // untimed steps before every trial

int prepare(gw_batch *B, char *err, size_t errlen)
{
    if (B->prepared) return GW_OK;
    hipLaunchKernelGGL(k_init_cells, dim3(B->ncells), dim3(kBlock), 0, B->ctx->stream, B->a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reset_obs, dim3((B->ncells + 255) / 256), dim3(256), 0, B->ctx->stream, B->d_dyn, B->ncells);
    HIP_TRY(hipGetLastError());
    B->prepared = true;
    B->steps_done = 0;
    return GW_OK;
}

template <int PRNG>
void launch_frozen(gw_batch *B, int64_t step0, int64_t n)
{
    if (B->precision == GW_PREC_FP32_LOCAL)
        hipLaunchKernelGGL((k_ssa<PRNG, false, float>), dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a, step0, n);
    else if (B->digest)
        hipLaunchKernelGGL((k_ssa<PRNG, true, double>), dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a, step0, n);
    else
        hipLaunchKernelGGL((k_ssa<PRNG, false, double>), dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a, step0, n);
    hipLaunchKernelGGL(k_frozen_close, dim3((B->ncells + 63) / 64), dim3(64), 0, B->ctx->stream, B->a, step0, n);
}

void launch_frozen_steps(gw_batch *B, int64_t step0, int64_t n)
{
    switch (B->prng) {
    case 0: launch_frozen<0>(B, step0, n); break;
    case 1: launch_frozen<1>(B, step0, n); break;
    case 2: launch_frozen<2>(B, step0, n); break;
    case 3: launch_frozen<3>(B, step0, n); break;
    case 4: launch_frozen<4>(B, step0, n); break;
    default: launch_frozen<5>(B, step0, n); break;
    }
}

// Grid of one launch: the closes, then one block per SSA tile.
unsigned step_grid(const gw_batch *B, int64_t ssa_step, int32_t nclose)
{
    return static_cast<unsigned>(nclose + (ssa_step >= 0 ? B->nblocks : 0));
}

// A step launch: with timing events through the dispatch packet, without them
// as a plain launch (the form stream capture records into a graph).
template <class K>
void dispatch(K kernel, unsigned grid, int threads, gw_batch *B, hipEvent_t e0, hipEvent_t e1, int64_t ssa_step,
              int32_t nclose, const int64_t *base)
{
    if (e0)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, B->ctx->stream, e0, e1, 0u, B->a, ssa_step, nclose, base);
    else
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, B->ctx->stream, B->a, ssa_step, nclose, base);
}

// One launch of k_step: SSA(ssa_step) (none if ssa_step < 0) beside the
// closes of the previous step (when nclose > 0). When e0/e1 are given, the
// dispatch itself records them (its start and end, no extra queue packets).
template <int PRNG, int BLOCK>
void launch_k_step_shape(gw_batch *B, int64_t ssa_step, int32_t nclose, hipEvent_t e0, hipEvent_t e1, const int64_t *base)
{
    const unsigned grid = step_grid(B, ssa_step, nclose);
    if (B->precision == GW_PREC_FP32_LOCAL) {
        dispatch(k_step<PRNG, false, float, BLOCK>, grid, BLOCK, B, e0, e1, ssa_step, nclose, base);
        return;
    }
    if constexpr (BLOCK == kBlock) {
        if (B->digest) {
            dispatch(k_step<PRNG, true, double, BLOCK>, grid, BLOCK, B, e0, e1, ssa_step, nclose, base);
            return;
        }
    }
    dispatch(k_step<PRNG, false, double, BLOCK>, grid, BLOCK, B, e0, e1, ssa_step, nclose, base);
}

// Shapes other than the default are built for xoshiro256** only (the engine
// of the production path); shape_allowed() keeps the others off them.
template <int PRNG>
void launch_k_step(gw_batch *B, int64_t ssa_step, int32_t nclose, hipEvent_t e0, hipEvent_t e1, const int64_t *base)
{
    if constexpr (PRNG == GW_PRNG_XOSHIRO256SS) {
        if (kConfigs[B->cfg].lanes == 1) {
            const unsigned grid = step_grid(B, ssa_step, nclose);
            switch (B->cfg) {
            case 4: dispatch(k_step_lane<PRNG, 64, 1>, grid, 64, B, e0, e1, ssa_step, nclose, base); break;
            case 5: dispatch(k_step_lane<PRNG, 128, 1>, grid, 128, B, e0, e1, ssa_step, nclose, base); break;
            case 6: dispatch(k_step_lane<PRNG, 256, 1>, grid, 256, B, e0, e1, ssa_step, nclose, base); break;
            case 7: dispatch(k_step_lane<PRNG, 64, 2>, grid, 64, B, e0, e1, ssa_step, nclose, base); break;
            case 8: dispatch(k_step_lane<PRNG, 128, 2>, grid, 128, B, e0, e1, ssa_step, nclose, base); break;
            case 9: dispatch(k_step_lane<PRNG, 256, 2>, grid, 256, B, e0, e1, ssa_step, nclose, base); break;
            case 10: dispatch(k_step_lane<PRNG, 64, 4>, grid, 64, B, e0, e1, ssa_step, nclose, base); break;
            case 11: dispatch(k_step_lane<PRNG, 128, 4>, grid, 128, B, e0, e1, ssa_step, nclose, base); break;
            case 12: dispatch(k_step_lane<PRNG, 256, 4>, grid, 256, B, e0, e1, ssa_step, nclose, base); break;
            case 13: dispatch(k_step_lane<PRNG, 64, 8>, grid, 64, B, e0, e1, ssa_step, nclose, base); break;
            case 14: dispatch(k_step_lane<PRNG, 128, 8>, grid, 128, B, e0, e1, ssa_step, nclose, base); break;
            default: dispatch(k_step_lane<PRNG, 256, 8>, grid, 256, B, e0, e1, ssa_step, nclose, base); break;
            }
            return;
        }
        if (kConfigs[B->cfg].cpl > 1) {
            const unsigned grid = step_grid(B, ssa_step, nclose);
            switch (B->cfg) {   // FP64 only (shape_allowed)
            case 16: dispatch(k_step_refill<PRNG, double, 64, 2>, grid, 64, B, e0, e1, ssa_step, nclose, base); break;
            case 17: dispatch(k_step_refill<PRNG, double, 64, 4>, grid, 64, B, e0, e1, ssa_step, nclose, base); break;
            case 18: dispatch(k_step_refill<PRNG, double, 128, 4>, grid, 128, B, e0, e1, ssa_step, nclose, base); break;
            default: dispatch(k_step_refill<PRNG, double, 256, 4>, grid, 256, B, e0, e1, ssa_step, nclose, base); break;
            }
            return;
        }
        switch (B->shape) {
        case 0: launch_k_step_shape<PRNG, 64>(B, ssa_step, nclose, e0, e1, base); return;
        case 1: launch_k_step_shape<PRNG, 128>(B, ssa_step, nclose, e0, e1, base); return;
        case 3: launch_k_step_shape<PRNG, 512>(B, ssa_step, nclose, e0, e1, base); return;
        default: break;
        }
    }
    launch_k_step_shape<PRNG, kBlock>(B, ssa_step, nclose, e0, e1, base);
}

void launch_one(gw_batch *B, int64_t ssa_step, int32_t nclose, hipEvent_t e0, hipEvent_t e1,
                const int64_t *base = nullptr)
{
    switch (B->prng) {
    case 0: launch_k_step<0>(B, ssa_step, nclose, e0, e1, base); break;
    case 1: launch_k_step<1>(B, ssa_step, nclose, e0, e1, base); break;
    case 2: launch_k_step<2>(B, ssa_step, nclose, e0, e1, base); break;
    case 3: launch_k_step<3>(B, ssa_step, nclose, e0, e1, base); break;
    case 4: launch_k_step<4>(B, ssa_step, nclose, e0, e1, base); break;
    default: launch_k_step<5>(B, ssa_step, nclose, e0, e1, base); break;
    }
}

// The replayable graph of graph_steps step launches (each SSA(base + k) beside
// the close of the step before) and the launch that advances base. It is
// captured again whenever the launch arguments or shape change.
bool graph_current(const gw_batch *B)
{
    return B->graph_exec && B->graph_steps_built == B->graph_steps && B->graph_cfg == B->cfg &&
           B->graph_digest == B->digest &&
           std::memcmp(&B->graph_args, &B->a, sizeof(KArgs)) == 0;
}

void graph_release(gw_batch *B)
{
    if (B->graph_exec) (void)hipGraphExecDestroy(B->graph_exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->graph_exec = nullptr;
    B->graph = nullptr;
}

hipError_t graph_build(gw_batch *B)
{
    graph_release(B);
    hipStream_t st = B->ctx->stream;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    for (int k = 0; k < B->graph_steps; ++k) launch_one(B, k, B->ncells, nullptr, nullptr, B->d_step_base);
    hipLaunchKernelGGL(k_advance_base, dim3(1), dim3(1), 0, st, B->d_step_base, static_cast<int64_t>(B->graph_steps));
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) return e;
    B->graph = g;
    e = hipGraphInstantiate(&B->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        graph_release(B);
        return e;
    }
    B->graph_steps_built = B->graph_steps;
    B->graph_cfg = B->cfg;
    B->graph_digest = B->digest;
    B->graph_args = B->a;
    return hipSuccess;
}

// The persistent step kernel of a batch's engine and precision.
template <int PRNG>
const void *persist_fn_prng(const gw_batch *B)
{
    if (B->precision == GW_PREC_FP32_LOCAL) return reinterpret_cast<const void *>(&k_persist<PRNG, false, float>);
    if (B->digest) return reinterpret_cast<const void *>(&k_persist<PRNG, true, double>);
    return reinterpret_cast<const void *>(&k_persist<PRNG, false, double>);
}

const void *persist_fn(const gw_batch *B)
{
    switch (B->prng) {
    case 0: return persist_fn_prng<0>(B);
    case 1: return persist_fn_prng<1>(B);
    case 2: return persist_fn_prng<2>(B);
    case 3: return persist_fn_prng<3>(B);
    case 4: return persist_fn_prng<4>(B);
    default: return persist_fn_prng<5>(B);
    }
}

// Whether every block of the persistent kernel (the closes and the SSA blocks
// of the default shape) can be resident at once on the batch's device.
bool persist_fits(gw_batch *B)
{
    if (B->persist_fit >= 0) return B->persist_fit != 0;
    int per_cu = 0, cus = 0;
    const int64_t grid = B->ncells + B->nblocks_shape[kShapeDefault];
    B->persist_fit = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_fn(B), kBlock, 0) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->ctx->device) == hipSuccess)
        B->persist_fit = static_cast<int64_t>(per_cu) * cus >= grid ? 1 : 0;
    return B->persist_fit != 0;
}

// The persistent kernel runs the default shape's free-running steps.
bool use_persist(gw_batch *B)
{
    return B->persist_mode != 0 && !B->frozen && B->cfg == kConfigDefault && persist_fits(B);
}

// Steps step0 .. step0+n-1 in one cooperative launch of the persistent kernel
// (which closes every step itself), after its counters are set for step0.
hipError_t launch_persist(gw_batch *B, int64_t step0, int64_t n)
{
    if (B->precision == GW_PREC_FP32_LOCAL)
        hipLaunchKernelGGL(k_sync_init<float>, dim3(B->ncells), dim3(64), 0, B->ctx->stream, B->a, step0);
    else
        hipLaunchKernelGGL(k_sync_init<double>, dim3(B->ncells), dim3(64), 0, B->ctx->stream, B->a, step0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t grid = B->ncells + B->nblocks_shape[kShapeDefault];
    void *args[] = {&B->a, &step0, &n};
    e = hipLaunchCooperativeKernel(persist_fn(B), dim3(static_cast<unsigned>(grid)), dim3(kBlock), args, 0,
                                   B->ctx->stream);
    return e;
}

// Whether the batch's plain launches run two SSA steps each (k_step_lane2):
// asked for, FP32-local free-running steps on a one-lane shape of 64 threads
// (xoshiro256**, no digests, as every one-lane shape), not persistent, no
// replayed graphs.
bool pair_capable(gw_batch *B)
{
    if (B->frozen || B->graph_steps > 0 || use_persist(B)) return false;
    const LaunchCfg &L = kConfigs[B->cfg];
    return B->precision == GW_PREC_FP32_LOCAL && L.lanes == 1 && L.threads == 64 && B->prng == GW_PRNG_XOSHIRO256SS;
}

// How many blocks of the batch's two-step kernel the GPU holds at once.
int64_t pair_capacity(gw_batch *B)
{
    int64_t &cap = B->pair_capacity[B->cfg];
    if (cap > 0) return cap;
    const void *fn = nullptr;
    switch (kConfigs[B->cfg].cpl) {
    case 1: fn = reinterpret_cast<const void *>(&k_step_lane2<GW_PRNG_XOSHIRO256SS, 64, 1>); break;
    case 2: fn = reinterpret_cast<const void *>(&k_step_lane2<GW_PRNG_XOSHIRO256SS, 64, 2>); break;
    case 4: fn = reinterpret_cast<const void *>(&k_step_lane2<GW_PRNG_XOSHIRO256SS, 64, 4>); break;
    default: fn = reinterpret_cast<const void *>(&k_step_lane2<GW_PRNG_XOSHIRO256SS, 64, 8>); break;
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->ctx->device) == hipSuccess)
        cap = static_cast<int64_t>(per_cu) * cus;
    return cap > 0 ? cap : INT64_MAX;
}

// Two steps per launch when asked for (2), or automatically (0) when the SSA
// blocks of a launch are more than the GPU holds at once: then the state's
// HBM traffic binds and halving it pays for the separate close kernel; a batch
// whose blocks all fit at once is bound by latency, which that kernel adds to.
bool use_pair(gw_batch *B)
{
    if (B->steps_per_launch == 1 || !pair_capable(B)) return false;
    return B->steps_per_launch == 2 || B->nblocks > pair_capacity(B);
}

// One launch of k_step_lane2: SSA(m .. m + nss - 1); with e0/e1 the dispatch
// records its own start and end.
template <class K>
void dispatch_lane2(K kernel, gw_batch *B, int64_t m, int32_t nss, hipEvent_t e0, hipEvent_t e1)
{
    const dim3 grid(static_cast<unsigned>(B->nblocks)), block(64);
    if (e0)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, B->ctx->stream, e0, e1, 0u, B->a, m, nss);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, B->ctx->stream, B->a, m, nss);
}

void launch_lane2(gw_batch *B, int64_t m, int32_t nss, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    constexpr int P = GW_PRNG_XOSHIRO256SS;
    switch (kConfigs[B->cfg].cpl) {
    case 1: dispatch_lane2(k_step_lane2<P, 64, 1>, B, m, nss, e0, e1); break;
    case 2: dispatch_lane2(k_step_lane2<P, 64, 2>, B, m, nss, e0, e1); break;
    case 4: dispatch_lane2(k_step_lane2<P, 64, 4>, B, m, nss, e0, e1); break;
    default: dispatch_lane2(k_step_lane2<P, 64, 8>, B, m, nss, e0, e1); break;
    }
}

void launch_closes(gw_batch *B, int32_t ncl)
{
    hipLaunchKernelGGL(k_close_cells, dim3(B->ncells), dim3(2 * kWave), 0, B->ctx->stream, B->a, ncl);
}

// Steps step0 .. step0+n-1 two per launch: the first launch runs SSA(step0)
// and SSA(step0+1) with no closes before it (the call starts with every step
// closed); each later launch is preceded by the closes its SSA steps need (up
// to two), an odd last step runs alone, and a last close kernel closes the
// rest. Returns the number of SSA launches through *launches. With ssa_ms,
// every SSA dispatch records its own start and end through a ring of event
// pairs (the host waits for the dispatch that last used a pair before reusing
// it, kRing launches back, so the queue never runs dry), and *ssa_ms is the
// sum of their durations: the SSA kernel's time without the close kernels.
hipError_t launch_steps_pair(gw_batch *B, int64_t step0, int64_t n, int64_t *launches, double *ssa_ms = nullptr)
{
    constexpr int kRing = 256;
    std::vector<hipEvent_t> ev(ssa_ms ? 2 * kRing : 0, nullptr);
    hipError_t e = hipSuccess;
    for (auto &x : ev)
        if ((e = hipEventCreate(&x)) != hipSuccess) break;
    double acc = 0.0;
    auto harvest = [&](int r) {
        float f = 0.f;
        hipError_t h = hipEventSynchronize(ev[2 * r + 1]);
        if (h == hipSuccess) h = hipEventElapsedTime(&f, ev[2 * r], ev[2 * r + 1]);
        acc += f;
        return h;
    };
    int64_t m = step0, c = step0, count = 0;   // next SSA step, first step not closed
    const int64_t end = step0 + n;
    while (e == hipSuccess && m < end) {
        const int32_t nss = end - m >= 2 ? 2 : 1;
        const int32_t ncl = static_cast<int32_t>(std::max<int64_t>(0, m + nss - 2 - c));   // closes up to m+nss-3
        if (ncl > 0) launch_closes(B, ncl);
        if (ssa_ms) {
            const int r = static_cast<int>(count % kRing);
            if (count >= kRing && (e = harvest(r)) != hipSuccess) break;
            launch_lane2(B, m, nss, ev[2 * r], ev[2 * r + 1]);
        } else {
            launch_lane2(B, m, nss);
        }
        c += ncl;
        m += nss;
        ++count;
    }
    if (e == hipSuccess && c < end) launch_closes(B, static_cast<int32_t>(end - c));
    for (int64_t k = std::max<int64_t>(0, count - kRing); ssa_ms && e == hipSuccess && k < count; ++k)
        e = harvest(static_cast<int>(k % kRing));
    for (auto &x : ev)
        if (x) (void)hipEventDestroy(x);
    if (ssa_ms) *ssa_ms = acc;
    if (launches) *launches = count;
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// Steps step0 .. step0+n-1: n launches, launch k running SSA(step0+k) beside
// close(step0+k-1), then one launch with the last closes. Between calls every
// step is closed, so the cell state read by the host is that of step0+n.
// With graphs enabled, the middle launches are replays of the captured chunk.
hipError_t launch_steps(gw_batch *B, int64_t step0, int64_t n, int64_t *launches = nullptr)
{
    if (launches) *launches = 0;
    if (n <= 0) return hipSuccess;
    if (use_persist(B)) {
        if (launches) *launches = 1;
        return launch_persist(B, step0, n);
    }
    if (use_pair(B)) return launch_steps_pair(B, step0, n, launches);
    if (launches) *launches = n;
    launch_one(B, step0, 0, nullptr, nullptr);
    int64_t k = 1;
    if (B->graph_steps > 0 && n - 1 >= B->graph_steps) {
        hipError_t e = hipSuccess;
        if (!graph_current(B)) e = graph_build(B);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_set_base, dim3(1), dim3(1), 0, B->ctx->stream, B->d_step_base, step0 + 1);
        for (; k + B->graph_steps <= n; k += B->graph_steps)
            if ((e = hipGraphLaunch(B->graph_exec, B->ctx->stream)) != hipSuccess) return e;
    }
    for (; k < n; ++k) launch_one(B, step0 + k, B->ncells, nullptr, nullptr);
    launch_one(B, -1, B->ncells, nullptr, nullptr);
    return hipGetLastError();
}

// launch_steps with every step's dispatch timed through a ring of event pairs:
// before a pair is reused, the host waits for the dispatch that last recorded
// it (kRing steps back), so the queue never runs dry. The final close-only
// launch is not timed.
int launch_steps_timed(gw_batch *B, int64_t step0, int64_t n, double *ssa_ms, char *err, size_t errlen)
{
    constexpr int kRing = 512;
    std::vector<hipEvent_t> ev(2 * kRing, nullptr);
    int rc = GW_OK;
    double acc = 0.0;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            set_err(err, errlen, "hipEventCreate failed");
            rc = GW_E_HIP;
            break;
        }
    auto harvest = [&](int r) {
        float f = 0.f;
        if (hipEventSynchronize(ev[2 * r + 1]) != hipSuccess || hipEventElapsedTime(&f, ev[2 * r], ev[2 * r + 1]) != hipSuccess)
            return false;
        acc += f;
        return true;
    };
    for (int64_t k = 0; rc == GW_OK && k < n; ++k) {
        const int r = static_cast<int>(k % kRing);
        if (k >= kRing && !harvest(r)) {
            set_err(err, errlen, "event timing failed");
            rc = GW_E_HIP;
            break;
        }
        launch_one(B, step0 + k, k == 0 ? 0 : B->ncells, ev[2 * r], ev[2 * r + 1]);
    }
    if (rc == GW_OK && n > 0) launch_one(B, -1, B->ncells, nullptr, nullptr);
    for (int64_t k = std::max<int64_t>(0, n - kRing); rc == GW_OK && k < n; ++k)
        if (!harvest(static_cast<int>(k % kRing))) {
            set_err(err, errlen, "event timing failed");
            rc = GW_E_HIP;
        }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (ssa_ms) *ssa_ms = acc;
    return rc;
}

// Status and trace row count of every cell, read with one copy.
int gather_dyn(gw_batch *B, std::vector<int64_t> &out, char *err, size_t errlen)
{
    out.assign(2 * static_cast<size_t>(B->ncells) + 1, 0);
    hipLaunchKernelGGL(k_gather_dyn, dim3((B->ncells + 255) / 256), dim3(256), 0, B->ctx->stream, B->d_dyn, B->ncells,
                       B->a.timeout, B->d_gather);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out.data(), B->d_gather, out.size() * sizeof(int64_t), hipMemcpyDeviceToHost, B->ctx->stream));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return GW_OK;
}

// The whole per-cell dynamic state, read with one copy.
int read_dyn(gw_batch *B, std::vector<CellDyn> &out, char *err, size_t errlen)
{
    out.resize(static_cast<size_t>(B->ncells));
    HIP_TRY(hipMemcpyAsync(out.data(), B->d_dyn, out.size() * sizeof(CellDyn), hipMemcpyDeviceToHost, B->ctx->stream));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return GW_OK;
}

int check_status(gw_batch *B, char *err, size_t errlen)
{
    std::vector<int64_t> v;
    const int rc = gather_dyn(B, v, err, errlen);
    if (rc) return rc;
    if (v[2 * static_cast<size_t>(B->ncells)] != 0) {
        set_err(err, errlen, "a bounded wait of the persistent step kernel expired; the steps of this call are incomplete");
        return GW_E_TIMEOUT;
    }
    for (int c = 0; c < B->ncells; ++c) {
        const int32_t st = static_cast<int32_t>(v[c]);
        if (st != 0) {
            const char *what = st == GW_E_NONPOSITIVE_CONC ? "a Nernst potential received a non-positive concentration"
                             : st == GW_E_NAN ? "a global state variable became NaN"
                             : st == GW_E_RANGE ? "a CRU flux, a RyR count or the SSA iteration count left the supported range"
                                                : "device error";
            set_err(err, errlen, "cell %d: %s (status %d)", c, what, st);
            return st;
        }
    }
    return GW_OK;
}

// After a per-cell override: before the first step the whole batch is
// prepared afresh at step 0 (nothing has run); after it only this cell's
// current step is formed again (k_refresh_cell).
int refresh_cell(gw_batch *B, int cell, int resum, char *err, size_t errlen)
{
    if (!B->prepared || B->steps_done == 0) {
        B->prepared = false;
        return GW_OK;
    }
    hipLaunchKernelGGL(k_refresh_cell, dim3(1), dim3(kBlock), 0, B->ctx->stream, B->a, cell, resum);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return GW_OK;
}

int sample_rest(gw_batch *B, char *err, size_t errlen)
{
    switch (B->prng) {
    case 0: hipLaunchKernelGGL(k_sample_rest<0>, dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a); break;
    case 1: hipLaunchKernelGGL(k_sample_rest<1>, dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a); break;
    case 2: hipLaunchKernelGGL(k_sample_rest<2>, dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a); break;
    case 3: hipLaunchKernelGGL(k_sample_rest<3>, dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a); break;
    case 4: hipLaunchKernelGGL(k_sample_rest<4>, dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a); break;
    default: hipLaunchKernelGGL(k_sample_rest<5>, dim3(B->nblocks), dim3(kBlock), 0, B->ctx->stream, B->a); break;
    }
    HIP_TRY(hipGetLastError());
    return GW_OK;
}

int ensure_staging(gw_batch *B, size_t bytes, char *err, size_t errlen)
{
    if (B->staging_bytes >= bytes) return GW_OK;
    if (B->arena) {
        set_err(err, errlen, "allocation-free batch: staging of %zu bytes exceeds the %zu carved", bytes, B->staging_bytes);
        return GW_E_INVALID_ARG;
    }
    if (B->d_staging) HIP_TRY(hipFree(B->d_staging));
    B->d_staging = nullptr;
    B->staging_bytes = 0;
    HIP_TRY(hipMalloc(&B->d_staging, bytes));
    B->staging_bytes = bytes;
    return GW_OK;
}

// staging layout for one cell of n CRUs: CaSS[4n] CaJSR[n] JLCC[4n] (f64), LCC LCCi ClCh [4n] RyR[24n] (i32)
size_t staging_size(int64_t n) { return static_cast<size_t>(n) * (9 * 8 + 36 * 4); }

// The reduced view in the same staging buffer (40 bytes per CRU).
ReducedView reduced_view(void *base, int64_t n)
{
    ReducedView v;
    char *p = static_cast<char *>(base);
    v.CaSS = reinterpret_cast<float *>(p); p += 4 * n * 4;
    v.CaJSR = reinterpret_cast<float *>(p); p += n * 4;
    v.RyR = reinterpret_cast<int16_t *>(p); p += 4 * n * 2;
    v.LCC = reinterpret_cast<int8_t *>(p); p += 4 * n;
    v.LCCi = reinterpret_cast<uint8_t *>(p); p += 4 * n;
    v.ClCh = reinterpret_cast<uint8_t *>(p);
    return v;
}

StateView staging_view(void *base, int64_t n)
{
    StateView v;
    char *p = static_cast<char *>(base);
    v.CaSS = reinterpret_cast<double *>(p); p += 4 * n * 8;
    v.CaJSR = reinterpret_cast<double *>(p); p += n * 8;
    v.JLCC = reinterpret_cast<double *>(p); p += 4 * n * 8;
    v.LCC = reinterpret_cast<int32_t *>(p); p += 4 * n * 4;
    v.LCCi = reinterpret_cast<int32_t *>(p); p += 4 * n * 4;
    v.ClCh = reinterpret_cast<int32_t *>(p); p += 4 * n * 4;
    v.RyR = reinterpret_cast<int32_t *>(p);
    return v;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

void gw_params_default(gw_params_t *p)
{
    if (p) *p = default_params();
}

void gw_globals_default(gw_globals_t *g)
{
    if (g) *g = default_globals();
}

int gw_prng_from_name(const char *name)
{
    if (!name) return -1;
    static const char *names[GW_PRNG_COUNT] = {"xoshiro256**", "xoshiro256++", "xoshiro256+",
                                               "xoroshiro128**", "xoroshiro128++", "xoroshiro128+"};
    for (int i = 0; i < GW_PRNG_COUNT; ++i)
        if (std::strcmp(name, names[i]) == 0) return i;
    return -1;
}

int gw_version(void) { return GW_CAPI_VERSION; }

int gw_rest_state(int32_t n_cru, uint64_t seed, int32_t prng_id, gw_cru_state_t *out)
{
    if (n_cru <= 0 || !out || prng_id < 0 || prng_id >= GW_PRNG_COUNT) return GW_E_INVALID_ARG;
    for (int64_t i = 0; i < n_cru; ++i) {
        double ca[4], jsr;
        uint64_t sub[4];
        sample_rest_cru_any(prng_id, seed, i, ca, jsr, sub);
        for (int j = 0; j < 4; ++j) {
            if (out->CaSS) out->CaSS[i * 4 + j] = ca[j];
            if (out->JLCC) out->JLCC[i * 4 + j] = 0.0;
            if (out->LCC) out->LCC[i * 4 + j] = sub_lcc(sub[j]);
            if (out->LCC_inactivation) out->LCC_inactivation[i * 4 + j] = sub_y(sub[j]);
            if (out->ClCh) out->ClCh[i * 4 + j] = sub_cl(sub[j]);
            if (out->RyR)
                for (int k = 0; k < 6; ++k) out->RyR[(i * 4 + j) * 6 + k] = ryr_get(sub[j], k);
        }
        if (out->CaJSR) out->CaJSR[i] = jsr;
    }
    return GW_OK;
}

int gw_ctx_create(gw_ctx **out, int device, char *err, size_t errlen)
{
    if (!out) {
        set_err(err, errlen, "out is null");
        return GW_E_INVALID_ARG;
    }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_err(err, errlen, "no HIP device is visible");
        return GW_E_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        set_err(err, errlen, "device %d out of range (%d devices)", device, n);
        return GW_E_INVALID_ARG;
    }
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_err(err, errlen, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
        return GW_E_NO_DEVICE;
    }
    HIP_TRY(hipSetDevice(device));
    if (device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void *ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                set_err(err, errlen, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
                return GW_E_HIP;
            }
            g_scratch_ptr[device] = ptr;
        }
    }
    gw_ctx *c = new gw_ctx;
    c->device = device;
    c->scratch = device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_err(err, errlen, "hipStreamCreate failed");
        return GW_E_HIP;
    }
    *out = c;
    return GW_OK;
}

int gw_ctx_destroy(gw_ctx *ctx)
{
    if (!ctx) return GW_OK;
    while (!ctx->batches.empty()) gw_batch_destroy(ctx->batches.back());
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GW_OK;
}

int gw_batch_destroy(gw_batch *B)
{
    if (!B) return GW_OK;
    auto &live = B->ctx->batches;
    live.erase(std::remove(live.begin(), live.end(), B), live.end());
    (void)hipSetDevice(B->ctx->device);
    (void)hipStreamSynchronize(B->ctx->stream);
    graph_release(B);
    if (B->arena) {   // carved from the scratch array: nothing to free
        delete B;
        return GW_OK;
    }
    void *ptrs[] = {B->a.CaSS, B->a.CaJSR, B->a.JLCC, B->a.JLCCf, B->a.sub, B->a.rng, B->a.dig, B->a.partials,
                    B->a.trace, B->a.iters_total, B->d_step_base, B->d_cells, B->d_dyn, B->d_block_cell, B->d_staging,
                    B->d_gather, B->a.sync, B->a.timeout};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (double *p : B->istim_dev)
        if (p) (void)hipFree(p);
    delete B;
    return GW_OK;
}

int gw_batch_create(gw_ctx *ctx, int32_t ncells, const gw_cell_desc_t *cells, int32_t precision, int32_t prng_id,
                    double dt, const gw_globals_t *g0, gw_batch **out, char *err, size_t errlen)
{
    return batch_create(ctx, ncells, cells, precision, prng_id, dt, g0, nullptr, 0, 0, out, err, errlen);
}

}  // extern "C"

namespace {

// gw_batch_create, or with arena != null an allocation-free batch whose device
// memory is carved from arena[0, arena_size), with a stimulus table of
// table_cap steps and the staging of one cell's state carved beside it.
int batch_create(gw_ctx *ctx, int32_t ncells, const gw_cell_desc_t *cells, int32_t precision, int32_t prng_id,
                 double dt, const gw_globals_t *g0, void *arena, size_t arena_size, int64_t table_cap,
                 gw_batch **out, char *err, size_t errlen)
{
    if (!ctx || !cells || !out || ncells <= 0) {
        set_err(err, errlen, "gw_batch_create: null argument or ncells <= 0");
        return GW_E_INVALID_ARG;
    }
    *out = nullptr;
    if (!(dt > 0.0) || !std::isfinite(dt)) {
        set_err(err, errlen, "step size must be > 0");
        return GW_E_INVALID_ARG;
    }
    if (prng_id < 0 || prng_id >= GW_PRNG_COUNT) {
        set_err(err, errlen, "unknown PRNG id %d", prng_id);
        return GW_E_INVALID_ARG;
    }
    if (precision != GW_PREC_FP64 && precision != GW_PREC_FP32_LOCAL) {
        set_err(err, errlen, "unknown precision %d", precision);
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    gw_batch *B = new gw_batch;
    B->ctx = ctx;
    B->arena = static_cast<char *>(arena);
    B->arena_size = arena ? arena_size : 0;
    B->ncells = ncells;
    B->precision = precision;
    B->prng = prng_id;
    B->dt = dt;
    B->cells_h.resize(ncells);
    B->istim_dev.assign(ncells, nullptr);
    int64_t slot = 0;
    for (int c = 0; c < ncells; ++c) {
        const gw_cell_desc_t &d = cells[c];
        if (d.n_cru <= 0 || d.n_cru > (1 << 20)) {
            set_err(err, errlen, "cell %d: n_cru must be in 1..%d", c, 1 << 20);
            delete B;
            return GW_E_INVALID_ARG;
        }
        int rc = validate_params(d.params, err, errlen);
        if (rc != GW_OK) {
            delete B;
            return rc;
        }
        if (d.protocol.kind < GW_PROTO_TABLE || d.protocol.kind > GW_PROTO_CLAMP) {
            set_err(err, errlen, "cell %d: unknown protocol kind %d", c, d.protocol.kind);
            delete B;
            return GW_E_INVALID_ARG;
        }
        DevCell &C = B->cells_h[c];
        std::memset(&C, 0, sizeof(C));
        C.p = d.params;
        cell_consts_init(C.c, C.p, d.n_cru);
        ssa_params_init(C.sp, C.p, C.c);
        ssa_params_f32(C.spf, C.sp);
        C.proto = d.protocol;
        C.seed = d.seed;
        C.n_cru = d.n_cru;
        C.slot0 = slot;
        for (int sh = 0; sh < kNumShapes; ++sh) {
            const int per = crus_per_block(sh);
            C.nblk[sh] = (d.n_cru + per - 1) / per;
        }
        slot += static_cast<int64_t>((d.n_cru + kSlotAlign - 1) / kSlotAlign) * kSlotAlign;
    }
    for (int sh = 0; sh < kNumShapes; ++sh) {
        B->bc_off[sh] = static_cast<int64_t>(B->block_cell_h.size());
        int64_t blk = 0;
        for (int c = 0; c < ncells; ++c) {
            DevCell &C = B->cells_h[c];
            C.blk0[sh] = blk;
            blk += C.nblk[sh];
            for (int b = 0; b < C.nblk[sh]; ++b) B->block_cell_h.push_back(c);
        }
        B->nblocks_shape[sh] = blk;
        B->max_nblocks = std::max(B->max_nblocks, blk);
    }
    B->ntot = slot;
    B->nblocks = B->nblocks_shape[kShapeDefault];
    const int64_t N = B->ntot;
    KArgs &a = B->a;
    int rc = GW_OK;
    const bool f32 = precision == GW_PREC_FP32_LOCAL;
    if ((rc = dmalloc(B, &a.CaSS, 4 * N, err, errlen)) || (rc = dmalloc(B, &a.CaJSR, N, err, errlen)) ||
        (!f32 && (rc = dmalloc(B, &a.JLCC, 4 * N, err, errlen))) ||
        (f32 && (rc = dmalloc(B, &a.JLCCf, 4 * N, err, errlen))) ||
        (rc = dmalloc(B, &a.sub, 4 * N, err, errlen)) ||
        (rc = dmalloc(B, &a.rng, 4 * N, err, errlen)) ||
        (!f32 && (rc = dmalloc(B, &a.dig, 2 * N, err, errlen))) ||
        (rc = dmalloc(B, &a.partials, 2 * B->max_nblocks, err, errlen)) ||
        (rc = dmalloc(B, &a.iters_total, 1, err, errlen)) || (rc = dmalloc(B, &B->d_step_base, 1, err, errlen)) ||
        (rc = dmalloc(B, &B->d_cells, ncells, err, errlen)) || (rc = dmalloc(B, &B->d_dyn, ncells, err, errlen)) ||
        (rc = dmalloc(B, &B->d_gather, 2 * static_cast<size_t>(ncells) + 1, err, errlen)) ||
        (rc = dmalloc(B, &a.sync, static_cast<size_t>(ncells), err, errlen)) ||
        (rc = dmalloc(B, &a.timeout, 1, err, errlen)) ||
        (rc = dmalloc(B, &B->d_block_cell, B->block_cell_h.size(), err, errlen))) {
        gw_batch_destroy(B);
        return rc;
    }
    if (B->arena) {
        // the staging of one cell's state (set/get_cru_state) and the stimulus table
        void *stg = nullptr;
        const size_t sb = staging_size(B->cells_h[0].n_cru);
        if ((rc = carve(B, &stg, sb, err, errlen)) ||
            (table_cap > 0 && (rc = dmalloc(B, &B->arena_table, static_cast<size_t>(table_cap), err, errlen)))) {
            gw_batch_destroy(B);
            return rc;
        }
        B->d_staging = stg;
        B->staging_bytes = sb;
        B->arena_table_cap = table_cap;
    }
    a.cells = B->d_cells;
    a.dyn = B->d_dyn;
    a.block_cell = B->d_block_cell + B->bc_off[kShapeDefault];
    a.shape = kShapeDefault;
    a.ntot = N;
    a.nblocks = B->nblocks;
    a.dt = dt;
    a.ncells = ncells;
    a.frozen = 0;
    a.trace_every = 0;
    a.trace_cap = 0;
    a.trace = nullptr;
    a.prng = prng_id;
    a.precision = precision;
    a.spin_limit = kSpinLimit;
    // per-cell stimulus tables are attached later; globals start at g0 or the defaults
    std::vector<CellDyn> dyn(ncells);
    B->g0.resize(ncells);
    for (int c = 0; c < ncells; ++c) {
        std::memset(&dyn[c], 0, sizeof(CellDyn));
        dyn[c].g = g0 ? *g0 : default_globals();
        dyn[c].clamp = (cells[c].protocol.kind == GW_PROTO_CLAMP) ? 1 : 0;
        B->g0[c] = dyn[c].g;
    }
    auto fail = [&](hipError_t e) {
        set_err(err, errlen, "HIP error during batch creation: %s", hipGetErrorString(e));
        gw_batch_destroy(B);
        return GW_E_HIP;
    };
    hipError_t e;
    if ((e = hipMemcpyAsync(B->d_cells, B->cells_h.data(), ncells * sizeof(DevCell), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(B->d_dyn, dyn.data(), ncells * sizeof(CellDyn), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(B->d_block_cell, B->block_cell_h.data(), B->block_cell_h.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, a.iters_total, int64_t(1), 0ull);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_fill_u64, dim3((ncells * (sizeof(CellSync) / 8) + 255) / 256), dim3(256), 0, ctx->stream,
                       reinterpret_cast<unsigned long long *>(a.sync), int64_t(ncells) * int64_t(sizeof(CellSync) / 8),
                       0ull);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, a.timeout, int64_t(1));
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_seed, dim3(B->nblocks), dim3(kBlock), 0, ctx->stream, a);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((rc = sample_rest(B, err, errlen)) != GW_OK) {
        gw_batch_destroy(B);
        return rc;
    }
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return fail(e);
    ctx->batches.push_back(B);
    *out = B;
    return GW_OK;
}

}  // namespace

extern "C" {

int gw_batch_set_cru_state(gw_batch *B, int32_t cell, const gw_cru_state_t *c, char *err, size_t errlen)
{
    if (!B || !c || cell < 0 || cell >= B->ncells || !c->CaSS || !c->CaJSR || !c->LCC || !c->LCC_inactivation ||
        !c->RyR || !c->ClCh) {
        set_err(err, errlen, "gw_batch_set_cru_state: bad arguments");
        return GW_E_INVALID_ARG;
    }
    const int64_t n = B->cells_h[cell].n_cru;
    for (int64_t i = 0; i < 4 * n; ++i) {
        if (c->LCC[i] < 1 || c->LCC[i] > 12) {
            set_err(err, errlen, "LCC states must be integers in 1..12 (got %d at flat index %lld)", c->LCC[i], static_cast<long long>(i));
            return GW_E_INVALID_ARG;
        }
        if ((c->LCC_inactivation[i] & ~1) || (c->ClCh[i] & ~1)) {
            set_err(err, errlen, "LCC_inactivation and ClCh must be 0 or 1 (flat index %lld)", static_cast<long long>(i));
            return GW_E_INVALID_ARG;
        }
        if (!std::isfinite(c->CaSS[i])) {
            set_err(err, errlen, "CaSS must be finite");
            return GW_E_INVALID_ARG;
        }
    }
    for (int64_t i = 0; i < 24 * n; ++i)
        if (c->RyR[i] < -100 || c->RyR[i] > 100) {
            set_err(err, errlen, "RyR counts must lie in -100..100");
            return GW_E_INVALID_ARG;
        }
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(c->CaJSR[i])) {
            set_err(err, errlen, "CaJSR must be finite");
            return GW_E_INVALID_ARG;
        }
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = ensure_staging(B, staging_size(n), err, errlen);
    if (rc) return rc;
    StateView v = staging_view(B->d_staging, n);
    hipStream_t s = B->ctx->stream;
    HIP_TRY(hipMemcpyAsync(v.CaSS, c->CaSS, 4 * n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.CaJSR, c->CaJSR, n * 8, hipMemcpyHostToDevice, s));
    if (c->JLCC) HIP_TRY(hipMemcpyAsync(v.JLCC, c->JLCC, 4 * n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.LCC, c->LCC, 4 * n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.LCCi, c->LCC_inactivation, 4 * n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.ClCh, c->ClCh, 4 * n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(v.RyR, c->RyR, 24 * n * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_put_state, dim3((n + 255) / 256), dim3(256), 0, s, B->a, cell, v, c->JLCC ? 1 : 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return refresh_cell(B, cell, 1, err, errlen);
}

int gw_batch_get_cru_state(gw_batch *B, int32_t cell, gw_cru_state_t *c, char *err, size_t errlen)
{
    if (!B || !c || cell < 0 || cell >= B->ncells) {
        set_err(err, errlen, "gw_batch_get_cru_state: bad arguments");
        return GW_E_INVALID_ARG;
    }
    const int64_t n = B->cells_h[cell].n_cru;
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = ensure_staging(B, staging_size(n), err, errlen);
    if (rc) return rc;
    StateView v = staging_view(B->d_staging, n);
    hipStream_t s = B->ctx->stream;
    hipLaunchKernelGGL(k_get_state, dim3((n + 255) / 256), dim3(256), 0, s, B->a, cell, v);
    HIP_TRY(hipGetLastError());
    if (c->CaSS) HIP_TRY(hipMemcpyAsync(c->CaSS, v.CaSS, 4 * n * 8, hipMemcpyDeviceToHost, s));
    if (c->CaJSR) HIP_TRY(hipMemcpyAsync(c->CaJSR, v.CaJSR, n * 8, hipMemcpyDeviceToHost, s));
    if (c->JLCC) HIP_TRY(hipMemcpyAsync(c->JLCC, v.JLCC, 4 * n * 8, hipMemcpyDeviceToHost, s));
    if (c->LCC) HIP_TRY(hipMemcpyAsync(c->LCC, v.LCC, 4 * n * 4, hipMemcpyDeviceToHost, s));
    if (c->LCC_inactivation) HIP_TRY(hipMemcpyAsync(c->LCC_inactivation, v.LCCi, 4 * n * 4, hipMemcpyDeviceToHost, s));
    if (c->ClCh) HIP_TRY(hipMemcpyAsync(c->ClCh, v.ClCh, 4 * n * 4, hipMemcpyDeviceToHost, s));
    if (c->RyR) HIP_TRY(hipMemcpyAsync(c->RyR, v.RyR, 24 * n * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GW_OK;
}

int gw_batch_get_cru_state_reduced(gw_batch *B, int32_t cell, gw_cru_state_reduced_t *c, char *err, size_t errlen)
{
    if (!B || !c || cell < 0 || cell >= B->ncells) {
        set_err(err, errlen, "gw_batch_get_cru_state_reduced: bad arguments");
        return GW_E_INVALID_ARG;
    }
    const int64_t n = B->cells_h[cell].n_cru;
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = ensure_staging(B, staging_size(n), err, errlen);
    if (rc) return rc;
    const ReducedView v = reduced_view(B->d_staging, n);
    hipStream_t s = B->ctx->stream;
    hipLaunchKernelGGL(k_get_state_reduced, dim3((n + 255) / 256), dim3(256), 0, s, B->a, cell, v);
    HIP_TRY(hipGetLastError());
    if (c->CaSS) HIP_TRY(hipMemcpyAsync(c->CaSS, v.CaSS, 4 * n * 4, hipMemcpyDeviceToHost, s));
    if (c->CaJSR) HIP_TRY(hipMemcpyAsync(c->CaJSR, v.CaJSR, n * 4, hipMemcpyDeviceToHost, s));
    if (c->RyR) HIP_TRY(hipMemcpyAsync(c->RyR, v.RyR, 4 * n * 2, hipMemcpyDeviceToHost, s));
    if (c->LCC) HIP_TRY(hipMemcpyAsync(c->LCC, v.LCC, 4 * n, hipMemcpyDeviceToHost, s));
    if (c->LCC_inactivation) HIP_TRY(hipMemcpyAsync(c->LCC_inactivation, v.LCCi, 4 * n, hipMemcpyDeviceToHost, s));
    if (c->ClCh) HIP_TRY(hipMemcpyAsync(c->ClCh, v.ClCh, 4 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GW_OK;
}

int gw_batch_set_globals(gw_batch *B, int32_t cell, const gw_globals_t *g, char *err, size_t errlen)
{
    if (!B || !g || cell < 0 || cell >= B->ncells) {
        set_err(err, errlen, "gw_batch_set_globals: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    HIP_TRY(hipMemcpyAsync(&B->d_dyn[cell].g, g, sizeof(gw_globals_t), hipMemcpyHostToDevice, B->ctx->stream));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return refresh_cell(B, cell, 0, err, errlen);
}

int gw_batch_get_globals(gw_batch *B, gw_globals_t *g_out, char *err, size_t errlen)
{
    if (!B || !g_out) {
        set_err(err, errlen, "gw_batch_get_globals: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    std::vector<CellDyn> d;
    const int rc = read_dyn(B, d, err, errlen);
    if (rc) return rc;
    for (int c = 0; c < B->ncells; ++c) g_out[c] = d[c].g;
    return GW_OK;
}

int gw_batch_set_stimulus_table(gw_batch *B, int32_t cell, const double *istim, int64_t n, char *err, size_t errlen)
{
    if (!B || cell < 0 || cell >= B->ncells || (n > 0 && !istim) || n < 0) {
        set_err(err, errlen, "gw_batch_set_stimulus_table: bad arguments");
        return GW_E_INVALID_ARG;
    }
    if (B->cells_h[cell].proto.kind == GW_PROTO_CLAMP) {
        set_err(err, errlen, "gw_batch_set_stimulus_table: cell %d is voltage-clamped; a stimulus table has no effect on it",
                cell);
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    double *d = nullptr;
    if (B->arena) {   // the table carved at creation (cell 0 only)
        if (cell != 0 || n > B->arena_table_cap) {
            set_err(err, errlen, "allocation-free batch: a stimulus table of %lld steps exceeds the %lld carved",
                    static_cast<long long>(n), static_cast<long long>(B->arena_table_cap));
            return GW_E_INVALID_ARG;
        }
        if (n > 0) {
            d = B->arena_table;
            HIP_TRY(hipMemcpyAsync(d, istim, n * sizeof(double), hipMemcpyHostToDevice, B->ctx->stream));
        }
    } else {
        if (B->istim_dev[cell]) HIP_TRY(hipFree(B->istim_dev[cell]));
        B->istim_dev[cell] = nullptr;
        if (n > 0) {
            HIP_TRY(hipMalloc(&d, n * sizeof(double)));
            HIP_TRY(hipMemcpyAsync(d, istim, n * sizeof(double), hipMemcpyHostToDevice, B->ctx->stream));
        }
        B->istim_dev[cell] = d;
    }
    B->cells_h[cell].istim = d;
    B->cells_h[cell].istim_len = n;
    B->cells_h[cell].proto.kind = GW_PROTO_TABLE;
    HIP_TRY(hipMemcpyAsync(&B->d_cells[cell], &B->cells_h[cell], sizeof(DevCell), hipMemcpyHostToDevice, B->ctx->stream));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return refresh_cell(B, cell, 0, err, errlen);
}

int gw_batch_step(gw_batch *B, int64_t nsteps, char *err, size_t errlen)
{
    if (!B || nsteps < 0) {
        set_err(err, errlen, "gw_batch_step: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = launch_cfg_check(B, err, errlen);
    if (rc) return rc;
    rc = prepare(B, err, errlen);
    if (rc) return rc;
    if (B->frozen) {
        // one launch per chunk; a chunk is at most ~2^26 CRU-steps of work
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(1 << 18, (1LL << 26) / std::max<int64_t>(1, B->ntot)));
        for (int64_t done = 0; done < nsteps;) {
            const int64_t n = std::min(chunk, nsteps - done);
            launch_frozen_steps(B, B->steps_done, n);
            HIP_TRY(hipGetLastError());
            done += n;
            B->steps_done += n;
            rc = check_status(B, err, errlen);
            if (rc) return rc;
        }
        return GW_OK;
    }
    const int64_t chunk = 4096;
    for (int64_t done = 0; done < nsteps;) {
        const int64_t n = std::min(chunk, nsteps - done);
        HIP_TRY(launch_steps(B, B->steps_done, n));
        done += n;
        B->steps_done += n;
        rc = check_status(B, err, errlen);
        if (rc) return rc;
    }
    return GW_OK;
}

int gw_batch_get_time(gw_batch *B, double *t, int64_t *steps_done)
{
    if (!B) return GW_E_INVALID_ARG;
    (void)hipSetDevice(B->ctx->device);
    if (!B->prepared) {
        if (t) *t = 0.0;
        if (steps_done) *steps_done = 0;
        return GW_OK;
    }
    // every running cell shares the clock; a cell that failed stopped its own
    std::vector<CellDyn> d;
    if (read_dyn(B, d, nullptr, 0) != GW_OK) return GW_E_HIP;
    int pick = -1;
    for (int c = 0; c < B->ncells && pick < 0; ++c)
        if (d[c].status == 0) pick = c;
    if (pick < 0) {
        pick = 0;
        for (int c = 1; c < B->ncells; ++c)
            if (d[c].step > d[pick].step) pick = c;
    }
    if (t) *t = d[pick].t;
    if (steps_done) *steps_done = d[pick].step;
    return GW_OK;
}

int gw_batch_get_obs(gw_batch *B, gw_cell_obs_t *obs, char *err, size_t errlen)
{
    if (!B || !obs) {
        set_err(err, errlen, "gw_batch_get_obs: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    std::vector<CellDyn> d;
    const int rc = read_dyn(B, d, err, errlen);
    if (rc) return rc;
    for (int c = 0; c < B->ncells; ++c) obs[c] = d[c].obs;
    return GW_OK;
}

int gw_batch_reset_obs(gw_batch *B, char *err, size_t errlen)
{
    if (!B) return GW_E_INVALID_ARG;
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = prepare(B, err, errlen);
    if (rc) return rc;
    hipLaunchKernelGGL(k_reset_obs, dim3((B->ncells + 255) / 256), dim3(256), 0, B->ctx->stream, B->d_dyn, B->ncells);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return GW_OK;
}

int gw_batch_enable_trace(gw_batch *B, int32_t every, int64_t capacity, char *err, size_t errlen)
{
    if (!B || every < 0 || capacity < 0) {
        set_err(err, errlen, "gw_batch_enable_trace: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    if (B->a.trace) HIP_TRY(hipFree(B->a.trace));
    B->a.trace = nullptr;
    B->a.trace_every = 0;
    B->a.trace_cap = 0;
    if (every > 0 && capacity > 0) {
        HIP_TRY(hipMalloc(&B->a.trace, static_cast<size_t>(capacity) * B->ncells * 4 * sizeof(double)));
        B->a.trace_every = every;
        B->a.trace_cap = capacity;
    }
    hipLaunchKernelGGL(k_zero_trace_rows, dim3((B->ncells + 255) / 256), dim3(256), 0, B->ctx->stream, B->d_dyn, B->ncells);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return GW_OK;
}

int gw_batch_get_trace(gw_batch *B, double *out, int64_t max_rows, int64_t *rows, char *err, size_t errlen)
{
    if (!B || !rows) {
        set_err(err, errlen, "gw_batch_get_trace: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    std::vector<int64_t> v;
    int rc = gather_dyn(B, v, err, errlen);
    if (rc) return rc;
    const int n = B->ncells;
    int64_t r = 0;
    for (int c = 0; c < n; ++c) r = std::max(r, v[n + c]);
    r = std::min(r, max_rows);
    if (r > 0 && out) {
        HIP_TRY(hipMemcpyAsync(out, B->a.trace, static_cast<size_t>(r) * n * 4 * sizeof(double), hipMemcpyDeviceToHost, B->ctx->stream));
        HIP_TRY(hipStreamSynchronize(B->ctx->stream));
        // a cell that stopped early (non-zero status) wrote fewer rows: its later rows read NaN
        for (int c = 0; c < n; ++c)
            for (int64_t k = std::max<int64_t>(0, v[n + c]); k < r; ++k)
                for (int q = 0; q < 4; ++q) out[(k * n + c) * 4 + q] = std::nan("");
    }
    hipLaunchKernelGGL(k_zero_trace_rows, dim3((n + 255) / 256), dim3(256), 0, B->ctx->stream, B->d_dyn, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    *rows = r;
    return GW_OK;
}

int gw_batch_get_event_digest(gw_batch *B, int32_t cell, uint64_t *count, uint64_t *hash, char *err, size_t errlen)
{
    if (!B || cell < 0 || cell >= B->ncells || !count || !hash) {
        set_err(err, errlen, "gw_batch_get_event_digest: bad arguments");
        return GW_E_INVALID_ARG;
    }
    if (!B->a.dig) {
        set_err(err, errlen, "event digests are not kept by this batch");
        return GW_E_INVALID_ARG;
    }
    const DevCell &C = B->cells_h[cell];
    HIP_TRY(hipSetDevice(B->ctx->device));
    std::vector<uint64_t> pairs(2 * static_cast<size_t>(C.n_cru));
    HIP_TRY(hipMemcpyAsync(pairs.data(), B->a.dig + 2 * C.slot0, pairs.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, B->ctx->stream));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    for (int64_t i = 0; i < C.n_cru; ++i) {
        count[i] = pairs[2 * i];
        hash[i] = pairs[2 * i + 1];
    }
    return GW_OK;
}

int gw_batch_get_iterations(gw_batch *B, uint64_t *iters)
{
    if (!B || !iters) return GW_E_INVALID_ARG;
    (void)hipSetDevice(B->ctx->device);
    unsigned long long v = 0;
    (void)hipMemcpyAsync(&v, B->a.iters_total, sizeof(v), hipMemcpyDeviceToHost, B->ctx->stream);
    if (hipStreamSynchronize(B->ctx->stream) != hipSuccess) return GW_E_HIP;
    *iters = v;
    return GW_OK;
}

int gw_launch_config(int32_t index, int32_t *lanes_per_cru, int32_t *block, int32_t *crus_per_lane)
{
    if (index < 0 || index >= kNumConfigs) return GW_E_INVALID_ARG;
    if (lanes_per_cru) *lanes_per_cru = kConfigs[index].lanes;
    if (block) *block = kConfigs[index].threads;
    if (crus_per_lane) *crus_per_lane = kConfigs[index].cpl;
    return GW_OK;
}

int gw_batch_set_launch(gw_batch *B, int32_t lanes_per_cru, int32_t block, int32_t crus_per_lane, char *err,
                        size_t errlen)
{
    if (!B) return GW_E_INVALID_ARG;
    const int lanes = lanes_per_cru ? lanes_per_cru : 4;
    const int threads = block ? block : (lanes == 4 ? kBlock : 64);
    const int cpl = crus_per_lane ? crus_per_lane : 1;
    int c = -1;
    for (int k = 0; k < kNumConfigs; ++k)
        if (kConfigs[k].lanes == lanes && kConfigs[k].threads == threads && kConfigs[k].cpl == cpl) c = k;
    if (c < 0) {
        set_err(err, errlen, "launch shape of %d lanes per CRU x %d threads x %d CRUs per lane is not available "
                "(4 lanes: 64, 128, 256 or 512 threads with 1 CRU per quad, 64 with 2, 64, 128 or 256 with 4; "
                "1 lane: 64, 128 or 256 threads, 1, 2, 4 or 8 CRUs per lane)", lanes_per_cru, block, crus_per_lane);
        return GW_E_INVALID_ARG;
    }
    if (!shape_allowed(B, c)) {
        set_err(err, errlen, "launch shape of %d lanes per CRU x %d threads is not available for this batch (other "
                "shapes than 4 x 256 need xoshiro256** without replay digests or frozen globals; 1 lane per CRU needs "
                "FP32-local; refilled quads need FP64)", lanes, threads);
        return GW_E_INVALID_ARG;
    }
    set_shape(B, c);
    return GW_OK;
}

int gw_batch_get_launch(gw_batch *B, int32_t *lanes_per_cru, int32_t *block, int32_t *crus_per_lane)
{
    if (!B) return GW_E_INVALID_ARG;
    return gw_launch_config(B->cfg, lanes_per_cru, block, crus_per_lane);
}

int gw_batch_autotune(gw_batch *B, int64_t trial_steps, int32_t *lanes_out, int32_t *block_out, int32_t *cpl_out,
                      int32_t *graph_out, double *ms_per_step, char *err, size_t errlen)
{
    if (!B || trial_steps <= 0) {
        set_err(err, errlen, "gw_batch_autotune: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = prepare(B, err, errlen);
    if (rc) return rc;
    hipStream_t st = B->ctx->stream;
    const int64_t N = B->ntot;
    KArgs &a = B->a;
    // everything a step changes; restored after every trial
    struct Piece {
        void *p;
        size_t n;
    };
    std::vector<Piece> pieces = {{a.CaSS, 32 * static_cast<size_t>(N)},
                                 {a.CaJSR, 8 * static_cast<size_t>(N)},
                                 {a.JLCC ? static_cast<void *>(a.JLCC) : static_cast<void *>(a.JLCCf),
                                  (a.JLCC ? 32 : 16) * static_cast<size_t>(N)},
                                 {a.sub, 32 * static_cast<size_t>(N)},
                                 {a.rng, 32 * static_cast<size_t>(N)},
                                 {B->d_dyn, sizeof(CellDyn) * static_cast<size_t>(B->ncells)},
                                 {a.iters_total, sizeof(unsigned long long)}};
    if (a.dig) pieces.push_back({a.dig, 16 * static_cast<size_t>(N)});
    size_t total = 0;
    for (const Piece &q : pieces) total += (q.n + 255) & ~static_cast<size_t>(255);
    char *snap = nullptr;
    HIP_TRY(hipMalloc(&snap, total));
    auto copy_all = [&](bool save) {
        size_t off = 0;
        for (const Piece &q : pieces) {
            const hipError_t e = save ? hipMemcpyAsync(snap + off, q.p, q.n, hipMemcpyDeviceToDevice, st)
                                      : hipMemcpyAsync(q.p, snap + off, q.n, hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return e;
            off += (q.n + 255) & ~static_cast<size_t>(255);
        }
        return hipSuccess;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = copy_all(true);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    const int64_t steps0 = B->steps_done;
    int best_cfg = B->cfg;
    double best_ms = 1e300;
    for (int c = 0; c < kNumConfigs && he == hipSuccess; ++c) {
        if (ms_per_step) ms_per_step[c] = -1.0;
        if (!shape_allowed(B, c)) continue;
        set_shape(B, c);
        if ((he = launch_steps(B, B->steps_done, kAutotuneWarm)) != hipSuccess) break;
        B->steps_done += kAutotuneWarm;
        if ((he = hipEventRecord(e0, st)) != hipSuccess) break;
        if ((he = launch_steps(B, B->steps_done, trial_steps)) != hipSuccess) break;
        if ((he = hipEventRecord(e1, st)) != hipSuccess) break;
        if ((he = hipEventSynchronize(e1)) != hipSuccess) break;
        float f = 0.f;
        if ((he = hipEventElapsedTime(&f, e0, e1)) != hipSuccess) break;
        const double ms = f / static_cast<double>(trial_steps);
        if (ms_per_step) ms_per_step[c] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best_cfg = c;
        }
        he = copy_all(false);
        B->steps_done = steps0;
    }
    // the fastest shape with the other number of SSA steps per launch, where
    // it can run both (the CRU state kept in registers across two steps, or
    // streamed through HBM every step)
    const int32_t spl0 = B->steps_per_launch;
    int32_t best_spl = spl0;
    set_shape(B, best_cfg);
    if (he == hipSuccess && pair_capable(B)) {
        const int32_t used = use_pair(B) ? 2 : 1;   // the choice the trials above ran with
        best_spl = used;
        B->steps_per_launch = 3 - used;
        if ((he = launch_steps(B, B->steps_done, kAutotuneWarm)) == hipSuccess) {
            B->steps_done += kAutotuneWarm;
            if ((he = hipEventRecord(e0, st)) == hipSuccess && (he = launch_steps(B, B->steps_done, trial_steps)) == hipSuccess &&
                (he = hipEventRecord(e1, st)) == hipSuccess && (he = hipEventSynchronize(e1)) == hipSuccess) {
                float f = 0.f;
                if ((he = hipEventElapsedTime(&f, e0, e1)) == hipSuccess) {
                    const double ms = f / static_cast<double>(trial_steps);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best_spl = B->steps_per_launch;
                    }
                }
            }
        }
        if (he == hipSuccess) he = copy_all(false);
        B->steps_done = steps0;
        B->steps_per_launch = best_spl;
    }
    // the fastest shape once more with its steps replayed as captured graphs
    const int32_t graph0 = B->graph_steps;
    int32_t best_graph = 0;
    if (ms_per_step) ms_per_step[kNumConfigs] = -1.0;
    if (he == hipSuccess && trial_steps > kAutotuneGraphSteps) {
        set_shape(B, best_cfg);
        B->graph_steps = kAutotuneGraphSteps;
        // the warm-up captures and replays the graph once, outside the timed window
        if ((he = launch_steps(B, B->steps_done, kAutotuneWarm)) == hipSuccess) {
            B->steps_done += kAutotuneWarm;
            if ((he = hipEventRecord(e0, st)) == hipSuccess && (he = launch_steps(B, B->steps_done, trial_steps)) == hipSuccess &&
                (he = hipEventRecord(e1, st)) == hipSuccess && (he = hipEventSynchronize(e1)) == hipSuccess) {
                float f = 0.f;
                if ((he = hipEventElapsedTime(&f, e0, e1)) == hipSuccess) {
                    const double ms = f / static_cast<double>(trial_steps);
                    if (ms_per_step) ms_per_step[kNumConfigs] = ms;
                    // graphs only when they save more than noise (launch-bound batches)
                    if (ms < 0.98 * best_ms) best_graph = kAutotuneGraphSteps;
                }
            }
        }
        if (he == hipSuccess) he = copy_all(false);
        B->steps_done = steps0;
    }
    B->graph_steps = (he == hipSuccess) ? best_graph : graph0;
    B->steps_per_launch = (he == hipSuccess) ? best_spl : spl0;
    if (he == hipSuccess) he = copy_all(false);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    B->steps_done = steps0;
    set_shape(B, best_cfg);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(snap);
    if (he != hipSuccess) {
        set_err(err, errlen, "autotune: %s", hipGetErrorString(he));
        return GW_E_HIP;
    }
    gw_batch_get_launch(B, lanes_out, block_out, cpl_out);
    if (graph_out) *graph_out = B->graph_steps;
    return check_status(B, err, errlen);
}

int gw_batch_time_steps(gw_batch *B, int64_t nsteps, double *ms, double *ssa_ms, int64_t *launches, char *err,
                        size_t errlen)
{
    if (!B || nsteps <= 0 || !ms) {
        set_err(err, errlen, "gw_batch_time_steps: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = prepare(B, err, errlen);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, B->ctx->stream));
    const bool persist = use_persist(B);
    int64_t nlaunch = nsteps;   // SSA launches
    if (ssa_ms && use_pair(B)) {   // two steps per launch, each SSA dispatch timed
        HIP_TRY(launch_steps_pair(B, B->steps_done, nsteps, &nlaunch, ssa_ms));
    } else if (ssa_ms && !persist) {   // one step per launch, each dispatch timed
        rc = launch_steps_timed(B, B->steps_done, nsteps, ssa_ms, err, errlen);
        if (rc) return rc;
    } else {
        HIP_TRY(launch_steps(B, B->steps_done, nsteps, &nlaunch));
    }
    HIP_TRY(hipGetLastError());
    B->steps_done += nsteps;
    HIP_TRY(hipEventRecord(e1, B->ctx->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f;
    if (ssa_ms && persist) *ssa_ms = f;   // one launch runs every step
    if (launches) *launches = nlaunch;
    return check_status(B, err, errlen);
}

int gw_batch_set_persistent(gw_batch *B, int32_t mode, char *err, size_t errlen)
{
    if (!B || mode < 0 || mode > 2) {
        set_err(err, errlen, "gw_batch_set_persistent: mode must be 0 (off), 1 (on) or 2 (when it fits)");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    if (mode == 1 && !persist_fits(B)) {
        set_err(err, errlen, "gw_batch_set_persistent: the batch's %lld SSA blocks and %d close blocks cannot all be "
                "resident at once", static_cast<long long>(B->nblocks_shape[kShapeDefault]), B->ncells);
        return GW_E_INVALID_ARG;
    }
    B->persist_mode = mode;
    return GW_OK;
}

int gw_batch_get_currents(gw_batch *B, double *ICaL, double *Ito2, char *err, size_t errlen)
{
    if (!B || (!ICaL && !Ito2)) {
        set_err(err, errlen, "gw_batch_get_currents: bad arguments");
        return GW_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(B->ctx->device));
    int rc = prepare(B, err, errlen);
    if (rc) return rc;
    const size_t bytes = 2 * sizeof(double) * static_cast<size_t>(B->ncells);
    rc = ensure_staging(B, bytes, err, errlen);
    if (rc) return rc;
    double *d = static_cast<double *>(B->d_staging);
    hipLaunchKernelGGL(k_posthoc_currents, dim3(B->ncells), dim3(kBlock), 0, B->ctx->stream, B->a, d);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(2 * static_cast<size_t>(B->ncells));
    HIP_TRY(hipMemcpyAsync(h.data(), d, bytes, hipMemcpyDeviceToHost, B->ctx->stream));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    for (int c = 0; c < B->ncells; ++c) {
        if (ICaL) ICaL[c] = h[2 * c];
        if (Ito2) Ito2[c] = h[2 * c + 1];
    }
    return GW_OK;
}

int gw_batch_set_spin_limit(gw_batch *B, int64_t polls, char *err, size_t errlen)
{
    if (!B || polls < 0) {
        set_err(err, errlen, "gw_batch_set_spin_limit: polls must be >= 0");
        return GW_E_INVALID_ARG;
    }
    B->a.spin_limit = polls > 0 ? polls : kSpinLimit;
    return GW_OK;
}

int gw_batch_get_persistent(gw_batch *B, int32_t *active)
{
    if (!B || !active) return GW_E_INVALID_ARG;
    (void)hipSetDevice(B->ctx->device);
    *active = use_persist(B) ? 1 : 0;
    return GW_OK;
}

int gw_batch_set_steps_per_launch(gw_batch *B, int32_t steps, char *err, size_t errlen)
{
    if (!B || steps < 0 || steps > 2) {
        set_err(err, errlen, "gw_batch_set_steps_per_launch: steps must be 1, 2 or 0 (automatic)");
        return GW_E_INVALID_ARG;
    }
    B->steps_per_launch = steps;
    return GW_OK;
}

int gw_batch_get_steps_per_launch(gw_batch *B, int32_t *steps)
{
    if (!B || !steps) return GW_E_INVALID_ARG;
    (void)hipSetDevice(B->ctx->device);
    *steps = use_pair(B) ? 2 : 1;
    return GW_OK;
}

int gw_batch_set_graph(gw_batch *B, int32_t steps_per_graph, char *err, size_t errlen)
{
    if (!B || steps_per_graph < 0 || steps_per_graph > 4096) {
        set_err(err, errlen, "gw_batch_set_graph: steps_per_graph must be in 0..4096");
        return GW_E_INVALID_ARG;
    }
    B->graph_steps = steps_per_graph;
    return GW_OK;
}

// Internal switches used by gw_run and the replay tests.
int gw_batch_set_mode(gw_batch *B, int32_t frozen, int32_t digest)
{
    if (!B) return GW_E_INVALID_ARG;
    B->frozen = frozen != 0;
    B->a.frozen = frozen != 0;
    B->digest = digest != 0;
    if (!shape_allowed(B, B->cfg)) set_shape(B, kConfigDefault);
    return GW_OK;
}

}  // extern "C"

namespace {

// The loop of run<PRNG> (src/pyGW.hpp:102-122) on a one-cell batch; record(B,
// k, t) writes record k (after step i with i % record_every == 0) while k is
// below num_records.
template <class Record>
int run_loop(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt, int64_t num_steps, const double *istim,
             int32_t record_every, const gw_globals_t *g0, const gw_cru_state_t *c0, uint64_t seed, int32_t prng_id,
             int32_t precision, int64_t num_records, Record &&record, char *err, size_t errlen,
             bool noalloc = false)
{
    if (!ctx || !p || n_cru <= 0 || num_steps < 0 || record_every <= 0) {
        set_err(err, errlen, "gw_run: invalid arguments (n_cru > 0, num_steps >= 0, record_every > 0 required)");
        return GW_E_INVALID_ARG;
    }
    gw_cell_desc_t d;
    std::memset(&d, 0, sizeof(d));
    d.n_cru = n_cru;
    d.seed = seed;
    d.params = *p;
    d.protocol.kind = GW_PROTO_TABLE;
    gw_batch *B = nullptr;
    int rc;
    if (noalloc) {
        if (!ctx->scratch) {
            set_err(err, errlen, "gw_run_noalloc: no scratch address for device %d", ctx->device);
            return GW_E_INVALID_ARG;
        }
        rc = batch_create(ctx, 1, &d, precision, prng_id, dt, g0, ctx->scratch, kScratchBytes,
                          istim ? num_steps : 0, &B, err, errlen);
    } else {
        rc = gw_batch_create(ctx, 1, &d, precision, prng_id, dt, g0, &B, err, errlen);
    }
    if (rc) return rc;
    gw_batch_set_mode(B, 0, 0);
    if (c0) {
        gw_cru_state_t c = *c0;
        c.JLCC = nullptr;  // a reference run starts with JLCC = 0 (GW.hpp:104-113)
        rc = gw_batch_set_cru_state(B, 0, &c, err, errlen);
        if (rc) { gw_batch_destroy(B); return rc; }
    }
    if (num_steps > 0 && istim) {
        rc = gw_batch_set_stimulus_table(B, 0, istim, num_steps, err, errlen);
        if (rc) { gw_batch_destroy(B); return rc; }
    }
    double t = 0.0;
    int64_t counter = 0;
    int64_t i = 0;
    while (i < num_steps) {
        // next record after step i_rec (i_rec % record_every == 0, i_rec >= i)
        const int64_t i_rec = ((i + record_every - 1) / record_every) * record_every;
        const int64_t upto = std::min(num_steps, i_rec + 1);
        rc = gw_batch_step(B, upto - i, err, errlen);
        if (rc) { gw_batch_destroy(B); return rc; }
        for (int64_t k = i; k < upto; ++k) t += dt;
        i = upto;
        if (i == i_rec + 1) {
            if (counter < num_records) {
                rc = record(B, counter, t);
                if (rc) { gw_batch_destroy(B); return rc; }
            }
            ++counter;
        }
    }
    gw_batch_destroy(B);
    return GW_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// gw_run and gw_run_noalloc: the same run, records in the full layout.
int run_full(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt, int64_t num_steps, const double *istim,
             int32_t record_every, const gw_globals_t *g0, const gw_cru_state_t *c0, uint64_t seed, int32_t prng_id,
             int32_t precision, gw_record_t *out, char *err, size_t errlen, bool noalloc)
{
    const int64_t n = n_cru;
    auto record = [&](gw_batch *B, int64_t k, double t) {
        gw_globals_t g;
        int rc = gw_batch_get_globals(B, &g, err, errlen);
        if (rc) return rc;
#define PUT(f, v) if (out->f) out->f[k] = (v)
        PUT(t, t); PUT(V, g.V); PUT(m, g.m); PUT(h, g.h); PUT(j, g.j); PUT(Nai, g.Nai); PUT(Ki, g.Ki);
        PUT(Cai, g.Cai); PUT(CaNSR, g.CaNSR); PUT(CaLTRPN, g.CaLTRPN); PUT(CaHTRPN, g.CaHTRPN); PUT(xKs, g.xKs);
#undef PUT
        if (out->XKr) for (int q = 0; q < 5; ++q) out->XKr[k * 5 + q] = g.Kr[q];
        if (out->XKv14) for (int q = 0; q < 10; ++q) out->XKv14[k * 10 + q] = g.Kv14[q];
        if (out->XKv43) for (int q = 0; q < 10; ++q) out->XKv43[k * 10 + q] = g.Kv43[q];
        gw_cru_state_t c;
        c.CaSS = out->CaSS ? out->CaSS + k * n * 4 : nullptr;
        c.CaJSR = out->CaJSR ? out->CaJSR + k * n : nullptr;
        c.LCC = out->LCC ? out->LCC + k * n * 4 : nullptr;
        c.LCC_inactivation = out->LCC_inactivation ? out->LCC_inactivation + k * n * 4 : nullptr;
        c.ClCh = out->ClCh ? out->ClCh + k * n * 4 : nullptr;
        c.RyR = out->RyR ? out->RyR + k * n * 24 : nullptr;
        c.JLCC = nullptr;
        return gw_batch_get_cru_state(B, 0, &c, err, errlen);
    };
    return run_loop(ctx, p, n_cru, dt, num_steps, istim, record_every, g0, c0, seed, prng_id, precision,
                    out ? out->num_records : 0, record, err, errlen, noalloc);
}

}  // namespace

extern "C" {

int gw_run(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt, int64_t num_steps, const double *istim,
           int32_t record_every, const gw_globals_t *g0, const gw_cru_state_t *c0, uint64_t seed, int32_t prng_id,
           int32_t precision, gw_record_t *out, char *err, size_t errlen)
{
    return run_full(ctx, p, n_cru, dt, num_steps, istim, record_every, g0, c0, seed, prng_id, precision, out, err,
                    errlen, false);
}

int gw_run_noalloc(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt, int64_t num_steps,
                   const double *istim, int32_t record_every, const gw_globals_t *g0, const gw_cru_state_t *c0,
                   uint64_t seed, int32_t prng_id, int32_t precision, gw_record_t *out, char *err, size_t errlen)
{
    return run_full(ctx, p, n_cru, dt, num_steps, istim, record_every, g0, c0, seed, prng_id, precision, out, err,
                    errlen, true);
}

int gw_run_reduced(gw_ctx *ctx, const gw_params_t *p, int32_t n_cru, double dt, int64_t num_steps,
                   const double *istim, int32_t record_every, const gw_globals_t *g0, const gw_cru_state_t *c0,
                   uint64_t seed, int32_t prng_id, int32_t precision, gw_record_reduced_t *out, char *err,
                   size_t errlen)
{
    const int64_t n = n_cru;
    auto record = [&](gw_batch *B, int64_t k, double t) {
        gw_globals_t g;
        int rc = gw_batch_get_globals(B, &g, err, errlen);
        if (rc) return rc;
#define PUT(f, v) if (out->f) out->f[k] = static_cast<float>(v)
        PUT(t, t); PUT(V, g.V); PUT(m, g.m); PUT(h, g.h); PUT(j, g.j); PUT(Nai, g.Nai); PUT(Ki, g.Ki);
        PUT(Cai, g.Cai); PUT(CaNSR, g.CaNSR); PUT(CaLTRPN, g.CaLTRPN); PUT(CaHTRPN, g.CaHTRPN); PUT(xKs, g.xKs);
#undef PUT
        if (out->XKr) for (int q = 0; q < 4; ++q) out->XKr[k * 4 + q] = static_cast<float>(g.Kr[q]);
        if (out->XKv14) for (int q = 0; q < 9; ++q) out->XKv14[k * 9 + q] = static_cast<float>(g.Kv14[q]);
        if (out->XKv43) for (int q = 0; q < 9; ++q) out->XKv43[k * 9 + q] = static_cast<float>(g.Kv43[q]);
        gw_cru_state_reduced_t c;
        c.CaSS = out->CaSS ? out->CaSS + k * n * 4 : nullptr;
        c.CaJSR = out->CaJSR ? out->CaJSR + k * n : nullptr;
        c.LCC = out->LCC ? out->LCC + k * n * 4 : nullptr;
        c.LCC_inactivation = out->LCC_inactivation ? out->LCC_inactivation + k * n * 4 : nullptr;
        c.ClCh = out->ClCh ? out->ClCh + k * n * 4 : nullptr;
        c.RyR = out->RyR ? out->RyR + k * n * 4 : nullptr;
        return gw_batch_get_cru_state_reduced(B, 0, &c, err, errlen);
    };
    return run_loop(ctx, p, n_cru, dt, num_steps, istim, record_every, g0, c0, seed, prng_id, precision,
                    out ? out->num_records : 0, record, err, errlen);
}

int gw_run_batch(gw_ctx *ctx, int32_t ncells, const gw_cell_desc_t *cells, int32_t precision, double dt, int64_t nsteps,
                 gw_globals_t *g_out, gw_cell_obs_t *obs_out, char *err, size_t errlen)
{
    gw_batch *B = nullptr;
    int rc = gw_batch_create(ctx, ncells, cells, precision, GW_PRNG_XOSHIRO256SS, dt, nullptr, &B, err, errlen);
    if (rc) return rc;
    gw_batch_set_mode(B, 0, 0);
    rc = gw_batch_step(B, nsteps, err, errlen);
    if (!rc && g_out) rc = gw_batch_get_globals(B, g_out, err, errlen);
    if (!rc && obs_out) rc = gw_batch_get_obs(B, obs_out, err, errlen);
    gw_batch_destroy(B);
    return rc;
}

}  // extern "C"
