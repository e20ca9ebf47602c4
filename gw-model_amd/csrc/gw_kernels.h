// gw_kernels.h -- device data layout shared by the kernels and the host side.
#pragma once

#include "gw_ssa.h"

namespace gw {

constexpr int kBlock = 256;        // threads per block of the setup kernels (64 CRUs)
constexpr int kWave = 64;
constexpr int kCruPerBlock = kBlock / 4;  // one quad of lanes per CRU
constexpr int64_t kNoStop = 0x7fffffffffffffffLL;   // CellDyn::stop_step of a cell that has not failed

// Block -> cell maps of the SSA launch shapes: map s gives blocks of 16 << s
// CRUs (16..2048). A quad kernel with T threads uses map log2(T / 64); a lane
// kernel with T threads and C CRUs per lane uses map log2(T * C / 16). Blocks
// stop at the cell's last CRU, so the maps need no padding; every cell's CRU
// slots start at a multiple of kSlotAlign for aligned, coalesced access. The
// setup kernels use map 2 (256 threads, 64 CRUs).
constexpr int kNumShapes = 8;
constexpr int kShapeDefault = 2;
constexpr int kSlotAlign = 128;
__host__ __device__ constexpr int crus_per_block(int s) { return 16 << s; }
__host__ __device__ constexpr int shape_threads(int s) { return 64 << s; }   // quad kernels
__host__ __device__ constexpr int shape_of_block(int threads) { return threads == 64 ? 0 : threads == 128 ? 1 : threads == 256 ? 2 : 3; }
__host__ __device__ constexpr int shape_of_crus(int crus) { return crus <= 16 ? 0 : 1 + shape_of_crus(crus / 2); }

// Static description of one cell on the device.
struct DevCell {
    gw_params_t p;
    CellConsts c;
    SsaParams sp;
    SsaParamsF spf;        // the same, rounded to FP32 (FP32-local batches)
    gw_protocol_t proto;
    const double *istim;   // per-step stimulus table (GW_PROTO_TABLE) or null
    int64_t istim_len;
    int64_t slot0;         // first CRU slot of this cell in the state arrays
    int64_t blk0[kNumShapes];  // first SSA block of this cell, per launch shape
    uint64_t seed;
    int32_t n_cru;
    int32_t nblk[kNumShapes];  // SSA blocks of this cell, per launch shape
};

// Dynamic per-cell state, advanced by the cell's close block of each step.
struct CellDyn {
    gw_globals_t g;      // globals at the start of the current step
    gw_globals_t dg;     // increments formed before the current step's SSA
    StepConsts k[2];     // SSA constants of steps with even / odd index
    StepConstsF kf[2];   // the same, rounded to FP32 (FP32-local batches)
    double t;            // accumulated time at the start of the current step
    double Istim;        // stimulus of the current step
    double sums[6];      // JLCC, Jxfer, Jtr, ClCh, open RyR, Jrel of the last SSA
    int64_t step;        // index of the current step
    int64_t trace_rows;
    int64_t stop_step;   // SSA steps from this index on are not run (set when the cell fails; kNoStop)
    int32_t status;
    int32_t clamp;
    gw_cell_obs_t obs;
};

// Block partial of the per-step reductions.
struct Partial {
    Fx jlcc, jxfer, jtr, jrel;
    int64_t clch, open;
    uint64_t iters;
    int32_t err;
    int32_t tag;   // the persistent kernel's step tag (step index + 1); 0 elsewhere
};

// Per-cell exchange of the persistent step kernel: a counter, then the SSA
// constants of the latest steps the cell's close formed and the cell's stop
// step. It is written and read only with device-coherent atomic
// accesses, so no cache writeback or invalidation is needed between blocks
// on different XCDs.
constexpr int kXchgWords = 64;   // 512 bytes: StepConsts (59 words) or StepConstsF (31), then stop_step
struct CellSync {
    unsigned long long kready;       // the SSA constants of every step below this index are formed
    unsigned long long _pad[7];
    unsigned long long k[2][kXchgWords];   // constants of steps by parity; word kXchgWords - 1 holds stop_step
};

struct KArgs {
    const DevCell *cells;
    CellDyn *dyn;
    const int32_t *block_cell;   // cell of each SSA block of the current shape
    double *CaSS;      // [ntot][4]   (both forms: Ca accumulates in FP64)
    double *CaJSR;     // [ntot]
    double *JLCC;      // [ntot][4]   (FP64 batches)
    float *JLCCf;      // [ntot][4]   (FP32-local batches)
    uint64_t *sub;     // [ntot][4] packed subunit words (RyR counts, LCC, y, ClCh)
    uint64_t *rng;     // [ntot][4] engine words
    uint64_t *dig;     // [ntot][2] (count, hash) when digests are kept
    Partial *partials; // [2][nblocks], by step parity
    double *trace;     // [cap][ncells][4]
    unsigned long long *iters_total;
    int64_t ntot;
    int64_t nblocks;     // SSA blocks of the current shape
    int64_t trace_cap;
    double dt;
    int32_t ncells;
    int32_t frozen;
    int32_t trace_every;
    int32_t prng;
    int32_t precision;   // GW_PREC_FP64 or GW_PREC_FP32_LOCAL
    int32_t shape;       // current SSA launch shape (0..3)
    int32_t _pad;
    CellSync *sync;      // [ncells] (persistent step kernel)
    int32_t *timeout;    // set when a bounded wait of the persistent kernel expires
    int64_t spin_limit;  // polls before a wait of the persistent kernel gives up
};

// CRU-local state of either precision, for the kernels outside the SSA.
__device__ __forceinline__ double ld_cass(const KArgs &a, int64_t k) { return a.CaSS[k]; }
__device__ __forceinline__ double ld_jlcc(const KArgs &a, int64_t k) { return a.JLCCf ? static_cast<double>(a.JLCCf[k]) : a.JLCC[k]; }
__device__ __forceinline__ double ld_cajsr(const KArgs &a, int64_t k) { return a.CaJSR[k]; }
__device__ __forceinline__ void st_cass(const KArgs &a, int64_t k, double v) { a.CaSS[k] = v; }
__device__ __forceinline__ void st_jlcc(const KArgs &a, int64_t k, double v)
{
    if (a.JLCCf) a.JLCCf[k] = static_cast<float>(v); else a.JLCC[k] = v;
}
__device__ __forceinline__ void st_cajsr(const KArgs &a, int64_t k, double v) { a.CaJSR[k] = v; }

}  // namespace gw
