#!/usr/bin/env python3
"""Throughput of the Greenstein-Winslow coupled integrator on one MI355X.

Default workload: 1,024 cells x 12,500 CRUs (12.8 M CRUs), each cell with
its own seed and with PCaL, GKr, GKs, GK1, GKv43 and kNaCa scaled
log-uniform in [0.5, 2], paced at 2 Hz with a +100 pA/pF, 1 ms stimulus (the
reference's sign: positive is depolarising), dt = 1e-3 ms, from the seeded rest
state. One bench step is one Euler step of the whole batch (every cell: global
ODE + SSA of all its CRUs). The batch is first advanced untimed to the start of
the timed window (--window-start-ms, default 1 ms: the upstroke, so the window
holds the upstroke and the Ca release, the busiest part of the beat), or by
--warmup steps if that is further; then exactly K steps are timed with HIP
events on the library's stream. The timed window is stated in `config`. Both
forms run on the same batch: the batched FP32-local form (CRU-local arithmetic
in FP32, global ODE and flux sums in FP64) and the plain FP64 form (bit-exact
with the reference's algorithm).

  value        CRU*ms simulated per second over the whole batch, FP32-local
               form (the plain FP64 form is under "fp64")
  cell_beats   cells * (simulated ms / 500 ms per beat) per second
  roofline     for the step kernel: algorithmic flops per launch (F_ITER per
               SSA iteration x iterations counted in-kernel) and algorithmic
               state bytes per launch (the CRU state read and written once per
               launch: bytes per CRU x CRUs, whether the launch runs one SSA
               step or two), each divided by the kernel's average duration (the
               timed region between its two HIP events over its SSA launches,
               which run back to back); against the vector peak of the form's
               precision and HBM. On the default workload the FP32-local form
               runs two SSA steps per launch (--steps-per-launch).
               `traffic` is the measured HBM bytes per launch from
               profiles/k_ssa_traffic.txt (rocprofv3 FETCH_SIZE and WRITE_SIZE,
               calibrated on a copy of known size) when present.
  cpu_baseline the CPU restatement built with the reference's release flags,
               OpenMP over CRUs on all host cores, on a bounded sample: the
               batch's first two cells (their parameters and seeds), one after
               another, over the same timed window; a one-thread sample of a
               smaller cell over the same window beside it (per-core
               throughput), and the host CPU model
  launch       the SSA launch shape: by default the autotuner's choice on
               this batch (1x64x4 FP32-local, 4x64 FP64), fixed so that a
               profile of the same command sees the same kernels;
               --autotune STEPS runs the tuner instead
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gw-model_amd")

# FP64 operations of one SSA loop iteration, counted on the oracle's
# SSA_single_CRU (add, sub, mul, div, int->fp each 1, log 1) for the iteration
# that closes a step at rest: fluxes 26, rates 4 x 56, total + waiting time 7,
# dt - t 1, CaSS 4 x 15, CaJSR 13.
F_ITER = 331
STATE_BYTES = 272       # bytes read + written per CRU-step by the FP64 SSA kernel: 136 in, 136 out
STATE_BYTES_F32 = 240   # the same for the FP32-local kernel: 120 in, 120 out (JLCC in FP32)
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def load_capi():
    spec = importlib.util.spec_from_file_location("gw_capi_binding", os.path.join(PKG, "capi.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def pacing(capi, period=1000.0):
    return {"kind": capi.PROTO_PACE, "period": period, "duration": 1.0, "amplitude": 100.0, "start": 0.0}


WORKLOADS = {
    "seeds64": "independent seeds of one cell type (default 64 x 12,500 CRUs), 1 Hz pacing",
    "batch256": "256 cells, CRU counts uniform in 2,000..12,500, PCaL/GKr/GKs scaled log-uniform in [0.5, 2], "
                "per-cell seeds, 1 Hz pacing",
    "large1024": "1,024 cells x 12,500 CRUs, PCaL/GKr/GKs/GK1/GKv43/kNaCa scaled log-uniform in [0.5, 2], "
                 "per-cell seeds, 2 Hz pacing",
}


def make_cells(capi, workload, cells, crus):
    """The cell descriptors of a workload; returns (cells, pacing period in ms)."""
    rng = np.random.default_rng(2024)

    def scaled(names):
        p = capi.Params.default()
        for name in names:
            setattr(p, name, getattr(p, name) * float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))))
        return p

    if workload == "seeds64":
        return [capi.make_cell(crus, 42 + c, None, pacing(capi)) for c in range(cells)], 1000.0
    if workload == "batch256":
        out = []
        for c in range(256):
            n = int(rng.integers(2000, 12501))
            out.append(capi.make_cell(n, 1000 + c, scaled(("PCaL", "GKr", "GKs")), pacing(capi)))
        return out, 1000.0
    out = [capi.make_cell(12500, 5000 + c, scaled(("PCaL", "GKr", "GKs", "GK1", "GKv43", "kNaCa")), pacing(capi, 500.0))
           for c in range(1024)]
    return out, 500.0


def gpu_bench(capi, cell_list, steps, warmup, precision=0, autotune=200, launch=None, steps_per_launch=1):
    """Time `steps` Euler steps of the batch after `warmup` untimed ones (steps_per_launch: 2 runs two
    SSA steps per launch where the launch shape supports it)."""
    ctx = capi.Context(0)
    tune = None
    if launch is not None:   # a fixed launch shape instead of the autotuner's choice
        tune = {"lanes_per_cru": launch[0], "block": launch[1], "crus_per_lane": launch[2],
                "crus_per_block": launch[1] * launch[2] // launch[0], "graph_steps": 0, "fixed": True}
    elif autotune > 0:
        # Tune on the state the timed region spends its time in: a copy of the
        # batch advanced to the middle of that region (the autotuner restores
        # the state it starts from after its trials).
        t = capi.Batch(ctx, cell_list, precision=precision)
        t.set_mode(frozen=False, digest=False)
        t.set_steps_per_launch(steps_per_launch)
        t.step(warmup + steps // 2)
        (lanes, blk, cpl), trials = t.autotune(autotune)
        tune = {"lanes_per_cru": lanes, "block": blk, "crus_per_lane": cpl, "crus_per_block": blk * cpl // lanes,
                "graph_steps": t.graph_steps, "tuned_steps_per_launch": t.steps_per_launch(),
                "trial_steps": autotune, "tuned_at_step": warmup + steps // 2, "trial_ms_per_step": trials}
        t.close()

    def fresh():
        x = capi.Batch(ctx, cell_list, precision=precision)
        x.set_mode(frozen=False, digest=False)
        x.set_steps_per_launch(tune.get("tuned_steps_per_launch", steps_per_launch) if tune else steps_per_launch)
        if tune is not None:
            x.set_launch(tune["lanes_per_cru"], tune["block"], tune["crus_per_lane"])
            x.set_graph(tune["graph_steps"])
        if warmup > 0:
            x.step(warmup)
        return x

    # The timed region runs the steps as a caller would (plain launches or
    # replayed graphs) between two HIP events on the batch's stream. With one
    # step per launch the step launches run back to back (no host gap: the
    # kernel trace of the same command shows them so), so the region's time
    # over its SSA launches is the step kernel's average duration, and no
    # per-dispatch events are used (each makes its kernel's completion write
    # back the caches, about 3 us per kernel). With two steps per launch a
    # close kernel runs between the SSA launches, so each SSA dispatch also
    # records its own events and the kernel's average comes from those.
    b = fresh()
    spl = b.steps_per_launch()
    i0 = b.iterations()
    ms, ssa, launches = b.time_steps(steps, per_kernel=spl > 1)
    iters = b.iterations() - i0
    ssa_ms = ssa if spl > 1 else ms
    if tune is not None:
        tune["steps_per_launch"] = spl
        tune["kernel_timing"] = (f"HIP events around the {steps} timed steps ({launches} SSA launches and one "
                                 "close-only launch)" if spl == 1 else
                                 f"HIP events of each of the {launches} SSA launches of 2 steps (the close "
                                 "kernels between them excluded); ms_per_step from HIP events around the "
                                 f"{steps} timed steps")
    g = b.get_globals()
    obs = b.obs()
    b.close()
    ctx.close()
    return ms, ssa_ms, launches, iters, g, obs, tune


def kernel_name(tune, f32):
    """The step kernel a launch shape runs (xoshiro256** instantiation)."""
    lanes, blk, cpl = (tune["lanes_per_cru"], tune["block"], tune["crus_per_lane"]) if tune else (4, 256, 1)
    if lanes == 1:
        if tune and tune.get("steps_per_launch", 1) == 2:
            return f"k_step_lane2<0, {blk}, {cpl}>"
        return f"k_step_lane<0, {blk}, {cpl}>"
    return f"k_step<0, false, {'float' if f32 else 'double'}, {blk}>"


def measured_traffic(total_crus, form):
    """HBM bytes per SSA launch from the committed PMC summary, scaled to this batch
    (form: fp64, fp32, or fp32_pair for two steps per launch)."""
    path = os.path.join(ROOT, "profiles", "k_ssa_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path, encoding="utf-8"):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        key = "hbm_bytes_per_cru_launch_fp32_pair" if form == "fp32_pair" else f"hbm_bytes_per_cru_step_{form}"
        return float(vals[key]) * total_crus
    except (KeyError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(crus, steps, warmup_steps=0, seed=42, threads=None, params=None, period=1000.0):
    """CRU*ms/s of the CPU restatement (OpenMP over CRUs) on one cell of `crus`
    CRUs over `steps` steps after `warmup_steps` untimed ones from t = 0 (paced
    with +100 pA/pF for 1 ms every `period` ms); `threads` caps the OpenMP team
    (None: every host core); `params` is the cell's parameter set (a
    gw_params_t of the library, default: the reference defaults)."""
    sys.path.insert(0, ROOT)
    from oracle.oracle import GwParams, OracleModel, lib

    L = lib("omp")
    gomp = C.CDLL("libgomp.so.1")
    all_threads = L.gwo_model_threads()
    gomp.omp_set_num_threads(int(threads or all_threads))
    threads = L.gwo_model_threads()
    p = GwParams.from_buffer_copy(bytes(params)) if params is not None else None
    m = OracleModel(crus, params=p, seed=seed, kind="omp")
    m.sample_rest_state()
    ist = np.zeros(steps + warmup_steps)
    t = 0.0
    for k in range(len(ist)):
        ist[k] = 100.0 if (t % period) < 1.0 else 0.0
        t += 1e-3
    if warmup_steps:
        m.run(1e-3, warmup_steps, ist[:warmup_steps])
    t0 = time.perf_counter()
    rc = L.gwo_model_run(m.h, 1e-3, steps, ist[warmup_steps:].ctypes.data_as(C.c_void_p), 0, None)
    dt = time.perf_counter() - t0
    gomp.omp_set_num_threads(int(all_threads))
    if rc != 0:
        raise RuntimeError(f"oracle run failed: {rc}")
    return crus * steps * 1e-3 / dt, threads, dt


def copy_calibration(nbytes, reps=5):
    """Device-to-device copies of nbytes (the runtime's copy kernel), for scaling
    the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def alternans_run(capi, args):
    """The FP32-local batch from rest through --alternans-beats whole beats with the
    on-device APD90 tracker: per-cell APD90 of the last two beats and the
    beat-to-beat APD90 differences, and the whole-run throughput. Prints one
    JSON line (rank 0); progress goes to stderr every 50 simulated ms."""
    cell_list, period = make_cells(capi, args.workload, args.cells, args.crus)
    crus = sum(c.n_cru for c in cell_list)
    ctx = capi.Context(0)
    b = capi.Batch(ctx, cell_list, precision=capi.PREC_FP32_LOCAL)
    b.set_mode(frozen=False, digest=False)
    shape = tuple(int(x) for x in (args.launch_fp32 + "x1").split("x")[:3])
    b.set_launch(*shape)
    b.set_steps_per_launch(args.steps_per_launch)
    total = int(round(args.alternans_beats * period / 1e-3))
    chunk = 50_000
    t0 = time.perf_counter()
    done = 0
    while done < total:
        n = min(chunk, total - done)
        b.step(n)
        done += n
        print(f"bench: {done * 1e-3:.0f} of {total * 1e-3:.0f} simulated ms", file=sys.stderr, flush=True)
    sec = time.perf_counter() - t0
    obs = b.obs()
    b.close()
    ctx.close()
    beats = np.array([o["beats_done"] for o in obs])
    last = np.array([o["apd90_last"] for o in obs])
    prev = np.array([o["apd90_prev"] for o in obs])
    both = (last > 0) & (prev > 0)
    alt = np.abs(last - prev)[both]
    mean_abs = np.array([o["alt_sum"] / o["alt_count"] for o in obs if o["alt_count"] > 0])

    def q(x):
        if len(x) == 0:
            return None
        p = np.percentile(x, [0, 5, 50, 95, 100])
        return {"min": p[0], "p5": p[1], "median": p[2], "p95": p[3], "max": p[4], "mean": float(np.mean(x))}

    return {
        "metric": "per-cell beat-to-beat APD90 alternans, FP32-local, 1 GPU",
        "config": {"workload": WORKLOADS[args.workload] + f", dt 1e-3 ms, {args.alternans_beats} beats from rest",
                   "name": args.workload, "cells": len(cell_list), "crus": crus, "beat_ms": period,
                   "beats": args.alternans_beats, "steps": total,
                   "launch": {"lanes_per_cru": shape[0], "block": shape[1], "crus_per_lane": shape[2],
                              "steps_per_launch": args.steps_per_launch}},
        "wall_s": sec,
        "cru_ms_per_s": crus * total * 1e-3 / sec,
        "cell_beats_per_s": len(cell_list) * args.alternans_beats / sec,
        "beats_done": {str(int(k)): int(np.sum(beats == k)) for k in np.unique(beats)},
        "apd90_last_ms": q(last[last > 0]),
        "apd90_prev_ms": q(prev[prev > 0]),
        "alternans_last_two_beats_ms": q(alt),
        "mean_abs_dapd90_ms": q(mean_abs),
        "cells_alternans_over_5ms": int(np.sum(alt > 5.0)),
        "data": "synthetic: seeded rest state, paced from rest",
    }


def ensemble_run(capi, args):
    """The seed ensemble in both forms: every cell from rest through --ensemble-beats
    whole beats, the on-device observables reset at the start of the last beat, and
    the last beat's APD90, peak Cai, peak I_CaL and mean RyR open fraction compared
    seed by seed (mean paired difference with its 95% interval, relative to the FP64
    mean) and as distributions (variance ratio, two-sample KS). One JSON line."""
    from scipy import stats

    cell_list, period = make_cells(capi, args.workload, args.cells, args.crus)
    beat_steps = int(round(period / 1e-3))
    shapes = {"fp64": tuple(int(x) for x in (args.launch_fp64 + "x1").split("x")[:3]),
              "fp32": tuple(int(x) for x in (args.launch_fp32 + "x1").split("x")[:3])}
    res, walls = {}, {}
    for form, prec in (("fp64", capi.PREC_FP64), ("fp32", capi.PREC_FP32_LOCAL)):
        ctx = capi.Context(0)
        b = capi.Batch(ctx, cell_list, precision=prec)
        b.set_mode(frozen=False, digest=False)
        b.set_launch(*shapes[form])
        b.set_steps_per_launch(args.steps_per_launch if form == "fp32" else 1)
        t0 = time.perf_counter()
        for beat in range(args.ensemble_beats):
            if beat == args.ensemble_beats - 1:
                b.reset_obs()
            for k in range(0, beat_steps, 50_000):
                b.step(min(50_000, beat_steps - k))
                print(f"bench: {form} beat {beat + 1}, {(k + 50_000) * 1e-3:.0f} ms", file=sys.stderr, flush=True)
        walls[form] = time.perf_counter() - t0
        obs = b.obs()
        b.close()
        ctx.close()
        res[form] = {"APD90": np.array([o["apd90_last"] for o in obs]),
                     "peak_Cai": np.array([o["peak_Cai"] for o in obs]),
                     "peak_ICaL": np.array([o["peak_ICaL_neg"] for o in obs]),
                     "mean_open_RyR": np.array([o["sum_open_RyR_frac"] / o["n_steps"] for o in obs])}
    out = {}
    for k in res["fp64"]:
        a, b_ = res["fp32"][k], res["fp64"][k]
        ok = (a > -1) & (b_ > -1) if k == "APD90" else np.ones(len(a), bool)
        a, b_ = a[ok], b_[ok]
        d = a - b_
        se = d.std(ddof=1) / np.sqrt(len(d))
        ratio = a.var(ddof=1) / b_.var(ddof=1)
        out[k] = {"seeds": int(len(d)), "fp64_mean": float(b_.mean()), "fp32_mean": float(a.mean()),
                  "fp64_sd": float(b_.std(ddof=1)), "fp32_sd": float(a.std(ddof=1)),
                  "paired_diff": float(d.mean()), "paired_diff_ci95": float(1.96 * se),
                  "rel_error": float(d.mean() / b_.mean()), "rel_error_ci95": float(1.96 * se / abs(b_.mean())),
                  "paired_diff_in_se": float(d.mean() / se) if se > 0 else 0.0,
                  "variance_ratio": float(ratio),
                  "variance_ratio_accept_1e-3": [float(stats.f.ppf(5e-4, len(a) - 1, len(b_) - 1)),
                                                 float(stats.f.ppf(1 - 5e-4, len(a) - 1, len(b_) - 1))],
                  "ks_p": float(stats.ks_2samp(a, b_).pvalue)}
    crus = sum(c.n_cru for c in cell_list)
    return {"metric": "seed-ensemble statistics of the last beat, FP32-local against FP64, 1 GPU",
            "config": {"workload": (f"{len(cell_list)} independent seeds x {args.crus} CRUs, 1 Hz pacing"
                                    if args.workload == "seeds64" else WORKLOADS[args.workload])
                                   + f", dt 1e-3 ms, {args.ensemble_beats} beats from rest; observables of the last",
                       "name": args.workload, "cells": len(cell_list), "crus": crus, "beat_ms": period,
                       "beats": args.ensemble_beats, "launch": {k: list(v) for k, v in shapes.items()}},
            "wall_s": walls, "last_beat": out, "data": "synthetic: seeded rest state, paced from rest"}


def advance_steps(args):
    """Untimed steps before the timed window: to --window-start-ms, or --warmup if that is further."""
    return max(int(args.warmup), int(round(args.window_start_ms / 1e-3)))


def measure(capi, precision, args, total_crus):
    """One form (FP64 or FP32-local) on the bench batch: throughput and the SSA kernel's roofline."""
    f32 = precision == capi.PREC_FP32_LOCAL
    cell_list, period = make_cells(capi, args.workload, args.cells, args.crus)
    fixed = args.launch_fp32 if f32 else args.launch_fp64
    launch = tuple(int(x) for x in (fixed + "x1").split("x")[:3]) if fixed and args.autotune <= 0 else None
    ms, ssa_ms, launches, iters, g, obs, tune = gpu_bench(capi, cell_list, args.steps, advance_steps(args),
                                                          precision=precision, autotune=args.autotune, launch=launch,
                                                          steps_per_launch=args.steps_per_launch if f32 else 1)
    pair = bool(tune and tune.get("steps_per_launch", 1) == 2)
    sim_ms = args.steps * 1e-3
    sec = ms / 1e3
    per_launch_s = ssa_ms / 1e3 / launches
    flops_per_launch = F_ITER * iters / launches
    achieved_tflops = flops_per_launch / per_launch_s / 1e12
    bytes_per_launch = (STATE_BYTES_F32 if f32 else STATE_BYTES) * total_crus   # the state crosses HBM once per launch
    achieved_gbs = bytes_per_launch / per_launch_s / 1e9
    peak_f = PEAK_FP32_TFLOPS if f32 else PEAK_FP64_TFLOPS
    frac_f = achieved_tflops / peak_f
    frac_b = achieved_gbs / PEAK_HBM_GBS
    valu = frac_f >= frac_b
    return {
        "value": total_crus * sim_ms / sec,
        "ms_per_step": ms / args.steps,
        "cell_beats_per_s": len(cell_list) * (sim_ms / period) / sec,
        "iterations_per_cru_step": iters / (args.steps * total_crus),
        "mean_V_end": float(np.mean([x.V for x in g])),
        "launch": tune,
        "roofline": {
            "bound": "valu" if valu else "hbm",
            "achieved": achieved_tflops if valu else achieved_gbs,
            "peak": peak_f if valu else PEAK_HBM_GBS,
            "unit": "TFLOP/s" if valu else "GB/s",
            "frac": max(frac_f, frac_b),
            "traffic": measured_traffic(total_crus, ("fp32_pair" if pair else "fp32") if f32 else "fp64"),
            "kernel": kernel_name(tune, f32),
            "avg_launch_us": per_launch_s * 1e6,
            "valu": {"achieved_tflops": achieved_tflops, "peak": peak_f, "frac": frac_f,
                     "flops_per_launch": flops_per_launch, "f_iter": F_ITER},
            "hbm": {"achieved_gbs": achieved_gbs, "peak": PEAK_HBM_GBS, "frac": frac_b,
                    "bytes_per_launch": bytes_per_launch},
        },
    }


LATENCY_BATCHES = ((1, 64), (1, 1250), (1, 12500), (4, 12500), (16, 12500))


def latency_run(capi, args):
    """Per-step time of small batches (a cell of the reference's sizes and a
    few of them) in both forms at the default launch shape: seeded rest state,
    1 Hz pacing, 1,000 untimed steps, then gw_batch_time_steps over
    --latency-steps steps."""
    ctx = capi.Context(0)
    rows = []
    for ncells, n in LATENCY_BATCHES:
        row = {"cells": ncells, "crus": n}
        for name, prec in (("fp64", capi.PREC_FP64), ("fp32_local", capi.PREC_FP32_LOCAL)):
            b = capi.Batch(ctx, [capi.make_cell(n, 42 + c, None, pacing(capi)) for c in range(ncells)], precision=prec)
            b.set_mode(False, False)
            b.step(1000)
            ms, _, _ = b.time_steps(args.latency_steps)
            b.close()
            row[f"{name}_us_per_step"] = ms / args.latency_steps * 1e3
        rows.append(row)
        print(f"bench: {ncells} x {n} CRUs: FP64 {row['fp64_us_per_step']:.1f} us, "
              f"FP32-local {row['fp32_local_us_per_step']:.1f} us per step", file=sys.stderr, flush=True)
    ctx.close()
    return {
        "metric": "time per Euler step of small batches, 1 GPU",
        "unit": "us/step",
        "higher_is_better": False,
        "data": "synthetic: seeded rest state, reference default parameters, 1 Hz pacing",
        "config": {"workload": "1 x 64, 1 x 1,250, 1 x 12,500, 4 x 12,500 and 16 x 12,500 CRUs, dt 1e-3 ms, "
                               f"{args.latency_steps} steps timed after 1,000 untimed, default launch shape",
                   "steps": args.latency_steps},
        "batches": rows,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000, help="timed Euler steps (dt = 1e-3 ms)")
    ap.add_argument("--warmup", type=int, default=1000,
                    help="untimed Euler steps before the timed region (at least up to --window-start-ms)")
    ap.add_argument("--window-start-ms", type=float, default=1.0,
                    help="simulated time at which the timed window starts (default 1 ms: upstroke and release)")
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--crus", type=int, default=12500)
    ap.add_argument("--cpu-crus", type=int, default=12500)
    ap.add_argument("--cpu-cells", type=int, default=2,
                    help="cells of the batch the CPU sample runs, one after another (their own parameters and seeds)")
    ap.add_argument("--cpu-steps", type=int, default=0, help="steps of the CPU sample (0: the timed window's)")
    ap.add_argument("--cpu-crus-1t", type=int, default=1500, help="CRUs of the one-thread CPU sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="large1024",
                    help="large1024 (default): 1,024 cells x 12,500 CRUs at 2 Hz; batch256: 256 uneven cells; "
                         "seeds64: --cells independent seeds of one cell")
    ap.add_argument("--autotune", type=int, default=0, metavar="STEPS",
                    help="run the autotuner with STEPS trial steps per launch shape, on a copy of the batch advanced "
                         "to the middle of the timed window, and time its choice instead of the fixed shapes")
    ap.add_argument("--launch-fp32", default="1x64x4", metavar="LxB[xC]",
                    help="FP32-local launch shape (lanes per CRU x threads x CRUs per lane); the default is the "
                         "autotuner's choice on the seeds64 batch")
    ap.add_argument("--launch-fp64", default="4x64", metavar="LxB",
                    help="FP64 launch shape; the default is the autotuner's choice on the seeds64 batch")
    ap.add_argument("--steps-per-launch", type=int, choices=[0, 1, 2], default=0,
                    help="SSA steps per launch of the FP32-local form (2: each CRU advances two steps per launch "
                         "with its state in registers, one-lane shapes of 64 threads; 0, the default: the "
                         "library's choice, two when a launch's blocks are more than the GPU holds at once)")
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both",
                    help="forms to time; with both, value is the batched FP32-local form and fp64 sits beside it")
    ap.add_argument("--ensemble-beats", type=int, default=0, metavar="N",
                    help="only run the batch in both forms from rest through N whole beats and compare the last "
                         "beat's observables seed by seed (one JSON line), instead of the throughput bench")
    ap.add_argument("--alternans-beats", type=int, default=0, metavar="N",
                    help="only run the FP32-local batch from rest through N whole beats and report the per-cell "
                         "APD90 alternans from the on-device tracker (one JSON line), instead of the throughput bench")
    ap.add_argument("--latency", action="store_true",
                    help="only time one step of small batches (1 to 16 cells) in both forms (one JSON line), "
                         "instead of the throughput bench")
    ap.add_argument("--latency-steps", type=int, default=5000, help="timed steps per batch of --latency")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (counter calibration) and exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (cells are independent: "
              "several GPUs would each take a range of whole cells, with no exchange during the run)",
              file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))
    capi = load_capi()
    if args.ensemble_beats > 0:
        out = ensemble_run(capi, args)
        if rank == 0:
            print(json.dumps(out), flush=True)
        return 0
    if args.alternans_beats > 0:
        out = alternans_run(capi, args)
        if rank == 0:
            print(json.dumps(out), flush=True)
        return 0
    if args.latency:
        out = latency_run(capi, args)
        if rank == 0:
            print(json.dumps(out), flush=True)
        return 0
    cell_list, period = make_cells(capi, args.workload, args.cells, args.crus)
    total_crus = sum(c.n_cru for c in cell_list)
    forms = {"fp32": capi.PREC_FP32_LOCAL, "fp64": capi.PREC_FP64}
    chosen = ["fp32", "fp64"] if args.precision == "both" else [args.precision]
    res = {}
    for k in chosen:
        res[k] = measure(capi, forms[k], args, total_crus)
        print(f"bench: {k} form timed, {res[k]['ms_per_step']:.4f} ms per step", file=sys.stderr, flush=True)
    head = res[chosen[0]]
    t0_ms = advance_steps(args) * 1e-3
    t1_ms = t0_ms + args.steps * 1e-3
    line = {
        "metric": "CRU\u00b7ms simulated/s and cell-beats/s, batched FP32-local vs plain FP64, 1 GPU",
        "value": head["value"],
        "unit": "CRU\u00b7ms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-local/f64-global" if chosen[0] == "fp32" else "f64",
        "data": "synthetic: seeded rest state, paced from rest",
        "config": {"workload": (f"{args.cells} cells (independent seeds) x {args.crus} CRUs, 1 Hz pacing"
                                if args.workload == "seeds64" else WORKLOADS[args.workload])
                               + f", dt 1e-3 ms, timed window t = {t0_ms:g}..{t1_ms:g} ms of beat 1; "
                               + ("batched FP32-local, plain FP64 beside it" if len(chosen) == 2 else chosen[0]),
                   "name": args.workload, "cells": len(cell_list), "crus": total_crus, "dt_ms": 1e-3,
                   "window_start_ms": t0_ms, "window_end_ms": t1_ms, "beat_ms": period,
                   "untimed_steps_before_window": advance_steps(args), "precision": chosen[0]},
        "cell_beats_per_s": head["cell_beats_per_s"],
        "iterations_per_cru_step": head["iterations_per_cru_step"],
        "launch": head["launch"],
        "roofline": head["roofline"],
    }
    for k in chosen[1:]:
        line[k] = {x: res[k][x] for x in ("value", "ms_per_step", "cell_beats_per_s", "iterations_per_cru_step",
                                          "launch", "roofline")}
    if len(chosen) == 2:
        line["speedup_fp32_local_over_fp64"] = res["fp32"]["value"] / res["fp64"]["value"]
    value = head["value"]
    if not args.no_cpu and rank == 0:
        # the batch's first cells (their parameters, seeds and pacing) over the same window
        sample = cell_list[:max(1, args.cpu_cells)]
        cpu_steps = args.cpu_steps or args.steps
        adv = advance_steps(args)
        secs, threads = 0.0, 0
        for c in sample:
            _, threads, dt_s = cpu_baseline(args.cpu_crus, cpu_steps, adv, seed=c.seed, params=c.params, period=period)
            secs += dt_s
        v = len(sample) * args.cpu_crus * cpu_steps * 1e-3 / secs
        c0 = sample[0]
        v1, _, secs1 = cpu_baseline(args.cpu_crus_1t, cpu_steps, adv, seed=c0.seed, threads=1, params=c0.params,
                                    period=period)
        window = f"t = {adv * 1e-3:g}..{(adv + cpu_steps) * 1e-3:g} ms of beat 1 after {adv} untimed steps"
        line["cpu_baseline"] = {"value": v, "unit": "CRU\u00b7ms/s", "cores": threads, "kind": "port",
                                "sample": f"the batch's first {len(sample)} cell{'s' if len(sample) > 1 else ''}, "
                                          f"one after another: {len(sample)} x {args.cpu_crus} CRUs x {cpu_steps} "
                                          f"steps ({window}), {secs:.1f} s",
                                "cpu_model": cpu_model(),
                                "one_thread": {"value": v1, "sample": f"1 cell x {args.cpu_crus_1t} CRUs x "
                                                                     f"{cpu_steps} steps ({window}), {secs1:.1f} s"},
                                "ratio_gpu_over_cpu": value / v,
                                "ratio_fp64_gpu_over_cpu": res["fp64"]["value"] / v if "fp64" in res else None}
    if rank == 0:
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
