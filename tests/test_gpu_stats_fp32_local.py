"""The FP32-local variant against the FP64 form.

CRU-local Ca, fluxes and propensities run in FP32; channel states, random
streams, flux sums (exact, FP64) and the global ODE (FP64) are those of the
FP64 form. The FP64 form is bit-identical to the CPU oracle (test_gpu_parity.py),
so it stands for the reference here. The variant is not bit-exact, so its
results are compared as ensembles: per-cell observables over independent seeds,
means within 3 standard errors.
"""
import math

import numpy as np
import pytest

from .test_gpu_parity import say

pytestmark = pytest.mark.gpu

PACE = {"kind": 1, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}


def _run(gwcapi, ctx, precision, cells, steps, launch=None):
    b = gwcapi.Batch(ctx, cells, precision=precision)
    b.set_mode(frozen=False, digest=False)
    if launch:
        b.set_launch(*launch)
        assert b.launch() == launch
    b.step(steps)
    obs = b.obs()
    g = b.get_globals()
    state0 = b.get_cru_state(0)
    b.close()
    return obs, g, state0


def _stats(obs):
    apd = np.array([o["apd90_last"] for o in obs])
    cai = np.array([o["peak_Cai"] for o in obs])
    ical = np.array([o["peak_ICaL_neg"] for o in obs])
    ryr = np.array([o["sum_open_RyR_frac"] / o["n_steps"] for o in obs])
    return {"APD90": apd, "peak_Cai": cai, "peak_ICaL": ical, "mean_open_RyR": ryr}


def _check_within_3se(a, b, name):
    se = math.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    diff = a.mean() - b.mean()
    say(f"  {name}: fp64 {b.mean():.6g} fp32 {a.mean():.6g} diff {diff:.3g} se {se:.3g} "
        f"rel {diff / b.mean():.2e}")
    assert abs(diff) <= 3 * se + 1e-12 * abs(b.mean()), name
    return diff / b.mean()


_FP64_ENSEMBLE = {}


@pytest.mark.parametrize("launch", [(4, 256, 1), (1, 64, 1), (1, 64, 4)], ids=["quad", "lane", "lane-refill"])
def test_fp32_local_free_mode_ensemble_matches_fp64(gwcapi, gpu_ctx, launch):
    """64 seeds, one paced beat (600 ms: past the longest APD90 of the FP64
    form, about 420 ms over 256 seeds, and of the FP32-local form's longer
    tail, about 500 ms, so no cell's beat is cut off): APD90, peak Cai, peak
    ICaL and mean RyR open fraction agree within 3 SE, for a quad of lanes per
    CRU, for one lane per CRU and for lanes refilled from four CRUs each."""
    seeds, n, steps = 64, 1000, 600_000
    cells = [gwcapi.make_cell(n, 500 + s, None, PACE) for s in range(seeds)]
    if "o64" not in _FP64_ENSEMBLE:
        say(f"fp64 ensemble {seeds} x {n} CRUs x {steps} steps")
        _FP64_ENSEMBLE["o64"] = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP64, cells, steps)[0]
    o64 = _FP64_ENSEMBLE["o64"]
    say(f"fp32-local ensemble, launch {launch}")
    o32, g32, s32 = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP32_LOCAL, cells, steps, launch)
    a64, a32 = _stats(o64), _stats(o32)
    for name, a, g in (("fp64", a64, None), ("fp32", a32, g32)):
        bad = np.flatnonzero(a["APD90"] <= 100)
        vb = [round(g[i].V, 2) for i in bad] if g is not None else []
        say(f"  {name}: cells without one full beat: {bad.tolist()} APD90 {a['APD90'][bad].tolist()} final V {vb}")
    assert np.all(a64["APD90"] > 100) and np.all(a32["APD90"] > 100)  # every cell repolarised once
    for k in a64:
        _check_within_3se(a32[k], a64[k], k)
    v32 = np.array([x.V for x in g32])
    say(f"  fp32 final V: min {v32.min():.2f} max {v32.max():.2f} mV")
    # every cell repolarised (APD90 above) and none has blown up since
    assert np.all(np.isfinite(v32)) and np.all(v32 < -60.0)


@pytest.mark.parametrize("launch", [(4, 256, 1), (1, 64, 1), (1, 256, 1), (1, 128, 2), (1, 64, 4), (1, 256, 8)],
                         ids=["quad", "lane64", "lane256", "lane128x2", "lane64x4", "lane256x8"])
def test_fp32_local_state_invariants(gwcapi, gpu_ctx, launch):
    """Channel states stay exact integers with conserved RyR counts; Ca stays finite and positive."""
    n = 4000
    say(f"fp32-local invariants, launch {launch}")
    _, g, s = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP32_LOCAL, [gwcapi.make_cell(n, 9, None, PACE)], 30_000, launch)
    R = s["RyR"]
    assert np.all(R.sum(axis=2) == 5)
    assert np.all(R[..., 2] + R[..., 3] >= 0) and np.all(R[..., 4] + R[..., 5] >= 0)
    assert np.all((s["LCC"] >= 1) & (s["LCC"] <= 12))
    open_lcc = ((s["LCC"] == 6) | (s["LCC"] == 12)) & (s["LCC_inactivation"] == 1)
    assert np.all(s["JLCC"][~open_lcc] == 0.0)
    assert np.all(np.isfinite(s["CaSS"])) and np.all(s["CaSS"] > 0) and np.all(s["CaJSR"] > 0)
    assert g[0].V > -40.0


def test_fp32_local_refilled_lanes_are_deterministic(gwcapi, gpu_ctx):
    """Which lane runs which CRU follows the CRUs' own iteration counts, so two
    runs of a refilled launch give the same bits; the first step, whose
    constants come from the initial state alone, leaves every CRU in the same
    state whatever the number of CRUs per lane."""
    cells = [gwcapi.make_cell(n, 30 + n, None, PACE) for n in (777, 3000, 64)]

    def run(launch, steps):
        b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
        b.set_mode(frozen=False, digest=False)
        b.set_launch(*launch)
        b.step(steps)
        out = ([x.as_array() for x in b.get_globals()], [b.get_cru_state(c) for c in range(len(cells))])
        b.close()
        return out

    say("refill determinism")
    r1, r2 = run((1, 128, 8), 3000), run((1, 128, 8), 3000)
    for x, y in zip(r1[0], r2[0]):
        assert np.array_equal(x, y)
    for x, y in zip(r1[1], r2[1]):
        for k in x:
            assert np.array_equal(x[k], y[k]), k
    one = [run(sh, 1)[1] for sh in ((1, 64, 1), (1, 64, 4), (1, 256, 8))]
    for st in one[1:]:
        for x, y in zip(st, one[0]):
            for k in ("LCC", "LCC_inactivation", "RyR", "ClCh"):
                assert np.array_equal(x[k], y[k]), k
            for k in ("CaSS", "CaJSR"):
                assert np.array_equal(x[k], y[k]), k


@pytest.mark.parametrize("launch", [(1, 64, 4), (4, 256, 1)], ids=["lane-refill", "quad"])
def test_fp32_local_channel_counts_stay_valid_at_scale(gwcapi, gpu_ctx, launch):
    """64 cells x 12,500 CRUs through the first 45 ms of a paced beat (the
    release), where FP32 rounding of the event choice once let a RyR leave an
    empty state and the resulting negative propensity blew a CRU up: the run
    completes, no count is negative outside a lumped pair, lumped pair sums
    are not negative, and each subunit holds 5 RyRs."""
    cells = [gwcapi.make_cell(12500, 42 + s, None, PACE) for s in range(64)]
    b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
    b.set_mode(frozen=False, digest=False)
    b.set_launch(*launch)
    say(f"fp32-local at scale, launch {launch}")
    b.step(45_000)
    for c in range(len(cells)):
        st = b.get_cru_state(c)
        R, ca = st["RyR"], st["CaSS"]
        assert np.all(R.sum(axis=2) == 5)
        assert np.all(R[..., 0] >= 0) and np.all(R[..., 1] >= 0)
        assert np.all(R[..., 2] + R[..., 3] >= 0) and np.all(R[..., 4] + R[..., 5] >= 0)
        for k, thr in ((2, 0.03685), (3, 0.03685), (4, 1.15e-4), (5, 1.15e-4)):
            assert not np.any((R[..., k] < 0) & (ca <= thr)), (c, k)
        assert np.all(np.isfinite(ca)) and np.all(ca > 0)
    b.close()


def test_fp32_local_whole_beat_at_scale(gwcapi, gpu_ctx):
    """64 cells x 12,500 CRUs through a whole paced beat (500 ms, about 30 s on
    one GPU): every channel count stays valid, every cell repolarises with an
    APD90 in the range of the FP64 ensemble, and the cells return to rest."""
    cells = [gwcapi.make_cell(12500, 42 + s, None, PACE) for s in range(64)]
    b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
    b.set_mode(frozen=False, digest=False)
    b.set_launch(1, 64, 4)
    for k in range(5):
        say(f"fp32-local whole beat at scale, t = {k * 100}..{(k + 1) * 100} ms")
        b.step(100_000)
    for c in range(len(cells)):
        R = b.get_cru_state(c)["RyR"]
        assert np.all(R.sum(axis=2) == 5)
        assert np.all(R[..., 0] >= 0) and np.all(R[..., 1] >= 0)
        assert np.all(R[..., 2] + R[..., 3] >= 0) and np.all(R[..., 4] + R[..., 5] >= 0)
    apd = np.array([o["apd90_last"] for o in b.obs()])
    v = np.array([x.V for x in b.get_globals()])
    say(f"  APD90 {apd.min():.1f}..{apd.max():.1f} ms, final V {v.min():.2f}..{v.max():.2f} mV")
    assert np.all((apd > 250) & (apd < 450))
    assert np.all(v < -85.0)
    b.close()


def test_fp32_local_clamp_family_errors(gwcapi, gpu_ctx):
    """Voltage-clamp steps from -80 mV to -40..+40 mV: peak ICaL and the
    integrated RyR release flux of the FP32-local form agree with FP64 within
    3 SE at every step potential."""
    seeds, n, hold_ms, step_ms = 8, 2000, 5.0, 50.0
    volts = [-40.0, -30.0, -20.0, -10.0, 0.0, 10.0, 20.0, 30.0, 40.0]
    steps = int(round((hold_ms + step_ms) * 1000))
    cells = []
    for v in volts:
        proto = {"kind": gwcapi.PROTO_CLAMP, "hold": -80.0, "step": v, "t_on": hold_ms - 5e-4,
                 "t_off": hold_ms + step_ms}
        cells += [gwcapi.make_cell(n, 900 + s, None, proto) for s in range(seeds)]
    say(f"clamp family {len(volts)} x {seeds} seeds x {n} CRUs")
    o64, _, _ = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP64, cells, steps)
    o32, _, _ = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP32_LOCAL, cells, steps)
    for i, v in enumerate(volts):
        sl = slice(i * seeds, (i + 1) * seeds)
        for key in ("peak_ICaL_neg", "int_Jrel"):
            a = np.array([o[key] for o in o32[sl]])
            b = np.array([o[key] for o in o64[sl]])
            _check_within_3se(a, b, f"{key} at {v:+.0f} mV")
