"""The FP32-local variant against the FP64 form.

CRU-local Ca, fluxes and propensities run in FP32; channel states, random
streams, flux sums (exact, FP64) and the global ODE (FP64) are those of the
FP64 form. The FP64 form is bit-identical to the CPU oracle (test_gpu_parity.py),
so it stands for the reference here. The variant is not bit-exact, so its
results are compared as ensembles over seeds: each seed runs in both forms and
the per-seed differences are tested (mean within 3 standard errors of the
paired differences), the spreads are compared (variance ratio, two-sample
Kolmogorov-Smirnov), and the relative error of each observable is printed with
its 95% confidence interval. The one-lane layouts draw their uniforms
differently from the FP64 quads (two per engine draw), so each seed's two runs
are different sample paths and a bias or a wider spread of the variant shows.
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
    """Unpaired: mean of a within 3 SE of mean of b."""
    se = math.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    diff = a.mean() - b.mean()
    say(f"  {name}: fp64 {b.mean():.6g} fp32 {a.mean():.6g} diff {diff:.3g} se {se:.3g} "
        f"rel {diff / b.mean():.2e}")
    assert abs(diff) <= 3 * se + 1e-12 * abs(b.mean()), name
    return diff / b.mean()


def _check_paired(a, b, name):
    """Paired by seed: the mean of the per-seed differences a - b is within 3
    of its standard errors of 0. Prints the relative error (mean difference
    over the FP64 mean) with its 95% confidence interval."""
    d = a - b
    n = len(d)
    se = d.std(ddof=1) / math.sqrt(n)
    m = d.mean()
    ref = b.mean()
    say(f"  {name}: fp64 {ref:.6g} fp32 {a.mean():.6g} | paired diff {m:.3g} +- {1.96 * se:.2g} (95% CI) "
        f"| rel error {m / ref:+.2e} +- {1.96 * se / abs(ref):.1e}")
    assert abs(m) <= 3 * se + 1e-12 * abs(ref), name
    return m / ref, 1.96 * se / abs(ref)


def _check_spread(a, b, name, alpha=1e-3):
    """The spread of a is that of b: variance ratio within the two-sided
    1 - alpha interval of the F distribution, and a two-sample
    Kolmogorov-Smirnov test not rejecting at level alpha."""
    from scipy import stats

    n = len(a)
    ratio = a.var(ddof=1) / b.var(ddof=1)
    lo, hi = stats.f.ppf(alpha / 2, n - 1, n - 1), stats.f.ppf(1 - alpha / 2, n - 1, n - 1)
    ks = stats.ks_2samp(a, b)
    say(f"  {name}: SD fp64 {b.std(ddof=1):.4g} fp32 {a.std(ddof=1):.4g} variance ratio {ratio:.3f} "
        f"(accept {lo:.2f}..{hi:.2f}), KS p {ks.pvalue:.3g}")
    assert lo <= ratio <= hi, name
    assert ks.pvalue > alpha, name


_FP64_ENSEMBLE = {}


@pytest.mark.parametrize("launch", [(1, 64, 1), (1, 64, 4)], ids=["lane", "lane-refill"])
def test_fp32_local_free_mode_ensemble_matches_fp64(gwcapi, gpu_ctx, launch):
    """64 seeds of a 1,000-CRU cell, one paced beat (600 ms: past the longest
    APD90 of either form, so no cell's beat is cut off): APD90, peak Cai, peak
    ICaL and mean RyR open fraction agree seed by seed within 3 SE, and the
    spread of APD90 and peak Cai is the same, for one lane per CRU and for
    lanes refilled from four CRUs each. (The quad layout takes the top 24 bits
    of the FP64 form's own draws, so its runs track the FP64 ones seed by seed
    and are checked for invariants only.)"""
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
        _check_paired(a32[k], a64[k], k)
    for k in ("APD90", "peak_Cai"):
        _check_spread(a32[k], a64[k], k)
    v32 = np.array([x.V for x in g32])
    say(f"  fp32 final V: min {v32.min():.2f} max {v32.max():.2f} mV")
    # every cell repolarised (APD90 above) and none has blown up since
    assert np.all(np.isfinite(v32)) and np.all(v32 < -60.0)


def test_fp32_local_seed_ensemble_at_12500_crus(gwcapi, gpu_ctx):
    """The 64-seed ensemble of a 12,500-CRU cell (the seed ensemble of the
    plain-method configuration), one paced beat (600 ms), the bench's launch
    shapes (FP64 quads 4 x 64, FP32-local lanes refilled from four CRUs,
    1 x 64 x 4): per-seed paired errors of APD90, peak Cai, peak ICaL and mean
    RyR open fraction within 3 SE, the same spread of APD90 and peak Cai, and
    in the FP32-local form every channel count valid and every cell back at
    rest."""
    seeds, n, steps = 64, 12500, 600_000
    cells = [gwcapi.make_cell(n, 42 + s, None, PACE) for s in range(seeds)]
    say(f"fp64 ensemble {seeds} x {n} CRUs x {steps} steps")
    o64, _, _ = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP64, cells, steps, (4, 64, 1))
    say("fp32-local ensemble, launch (1, 64, 4)")
    b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
    b.set_mode(frozen=False, digest=False)
    b.set_launch(1, 64, 4)
    for k in range(6):
        b.step(steps // 6)
        say(f"  fp32-local t = {(k + 1) * steps // 6 // 1000} ms")
    o32 = b.obs()
    for c in range(len(cells)):
        R = b.get_cru_state(c)["RyR"]
        assert np.all(R.sum(axis=2) == 5), c
        assert np.all(R[..., 0] >= 0) and np.all(R[..., 1] >= 0), c
        assert np.all(R[..., 2] + R[..., 3] >= 0) and np.all(R[..., 4] + R[..., 5] >= 0), c
    g32 = b.get_globals()
    b.close()
    a64, a32 = _stats(o64), _stats(o32)
    say(f"  APD90 fp64 {a64['APD90'].min():.1f}..{a64['APD90'].max():.1f} ms, "
        f"fp32 {a32['APD90'].min():.1f}..{a32['APD90'].max():.1f} ms")
    assert np.all(a64["APD90"] > 100) and np.all(a32["APD90"] > 100)
    for k in a64:
        _check_paired(a32[k], a64[k], k)
    for k in ("APD90", "peak_Cai"):
        _check_spread(a32[k], a64[k], k)
    v32 = np.array([x.V for x in g32])
    assert np.all(np.isfinite(v32)) and np.all(v32 < -80.0)


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


def test_fp32_local_clamp_family_errors(gwcapi, gpu_ctx):
    """The voltage-clamp family at full cell size: 12,500 CRUs held at -80 mV
    for 10 ms, then stepped to -40..+40 mV in 10 mV increments for 200 ms,
    4 seeds per potential in both forms. Peak ICaL and the integrated RyR
    release flux of the FP32-local form agree with FP64 seed by seed within
    3 SE at every step potential; the relative errors are printed with their
    95% confidence intervals."""
    seeds, n, hold_ms, step_ms = 4, 12500, 10.0, 200.0
    volts = [-40.0, -30.0, -20.0, -10.0, 0.0, 10.0, 20.0, 30.0, 40.0]
    steps = int(round((hold_ms + step_ms) * 1000))
    cells = []
    for v in volts:
        proto = {"kind": gwcapi.PROTO_CLAMP, "hold": -80.0, "step": v, "t_on": hold_ms - 5e-4,
                 "t_off": hold_ms + step_ms}
        cells += [gwcapi.make_cell(n, 900 + s, None, proto) for s in range(seeds)]
    say(f"clamp family {len(volts)} x {seeds} seeds x {n} CRUs x {steps} steps")
    o64, _, _ = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP64, cells, steps, (4, 64, 1))
    o32, _, _ = _run(gwcapi, gpu_ctx, gwcapi.PREC_FP32_LOCAL, cells, steps, (1, 64, 4))
    for i, v in enumerate(volts):
        sl = slice(i * seeds, (i + 1) * seeds)
        for key in ("peak_ICaL_neg", "int_Jrel"):
            a = np.array([o[key] for o in o32[sl]])
            b = np.array([o[key] for o in o64[sl]])
            _check_paired(a, b, f"{key} at {v:+.0f} mV")
