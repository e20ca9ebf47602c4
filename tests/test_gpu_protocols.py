"""GPU parity of the protocol paths and of the drop-in Python surface.

* the voltage-clamp protocol (hold, step, hold), including a step to exactly
  0 mV, where the LCC flux, Ito2 and IKv14 take their analytic limits;
* GreensteinWinslow.run (src/pyGW.cpp:341-344) and GWModel.simulate
  (src/model.py:940-982), checked record by record against the oracle.

All comparisons are exact."""
import importlib
import os
import sys

import numpy as np
import pytest

from oracle.oracle import GwGlobals, OracleModel

from .test_gpu_parity import STATE_KEYS, assert_globals_equal, assert_states_equal, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pkg():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    return importlib.import_module("gw-model_amd")


@pytest.fixture(scope="module")
def gw(pkg):
    import build.GreensteinWinslow as m  # resolved through the package's sys.path entry

    return m


@pytest.mark.parametrize("vstep", [10.0, 0.0])
def test_voltage_clamp_steps_bit_exact(gwcapi, gpu_ctx, vstep):
    n, steps, t_on, t_off, hold = 200, 4000, 1.0005, 3.0005, -80.0
    say(f"clamp step to {vstep} mV")
    proto = {"kind": gwcapi.PROTO_CLAMP, "hold": hold, "step": vstep, "t_on": t_on, "t_off": t_off}
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 21, None, proto)])
    b.set_mode(frozen=False, digest=True)
    b.step(steps)
    m = OracleModel(n, seed=21)
    m.sample_rest_state()
    t = 0.0
    for _ in range(steps):
        m.set_clamp(True, vstep if t_on <= t < t_off else hold)
        assert m.step(1e-3, 0.0) == 0
        t += 1e-3
    assert_globals_equal(b.get_globals()[0], m.globals())
    c_g, h_g = b.digest(0)
    c_o, h_o = m.digest()
    assert np.array_equal(c_g, c_o) and np.array_equal(h_g, h_o)
    s = b.get_cru_state(0)
    assert_states_equal(s, m.state())
    b.close()
    # the depolarising step opened L-type channels and released calcium
    assert int(c_o.sum()) > n * 20


def _oracle_globals_from_binding(gs):
    o = GwGlobals()
    for k in ("V", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "m", "h", "j", "xKs"):
        setattr(o, k, getattr(gs, k))
    o.Kr[:] = list(np.asarray(gs.XKr).ravel())
    o.Kv14[:] = list(np.asarray(gs.XKv14).ravel())
    o.Kv43[:] = list(np.asarray(gs.XKv43).ravel())
    return o


def _stimulus_table(fn, steps, dt=1e-3):
    out = np.zeros(steps)
    t = 0.0
    for k in range(steps):
        out[k] = fn(t)
        t += dt
    return out


RECORD_FIELDS = ("t", "V", "m", "h", "j", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "xKs",
                 "CaJSR", "CaSS", "LCC", "LCC_inactivation", "RyR", "ClCh")


def test_pybind_run_matches_oracle(gw):
    n, steps, rec, seed = 60, 1500, 300, 5
    say("pybind run")
    stim = lambda t: 100.0 if t < 1.0 else 0.0  # noqa: E731
    p = gw.GWParameters()
    c = gw.rest_state(n, seed)
    gs = gw.GWGlobalState()
    out = gw.run(p, n, 1e-3, steps, stim, rec, c, gs, PRNG="xoshiro256**", seed=seed)
    m = OracleModel(n, seed=seed)
    m.set_state(globals_=_oracle_globals_from_binding(gs), CaSS=c.CaSS, CaJSR=c.CaJSR, LCC=c.LCC,
                LCC_inactivation=c.LCC_inactivation, RyR=c.RyR, ClCh=c.ClCh)
    ref = m.run(1e-3, steps, _stimulus_table(stim, steps), record_every=rec)
    assert out.V.shape == (steps // rec,)
    for k in RECORD_FIELDS:
        assert np.array_equal(np.asarray(getattr(out, k)), ref[k]), k
    assert np.array_equal(np.asarray(out.XKr), ref["XKr"]) and np.array_equal(np.asarray(out.XKv14), ref["XKv14"])


def test_gwmodel_simulate_matches_oracle(pkg):
    n, steps, rec, seed = 40, 1200, 400, 8
    say("GWModel.simulate")
    params = pkg.GWParameters(NCaRU_sim=n)
    stim = lambda t: 100.0 if t < 1.0 else 0.0  # noqa: E731
    model = pkg.GWModel(parameters=params, stimulus_fn=stim)
    sol = model.simulate(1e-3, steps, record_every=rec, PRNG="xoshiro256++", seed=seed)
    import build.GreensteinWinslow as gw

    c0 = gw.GWCRUState(n)
    m = OracleModel(n, seed=seed, prng=1)
    m.set_state(globals_=_oracle_globals_from_binding(gw.GWGlobalState()), CaSS=c0.CaSS, CaJSR=c0.CaJSR,
                LCC=c0.LCC, LCC_inactivation=c0.LCC_inactivation, RyR=c0.RyR, ClCh=c0.ClCh)
    ref = m.run(1e-3, steps, _stimulus_table(stim, steps), record_every=rec)
    for k in RECORD_FIELDS:
        assert np.array_equal(np.asarray(getattr(sol, k)), ref[k]), k
    # post-hoc ICaL (src/model.py:414-449) from the recorded state
    ical = sol.ICaL
    assert ical.shape == (steps // rec,)
    assert np.array_equal(ical, pkg.ICaL(ref["V"], ref["LCC"], ref["LCC_inactivation"], ref["CaSS"], params))
    assert float(sol.V[-1]) > -40.0


def test_state_round_trip_and_restart(gwcapi, gpu_ctx):
    """get_cru_state / set_cru_state round-trip exactly, and a run split in two
    calls equals one call (every call ends with all steps closed)."""
    n = 300
    say("state round trip")
    pace = {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    a = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 4, None, pace)])
    a.set_mode(frozen=False, digest=True)
    a.step(1500)
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 4, None, pace)])
    b.set_mode(frozen=False, digest=True)
    for k in (1, 2, 7, 490, 1000):
        b.step(k)
    assert_globals_equal(a.get_globals()[0], b.get_globals()[0])
    assert_states_equal(a.get_cru_state(0), b.get_cru_state(0))
    s = a.get_cru_state(0)
    a.set_cru_state(0, **s)
    assert_states_equal(a.get_cru_state(0), s, STATE_KEYS)
    a.close()
    b.close()


def _snapshot(b, ncells):
    g = [x.as_array() for x in b.get_globals()]
    s = [b.get_cru_state(c) for c in range(ncells)]
    return g, s


@pytest.mark.parametrize("precision", [0, 1])
def test_launch_shapes_give_identical_results(gwcapi, gpu_ctx, precision):
    """Every quad launch shape, refilled or not, gives the same bits (exact flux
    sums, per-CRU streams)."""
    say(f"launch shapes, precision {precision}")
    pace = {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    specs = ((37, 1), (300, 2), (1000, 3))
    ref = None
    shapes = ((64, 1), (128, 1), (256, 1), (512, 1))
    if precision == 0:   # refilled quads are built for the FP64 form
        shapes += ((64, 2), (64, 4), (128, 4), (256, 4))
    for blk, cpl in shapes:
        b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, s, None, pace) for n, s in specs], precision=precision)
        b.set_mode(frozen=False, digest=False)
        b.set_launch(4, blk, cpl)
        assert b.launch() == (4, blk, cpl)
        b.step(1500)
        snap = _snapshot(b, len(specs))
        b.close()
        if ref is None:
            ref = snap
            continue
        for x, y in zip(snap[0], ref[0]):
            assert np.array_equal(x, y), (blk, cpl)
        for x, y in zip(snap[1], ref[1]):
            assert_states_equal(x, y)


def test_autotune_leaves_the_state_unchanged(gwcapi, gpu_ctx):
    say("autotune")
    pace = {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    cells = [gwcapi.make_cell(2000, 10 + c, None, pace) for c in range(4)]
    a = gwcapi.Batch(gpu_ctx, cells)
    a.set_mode(frozen=False, digest=False)
    a.step(300)
    shape, ms = a.autotune(40)
    say(f"  chose {shape}: {ms}")
    assert shape[0] == 4 and a.launch() == shape   # quads, refilled or not: FP64
    assert all(ms[k] is not None and ms[k] > 0 for k in ("4x64", "4x128", "4x256", "4x512", "4x64x2", "4x64x4"))
    lane = [k for k in ms if k.startswith("1x")]
    assert len(lane) == 12 and all(ms[k] is None for k in lane)  # one lane per CRU is FP32-local only
    a.step(300)
    b = gwcapi.Batch(gpu_ctx, cells)
    b.set_mode(frozen=False, digest=False)
    b.step(600)
    ga, sa = _snapshot(a, len(cells))
    gb, sb = _snapshot(b, len(cells))
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y)
    for x, y in zip(sa, sb):
        assert_states_equal(x, y)
    t_a, n_a = a.time()
    t_b, n_b = b.time()
    assert n_a == n_b == 600 and t_a == t_b
    a.close()
    b.close()


@pytest.mark.parametrize("precision,launch,digest",
                         [(0, (4, 256, 1), True), (0, (4, 64, 1), False), (1, (1, 128, 1), False), (1, (1, 64, 4), False)],
                         ids=["fp64-digest", "fp64-4x64", "fp32-1x128", "fp32-1x64x4"])
def test_graph_replays_match_plain_launches(gwcapi, gpu_ctx, precision, launch, digest):
    """Steps replayed as captured graphs give the same bits as plain launches of
    one step each (graphs replay one step per launch), over calls of several
    lengths with the host reading the state between them."""
    say(f"graphs {precision} {launch}")
    pace = {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    cells = [gwcapi.make_cell(n, s, None, pace) for n, s in ((300, 5), (1000, 6))]
    runs = []
    for graph in (0, 16):
        b = gwcapi.Batch(gpu_ctx, cells, precision=precision)
        b.set_mode(frozen=False, digest=digest)
        b.set_launch(*launch)
        b.set_steps_per_launch(1)
        b.set_graph(graph)
        runs.append(b)
    for chunk in (1, 40, 17, 100, 3, 250):
        for b in runs:
            b.step(chunk)
        ga, sa = _snapshot(runs[0], len(cells))
        gb, sb = _snapshot(runs[1], len(cells))
        for x, y in zip(ga, gb):
            assert np.array_equal(x, y), chunk
        for x, y in zip(sa, sb):
            assert_states_equal(x, y)
        if digest:
            for c in range(len(cells)):
                da, db = runs[0].digest(c), runs[1].digest(c)
                assert np.array_equal(da[0], db[0]) and np.array_equal(da[1], db[1])
        assert runs[0].time() == runs[1].time()
    for b in runs:
        b.close()


def _apd90_from_trace(V, t, v0):
    """APD90 of every beat from a sampled V trace, by the definition the device
    tracker uses: upstroke at the upward -40 mV crossing, end at the downward
    crossing of V_rest + 0.1 (V_peak - V_rest), both interpolated linearly."""
    out = []
    t_up, v_rest, v_peak, v_old, t_old = -1.0, v0, v0, v0, 0.0
    for v, tt in zip(V, t):
        if t_up < 0:
            v_rest = min(v_rest, v)
            if v_old < -40.0 <= v:
                t_up = t_old + (tt - t_old) * (-40.0 - v_old) / (v - v_old)
                v_peak = v
        else:
            v_peak = max(v_peak, v)
            thr = v_rest + 0.1 * (v_peak - v_rest)
            if v_old > thr and v <= thr and v < v_peak - 10.0:
                out.append(t_old + (tt - t_old) * (v_old - thr) / (v_old - v) - t_up)
                t_up, v_rest = -1.0, v
        v_old, t_old = v, tt
    return out


def test_on_device_apd_and_alternans_match_the_trace(gwcapi, gpu_ctx):
    """The per-cell observables (APD90 of the last two beats, mean beat-to-beat
    APD90 difference, beats, peaks) agree with the same quantities computed on
    the host from a V trace taken every second step, over three beats at 2 Hz."""
    ncell, n, period, beats, every = 4, 500, 500.0, 3, 2
    steps = int(beats * period * 1000)
    say(f"observables {ncell} cells x {n} CRUs, {steps} steps")
    pace = {"kind": gwcapi.PROTO_PACE, "period": period, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 70 + c, None, pace) for c in range(ncell)])
    b.set_mode(frozen=False, digest=False)
    b.enable_trace(every, steps // every + 8)
    b.step(steps)
    tr = b.get_trace()
    obs = b.obs()
    b.close()
    assert tr.shape[0] == steps // every
    t = np.arange(1, tr.shape[0] + 1) * every * 1e-3
    for c in range(ncell):
        apds = _apd90_from_trace(tr[:, c, 0], t, gwcapi.default_globals().V)
        o = obs[c]
        say(f"  cell {c}: device beats {o['beats_done']:.0f} APD90 {o['apd90_prev']:.3f}, {o['apd90_last']:.3f}; "
            f"trace {[round(x, 3) for x in apds]}")
        assert len(apds) >= 2 and o["beats_done"] == len(apds)
        assert abs(o["apd90_last"] - apds[-1]) < 0.02 and abs(o["apd90_prev"] - apds[-2]) < 0.02
        assert abs(o["alt_sum"] / o["alt_count"] - float(np.mean(np.abs(np.diff(apds))))) < 0.02
        assert o["peak_Cai"] >= tr[:, c, 1].max() and o["peak_ICaL_neg"] <= tr[:, c, 2].min()
        assert abs(o["sum_open_RyR_frac"] / o["n_steps"] - tr[:, c, 3].mean()) < 2e-3


@pytest.mark.parametrize("field, value, codes", [("Nai", -1.0, ("GW_E_NONPOSITIVE_CONC",)),
                                                  ("V", float("nan"), ("GW_E_NAN", "GW_E_RANGE"))],
                         ids=["nonpositive-conc", "nan"])
def test_a_failing_cell_stops_alone(gwcapi, gpu_ctx, field, value, codes):
    """A cell whose globals become invalid stops with its own status, and the
    step call reports it, naming the cell, instead of ending the process (the
    reference calls exit() on a non-positive Nernst argument, common.hpp:72-75).
    The other cells of the batch run on bit-identical to a batch without it."""
    pace = {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    cells = [gwcapi.make_cell(256, 70 + k, None, pace) for k in range(3)]
    say(f"failing cell: {field} = {value}")
    ref = gwcapi.Batch(gpu_ctx, cells)
    ref.step(2000)
    want = [g.as_array() for g in ref.get_globals()]
    ref.close()
    b = gwcapi.Batch(gpu_ctx, cells)
    g = b.get_globals()[1]
    setattr(g, field, value)
    b.set_globals(1, g)
    with pytest.raises(gwcapi.GwError) as e:
        b.step(2000)
    assert gwcapi.STATUS[e.value.code] in codes
    assert "cell 1" in str(e.value)
    got = [g.as_array() for g in b.get_globals()]
    for c in (0, 2):
        assert np.array_equal(got[c], want[c]), c
    b.close()
