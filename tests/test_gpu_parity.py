"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

FP64 mode is bit-exact: the same per-CRU xoshiro streams, the same operation
order and libm, exact flux sums. Every comparison below is exact equality of
doubles and integers, plus equality of per-CRU event counts and 64-bit event
sequence hashes (so the full integer channel-state sequences agree)."""
import math
import sys

import numpy as np
import pytest

from oracle.oracle import GwGlobals, GwParams, OracleModel

pytestmark = pytest.mark.gpu

STATE_KEYS = ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "RyR", "ClCh", "JLCC")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_params(cap_params):
    p = GwParams.default()
    for name in cap_params.__class__._fields_:
        setattr(p, name[0], getattr(cap_params, name[0]))
    return p


def to_oracle_globals(g):
    o = GwGlobals()
    for name, _ in o._fields_:
        v = getattr(g, name)
        if hasattr(v, "__len__"):
            getattr(o, name)[:] = list(v)
        else:
            setattr(o, name, v)
    return o


def assert_states_equal(a, b, keys=STATE_KEYS):
    for k in keys:
        assert a[k].shape == b[k].shape, k
        if not np.array_equal(a[k], b[k]):
            bad = np.argwhere(a[k] != b[k])
            raise AssertionError(f"{k}: {len(bad)} mismatches, first at {bad[0].tolist()}: "
                                 f"gpu={a[k][tuple(bad[0])]!r} oracle={b[k][tuple(bad[0])]!r}")


def assert_globals_equal(g_gpu, g_orc):
    a = g_gpu.as_array()
    b = g_orc.as_array()
    if not np.array_equal(a, b):
        i = int(np.argwhere(a != b)[0][0])
        raise AssertionError(f"global {i}: gpu={a[i]!r} oracle={b[i]!r}")


@pytest.mark.parametrize("n,seed", [(1, 0), (300, 42), (1250, 7)])
def test_rest_sampler_matches_oracle(gwcapi, gpu_ctx, n, seed):
    say(f"rest sampler n={n} seed={seed}")
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, seed)])
    m = OracleModel(n, seed=seed)
    m.sample_rest_state()
    assert_states_equal(b.get_cru_state(0), m.state())
    b.close()


def _frozen_pair(gwcapi, gpu_ctx, n, seed, V, Cai, prng=0, kind="strict"):
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, seed)], prng=prng)
    g = gwcapi.default_globals()
    g.V, g.Cai = V, Cai
    b.set_globals(0, g)
    b.set_mode(frozen=True, digest=True)
    m = OracleModel(n, seed=seed, prng=prng, kind=kind)
    m.sample_rest_state()
    m.set_state(globals_=to_oracle_globals(g))
    m.set_frozen(True)
    return b, m


def test_frozen_replay_short(gwcapi, gpu_ctx):
    say("frozen replay n=300 steps=3000")
    b, m = _frozen_pair(gwcapi, gpu_ctx, 300, 42, -20.0, 5e-4)
    b.step(3000)
    for _ in range(3000):
        assert m.step(1e-3, 0.0) == 0
    c_g, h_g = b.digest(0)
    c_o, h_o = m.digest()
    assert c_o.sum() > 300 * 30
    assert np.array_equal(c_g, c_o)
    assert np.array_equal(h_g, h_o)
    assert_states_equal(b.get_cru_state(0), m.state())


@pytest.mark.parametrize("prng", range(6))
def test_frozen_replay_every_engine(gwcapi, gpu_ctx, prng):
    say(f"frozen replay prng={prng}")
    b, m = _frozen_pair(gwcapi, gpu_ctx, 64, 1000 + prng, 0.5, 2e-3, prng=prng)
    b.step(800)
    for _ in range(800):
        assert m.step(1e-3, 0.0) == 0
    c_g, h_g = b.digest(0)
    c_o, h_o = m.digest()
    assert np.array_equal(c_g, c_o) and np.array_equal(h_g, h_o)
    assert_states_equal(b.get_cru_state(0), m.state())


def test_frozen_replay_1e5_events_per_cru(gwcapi, gpu_ctx):
    """Fixed-uniform replay: every CRU's channel-state sequence agrees over at
    least its first 10^5 events (the run covers more for most CRUs)."""
    n, steps = 32, 6_000_000
    say(f"long frozen replay n={n} steps={steps}")
    b, m = _frozen_pair(gwcapi, gpu_ctx, n, 2026, -40.0, 2e-3, kind="par")
    chunk = 500_000
    for k in range(0, steps, chunk):
        b.step(chunk)
        say(f"  gpu {k + chunk} steps")
    rc = m.L.gwo_model_run(m.h, 1e-3, steps, None, 0, None)
    assert rc == 0
    c_g, h_g = b.digest(0)
    c_o, h_o = m.digest()
    say(f"  events per CRU: min {int(c_o.min())} mean {float(c_o.mean()):.0f}")
    assert int(c_o.min()) >= 100_000
    assert np.array_equal(c_g, c_o)
    assert np.array_equal(h_g, h_o)
    assert_states_equal(b.get_cru_state(0), m.state())


def _pulse(steps, dt=1e-3, amp=100.0, dur=1.0):
    t = 0.0
    out = np.zeros(steps)
    for k in range(steps):
        out[k] = amp if t < dur else 0.0
        t += dt
    return out


@pytest.mark.parametrize("n,steps", [(128, 20000), (1250, 3000)])
def test_coupled_paced_ap_bit_exact(gwcapi, gpu_ctx, n, steps):
    """Coupled run (global ODE + SSA every step), +100 pA/pF for 1 ms from rest."""
    say(f"coupled n={n} steps={steps}")
    istim = _pulse(steps)
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 3)])
    b.set_mode(frozen=False, digest=True)
    b.set_stimulus_table(0, istim)
    b.step(steps)
    m = OracleModel(n, seed=3)
    m.sample_rest_state()
    m.run(1e-3, steps, istim)
    g_gpu = b.get_globals()[0]
    assert_globals_equal(g_gpu, m.globals())
    c_g, h_g = b.digest(0)
    c_o, h_o = m.digest()
    assert np.array_equal(c_g, c_o) and np.array_equal(h_g, h_o)
    assert_states_equal(b.get_cru_state(0), m.state())
    b.close()
    if steps >= 20000:
        assert g_gpu.V > -40.0  # the stimulus produced an action potential (plateau)


def test_gw_run_records_match_oracle(gwcapi, gpu_ctx):
    """gw_run (the run<PRNG> loop) records every array of GWVariables exactly;
    num_steps not a multiple of record_every gives ceil(num_steps/record_every) rows."""
    n, steps, rec = 50, 2501, 500
    say("gw_run records")
    m = OracleModel(n, seed=9)
    m.sample_rest_state()
    crus = m.state()
    istim = _pulse(steps, amp=35.0, dur=2.0)
    g0 = gwcapi.default_globals()
    out_g = gwcapi.run(gpu_ctx, gwcapi.Params.default(), n, 1e-3, steps, istim, rec, g0, crus, seed=9)
    m2 = OracleModel(n, seed=9)
    m2.set_state(globals_=to_oracle_globals(g0), **{k: crus[k] for k in STATE_KEYS if k != "JLCC"})
    out_o = m2.run(1e-3, steps, istim, record_every=rec)
    assert out_g["t"].shape[0] == 6
    for k, v in out_o.items():
        assert np.array_equal(out_g[k], v), k


@pytest.mark.parametrize("steps,rec,rows", [(0, 10, 0), (3, 1000, 1), (10, 10, 1)], ids=["empty", "short", "exact"])
def test_gw_run_degenerate_lengths(gwcapi, gpu_ctx, steps, rec, rows):
    """gw_run with no steps, with fewer steps than record_every and with
    exactly one record interval: ceil(num_steps / record_every) rows, each
    equal to the oracle's."""
    n = 20
    say(f"gw_run steps={steps} record_every={rec}")
    m = OracleModel(n, seed=5)
    m.sample_rest_state()
    crus = m.state()
    istim = _pulse(steps)
    g0 = gwcapi.default_globals()
    out_g = gwcapi.run(gpu_ctx, gwcapi.Params.default(), n, 1e-3, steps, istim, rec, g0, crus, seed=5)
    assert out_g["t"].shape[0] == rows
    m2 = OracleModel(n, seed=5)
    m2.set_state(globals_=to_oracle_globals(g0), **{k: crus[k] for k in STATE_KEYS if k != "JLCC"})
    out_o = m2.run(1e-3, steps, istim, record_every=rec)
    for k, v in out_o.items():
        assert out_g[k].shape == v.shape, k
        assert np.array_equal(out_g[k], v), k


def test_uneven_batch_matches_independent_oracles(gwcapi, gpu_ctx):
    """Cells with different sizes, parameters, seeds and pacing advance in one
    launch exactly as each would alone."""
    say("uneven batch")
    specs = [(37, 11, {"PCaL": 1.5}), (300, 12, {"GKr": 0.7, "GKs": 1.9}), (513, 13, {})]
    steps = 3000
    cells, params = [], []
    for n, seed, scale in specs:
        p = gwcapi.Params.default()
        for k, f in scale.items():
            setattr(p, k, getattr(p, k) * f)
        params.append(p)
        cells.append(gwcapi.make_cell(n, seed, p, {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0,
                                                   "amplitude": 100.0, "start": 0.0}))
    b = gwcapi.Batch(gpu_ctx, cells)
    b.set_mode(frozen=False, digest=True)
    b.step(steps)
    gg = b.get_globals()
    for c, ((n, seed, _), p) in enumerate(zip(specs, params)):
        t = 0.0
        ist = np.zeros(steps)
        for k in range(steps):
            ist[k] = 100.0 if math.fmod(t, 1000.0) < 1.0 else 0.0
            t += 1e-3
        m = OracleModel(n, params=oracle_params(p), seed=seed)
        m.sample_rest_state()
        m.run(1e-3, steps, ist)
        assert_globals_equal(gg[c], m.globals())
        c_g, h_g = b.digest(c)
        c_o, h_o = m.digest()
        assert np.array_equal(c_g, c_o) and np.array_equal(h_g, h_o)
        assert_states_equal(b.get_cru_state(c), m.state())


def test_extreme_cell_sizes(gwcapi, gpu_ctx):
    """CRU counts at both ends of the supported range, 1 and 2^20, advance in
    one batch exactly as the oracle; 0 and 2^20 + 1 are rejected."""
    pace = {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    for bad in (0, (1 << 20) + 1):
        with pytest.raises(gwcapi.GwError) as e:
            gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(bad, 1, None, pace)])
        assert gwcapi.STATUS[e.value.code] == "GW_E_INVALID_ARG"
    specs = [(1, 21), (1 << 20, 22)]
    steps = 20
    say("extreme cell sizes 1 and 2^20")
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, seed, None, pace) for n, seed in specs])
    b.set_mode(frozen=False, digest=True)
    b.step(steps)
    gg = b.get_globals()
    ist = np.zeros(steps)
    t = 0.0
    for k in range(steps):
        ist[k] = 100.0 if t < 1.0 else 0.0
        t += 1e-3
    for c, (n, seed) in enumerate(specs):
        m = OracleModel(n, seed=seed, kind="par")   # the strict restatement with OpenMP over CRUs
        m.sample_rest_state()
        m.run(1e-3, steps, ist)
        assert_globals_equal(gg[c], m.globals())
        c_g, h_g = b.digest(c)
        c_o, h_o = m.digest()
        assert np.array_equal(c_g, c_o) and np.array_equal(h_g, h_o)
        assert_states_equal(b.get_cru_state(c), m.state())
    b.close()


def test_full_size_cell_invariants_and_determinism(gwcapi, gpu_ctx):
    """12,500-CRU cell: channel-state invariants hold after an upstroke, and
    the same seed reproduces the same bits."""
    n, steps = 12500, 2000
    say("full-size invariants")
    runs = []
    for _ in range(2):
        b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 77, None, {"kind": gwcapi.PROTO_PACE, "period": 1000.0,
                                                                  "duration": 1.0, "amplitude": 100.0})])
        b.set_mode(frozen=False, digest=True)
        b.step(steps)
        runs.append((b.get_globals()[0].as_array(), b.get_cru_state(0), b.digest(0)))
        b.close()
    (g1, s1, d1), (g2, s2, d2) = runs
    assert np.array_equal(g1, g2) and np.array_equal(d1[1], d2[1])
    assert_states_equal(s1, s2)
    R = s1["RyR"]
    assert np.all(R.sum(axis=2) == 5)
    assert np.all(R[..., 2] + R[..., 3] >= 0) and np.all(R[..., 4] + R[..., 5] >= 0)
    assert np.all((s1["LCC"] >= 1) & (s1["LCC"] <= 12))
    assert set(np.unique(s1["LCC_inactivation"])) <= {0, 1} and set(np.unique(s1["ClCh"])) <= {0, 1}
    open_lcc = ((s1["LCC"] == 6) | (s1["LCC"] == 12)) & (s1["LCC_inactivation"] == 1)
    assert np.all(s1["JLCC"][~open_lcc] == 0.0)
    assert np.all(np.isfinite(s1["CaSS"])) and np.all(s1["CaSS"] > 0)
    assert g1[0] > -40.0  # depolarised by the 1 ms stimulus
