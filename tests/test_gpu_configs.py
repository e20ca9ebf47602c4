"""The workloads at their full shapes.

* One cell of 1,250 CRUs, default parameters, a 1 ms stimulus at t = 0 and
  500 ms of simulated time with seed 42 (the plain method's CPU-runnable
  case): the FP64 form bit-identical to the oracle over the whole beat, every
  record and the final CRU state.

* 256 cells with CRU counts uniform in 2,000..12,500, PCaL, GKr and GKs scaled
  log-uniform in [0.5, 2], per-cell seeds, 1 Hz pacing, FP32-local with the
  steps replayed as captured graphs and the launch shape picked by the
  autotuner, through its two beats: at 455 ms every cell has fired, the cells
  with a completed beat are back at rest and the others (more L-type current)
  are still in their action potential; at 2 s every cell that completed its
  first beat by 455 ms has completed its second and is back at rest; channel
  counts stay valid; the same batch in FP64 gives its smallest cells
  bit-identical to the oracle.
* 1,024 cells x 12,500 CRUs (12.8 M CRUs) with six conductances scaled
  log-uniform in [0.5, 2], 2 Hz pacing: FP32-local through the upstroke and the
  Ca release (50 ms) with every cell depolarised and releasing and the
  channel counts of sampled cells valid; FP64 over the first 5 ms with the
  first cell bit-identical to the oracle at full size.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleModel

from .test_gpu_parity import (STATE_KEYS, _pulse, assert_globals_equal, assert_states_equal, oracle_params, say,
                               to_oracle_globals)

pytestmark = pytest.mark.gpu


def test_plain_cell_whole_beat_matches_the_oracle(gwcapi, gpu_ctx):
    n, steps, rec = 1250, 500_000, 10_000
    say(f"plain cell: {n} CRUs x {steps} steps against the oracle")
    m = OracleModel(n, seed=42, kind="par")
    m.sample_rest_state()
    crus = m.state()
    istim = _pulse(steps)
    g0 = gwcapi.default_globals()
    out_g = gwcapi.run(gpu_ctx, gwcapi.Params.default(), n, 1e-3, steps, istim, rec, g0, crus, seed=42)
    say("  GPU done; oracle")
    m2 = OracleModel(n, seed=42, kind="par")
    m2.set_state(globals_=to_oracle_globals(g0), **{k: crus[k] for k in STATE_KEYS if k != "JLCC"})
    out_o = m2.run(1e-3, steps, istim, record_every=rec)
    assert out_g["t"].shape[0] == steps // rec
    for k, v in out_o.items():
        assert np.array_equal(out_g[k], v), k
    v = out_g["V"]
    say(f"  V: peak {v.max():.1f} mV, last {v[-1]:.2f} mV")
    assert v.max() > 0.0 and v[-1] < -80.0   # an action potential, back near rest by 490 ms


def _pace(gwcapi, period):
    return {"kind": gwcapi.PROTO_PACE, "period": period, "duration": 1.0, "amplitude": 100.0, "start": 0.0}


def _scaled(gwcapi, rng, names):
    p = gwcapi.Params.default()
    for name in names:
        setattr(p, name, getattr(p, name) * float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))))
    return p


def _paced_table(steps, period):
    out = np.zeros(steps)
    t = 0.0
    for k in range(steps):
        out[k] = 100.0 if math.fmod(t, period) < 1.0 else 0.0
        t += 1e-3
    return out


def _check_counts(st, key=""):
    R = st["RyR"]
    assert np.all(R.sum(axis=2) == 5), key
    assert np.all(R[..., 0] >= 0) and np.all(R[..., 1] >= 0), key
    assert np.all(R[..., 2] + R[..., 3] >= 0) and np.all(R[..., 4] + R[..., 5] >= 0), key
    assert np.all((st["LCC"] >= 1) & (st["LCC"] <= 12)), key
    assert np.all(np.isfinite(st["CaSS"])) and np.all(st["CaSS"] > 0) and np.all(st["CaJSR"] > 0), key


def _batch256(gwcapi):
    rng = np.random.default_rng(256)
    cells = []
    for c in range(256):
        n = int(rng.integers(2000, 12501))
        cells.append(gwcapi.make_cell(n, 1000 + c, _scaled(gwcapi, rng, ("PCaL", "GKr", "GKs")), _pace(gwcapi, 1000.0)))
    return cells


def test_uneven_256_cell_batch_through_two_beats(gwcapi, gpu_ctx):
    cells = _batch256(gwcapi)
    sizes = np.array([c.n_cru for c in cells])
    say(f"256-cell batch: {sizes.sum()} CRUs, {sizes.min()}..{sizes.max()} per cell")
    b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
    b.set_mode(frozen=False, digest=False)
    b.step(5000)   # tune in the release, where the beat spends its busiest steps
    shape, trials = b.autotune(40)
    say(f"  autotuner: {shape}, graph {b.graph_steps}; trials {trials}")
    assert shape in gwcapi.launch_configs()
    b.set_launch(*shape)
    b.set_graph(64)   # replayed captured graphs of 64 steps
    for k in range(9):
        b.step(50_000)
        say(f"  t = {5 + (k + 1) * 50} ms")
    obs = b.obs()
    apd = np.array([o["apd90_last"] for o in obs])
    peak_v = np.array([o["peak_V"] for o in obs])
    v = np.array([g.V for g in b.get_globals()])
    base = gwcapi.Params.default()
    pcal = np.array([c.params.PCaL / base.PCaL for c in cells])
    gk = np.array([(c.params.GKr / base.GKr) * (c.params.GKs / base.GKs) for c in cells])
    done = apd > 0
    say(f"  fired {np.sum(peak_v > 0)}/256; repolarised {done.sum()}: APD90 {apd[done].min():.1f}.."
        f"{apd[done].max():.1f} ms (median {np.median(apd[done]):.1f}); still depolarised {np.sum(~done)}")
    for name, x in (("PCaL scale", pcal), ("GKr x GKs scale", gk)):
        say(f"  median {name}: repolarised {np.median(x[done]):.2f}, still depolarised {np.median(x[~done]):.2f}")
    assert np.all(peak_v > 0.0)                            # every cell fired
    assert np.all((apd[done] > 30) & (apd[done] < 455))     # a completed beat has an APD90 within the window
    assert np.all(v[done] < -70.0)                          # and its cell is back near rest
    assert np.all(v[~done] > -70.0)                         # the others are still in their action potential
    # the cells still in their action potential carry more L-type current (PCaL scaled up): in this
    # model the plateau is held by I_CaL, and with PCaL scaled down the dome is lost (APD90 below 150 ms)
    assert np.median(pcal[~done]) > 1.2 * np.median(pcal[done])
    assert np.median(pcal[done & (apd < 150)]) < 0.8
    for c in np.argsort(sizes)[:4].tolist() + np.argsort(sizes)[-4:].tolist():
        _check_counts(b.get_cru_state(c), c)
    # the second beat: on to t = 2 s, just before the third stimulus
    for k in range(31):
        b.step(50_000 if k < 30 else 45_000)
        if k % 5 == 4:
            say(f"  t = {455 + min(k + 1, 30) * 50 + (45 if k == 30 else 0)} ms")
    obs = b.obs()
    beats = np.array([o["beats_done"] for o in obs])
    apd2 = np.array([o["apd90_last"] for o in obs])
    v2 = np.array([g.V for g in b.get_globals()])
    say(f"  at 2 s: beats completed {dict(zip(*np.unique(beats, return_counts=True)))}; APD90 of the second beat "
        f"of the first beat's repolarised cells {apd2[done].min():.1f}..{apd2[done].max():.1f} ms")
    assert np.all(beats[done] == 2)                        # a normal first beat is followed by a second
    assert np.all((apd2[done] > 30) & (apd2[done] < 1000))
    assert np.all(v2[beats == 2] < -70.0)                  # and the cell is back at rest before the third
    for c in np.argsort(sizes)[:4].tolist() + np.argsort(sizes)[-4:].tolist():
        _check_counts(b.get_cru_state(c), c)
    b.close()


def test_uneven_256_cell_batch_smallest_cells_match_the_oracle(gwcapi, gpu_ctx):
    """The same 256 cells in FP64 over the first 2 ms (the upstroke): the three
    smallest cells, with their own sizes, parameters and seeds, are
    bit-identical to independent oracles."""
    cells = _batch256(gwcapi)
    steps = 2000
    b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP64)
    b.set_mode(frozen=False, digest=False)
    b.set_launch(4, 64, 1)
    b.step(steps)
    g = b.get_globals()
    sizes = np.array([c.n_cru for c in cells])
    ist = _paced_table(steps, 1000.0)
    for c in np.argsort(sizes)[:3].tolist():
        say(f"  cell {c}: {cells[c].n_cru} CRUs against the oracle")
        m = OracleModel(cells[c].n_cru, params=oracle_params(cells[c].params), seed=cells[c].seed, kind="par")
        m.sample_rest_state()
        m.run(1e-3, steps, ist)
        assert_globals_equal(g[c], m.globals())
        assert_states_equal(b.get_cru_state(c), m.state())
    b.close()


def _batch1024(gwcapi):
    rng = np.random.default_rng(1024)
    names = ("PCaL", "GKr", "GKs", "GK1", "GKv43", "kNaCa")
    return [gwcapi.make_cell(12500, 5000 + c, _scaled(gwcapi, rng, names), _pace(gwcapi, 500.0)) for c in range(1024)]


def test_large_1024_cell_batch_through_the_release(gwcapi, gpu_ctx):
    cells = _batch1024(gwcapi)
    b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
    b.set_mode(frozen=False, digest=False)
    b.set_launch(1, 64, 4)
    for k in range(5):
        b.step(10_000)
        say(f"1,024 x 12,500 FP32-local: t = {(k + 1) * 10} ms")
    obs = b.obs()
    peak_v = np.array([o["peak_V"] for o in obs])
    peak_cai = np.array([o["peak_Cai"] for o in obs])
    say(f"  peak V {peak_v.min():.1f}..{peak_v.max():.1f} mV, peak Cai {peak_cai.min():.3g}..{peak_cai.max():.3g} mM")
    assert np.all(peak_v > 20.0)                 # every cell fired
    assert np.all(peak_cai > 1.5 * 8.5e-5)       # and released Ca (rest is about 8.5e-5 mM; the cells with
                                                 # PCaL scaled down release least)
    g = b.get_globals()
    assert all(np.all(np.isfinite(x.as_array())) for x in g)
    for c in (0, 1, 511, 1022, 1023):
        _check_counts(b.get_cru_state(c), c)
    b.close()


def test_large_1024_cell_batch_fp64_first_cell_matches_the_oracle(gwcapi, gpu_ctx):
    cells = _batch1024(gwcapi)
    steps = 5000
    b = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP64)
    b.set_mode(frozen=False, digest=False)
    b.set_launch(4, 64, 1)
    b.step(steps)
    say("1,024 x 12,500 FP64: 5 ms done; oracle of cell 0")
    g0 = b.get_globals()[0]
    st0 = b.get_cru_state(0)
    b.close()
    m = OracleModel(12500, params=oracle_params(cells[0].params), seed=cells[0].seed, kind="par")
    m.sample_rest_state()
    m.run(1e-3, steps, _paced_table(steps, 500.0))
    assert_globals_equal(g0, m.globals())
    assert_states_equal(st0, m.state())
