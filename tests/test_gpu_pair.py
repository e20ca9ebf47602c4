"""Two SSA steps per launch (k_step_lane2, FP32-local one-lane shapes of 64
threads): each CRU advances two steps with its state in registers, after a
close kernel has formed the constants both steps need. With one CRU per lane
the results are bit-identical to one step per launch, over calls of odd and
even lengths with the state read between them, and when a cell fails mid-run;
with refilled lanes they are deterministic and keep every channel count
valid."""
import numpy as np
import pytest

from .test_gpu_parity import assert_states_equal, say

pytestmark = pytest.mark.gpu


def _pace(gwcapi, period=1000.0):
    return {"kind": gwcapi.PROTO_PACE, "period": period, "duration": 1.0, "amplitude": 100.0, "start": 0.0}


def _batch(gwcapi, ctx, cells, launch, steps_per_launch):
    b = gwcapi.Batch(ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
    b.set_mode(frozen=False, digest=False)
    b.set_launch(*launch)
    b.set_steps_per_launch(steps_per_launch)
    assert b.steps_per_launch() == steps_per_launch
    return b


def _snapshot(b, ncells):
    return ([g.as_array() for g in b.get_globals()], [b.get_cru_state(c) for c in range(ncells)], b.obs(),
            b.time(), b.iterations())


def _assert_same(x, y, what):
    for p, q in zip(x[0], y[0]):
        assert np.array_equal(p, q), what
    for p, q in zip(x[1], y[1]):
        assert_states_equal(p, q)
    assert x[2] == y[2], what
    assert x[3] == y[3] and x[4] == y[4], what


def test_two_steps_per_launch_bit_identical_with_one_cru_per_lane(gwcapi, gpu_ctx):
    cells = [gwcapi.make_cell(n, 900 + n, None, _pace(gwcapi)) for n in (37, 300, 1250, 12500)]
    one = _batch(gwcapi, gpu_ctx, cells, (1, 64, 1), 1)
    two = _batch(gwcapi, gpu_ctx, cells, (1, 64, 1), 2)
    done = 0
    for chunk in (1, 2, 3, 64, 501, 1429, 2000):   # odd and even calls, each resuming from a closed state
        one.step(chunk)
        two.step(chunk)
        done += chunk
        say(f"  two steps per launch, one CRU per lane: {done} steps")
        _assert_same(_snapshot(one, len(cells)), _snapshot(two, len(cells)), done)
        if done == 571:   # the same mid-run change of one cell's globals and another's CRU state in both
            for b in (one, two):
                g = b.get_globals()[1]
                g.Cai *= 1.5
                b.set_globals(1, g)
                st = b.get_cru_state(2)
                st["CaJSR"] = st["CaJSR"] * 0.9
                b.set_cru_state(2, **st)
    one.close()
    two.close()


def test_two_steps_per_launch_a_failing_cell_stops_at_the_same_step(gwcapi, gpu_ctx):
    """A cell given a NaN voltage mid-run stops with its status; with one CRU
    per lane every cell, the failed one included, ends bit-identical to one
    step per launch (the same steps run), and the call names the cell."""
    cells = [gwcapi.make_cell(256, 70 + k, None, _pace(gwcapi)) for k in range(3)]
    out = []
    for spl in (1, 2):
        b = _batch(gwcapi, gpu_ctx, cells, (1, 64, 1), spl)
        b.step(1001)
        g = b.get_globals()[1]
        g.V = float("nan")
        b.set_globals(1, g)
        with pytest.raises(gwcapi.GwError) as e:
            b.step(1000)
        assert gwcapi.STATUS[e.value.code] in ("GW_E_NAN", "GW_E_RANGE")
        assert "cell 1" in str(e.value)
        out.append(([g.as_array() for g in b.get_globals()], [b.get_cru_state(c) for c in range(3)]))
        b.close()
    for c in range(3):
        assert np.array_equal(out[0][0][c], out[1][0][c], equal_nan=True), c
        for k, v in out[0][1][c].items():   # the failed cell's CRU state holds NaNs in both
            assert np.array_equal(v, out[1][1][c][k], equal_nan=v.dtype.kind == "f"), (c, k)


def test_two_steps_per_launch_refilled_lanes_deterministic_and_valid(gwcapi, gpu_ctx):
    cells = [gwcapi.make_cell(12500, 40 + k, None, _pace(gwcapi, 500.0)) for k in range(4)]
    snaps = []
    for run in range(2):
        b = _batch(gwcapi, gpu_ctx, cells, (1, 64, 4), 2)
        for chunk in (1, 2999, 7000):   # through the upstroke and the release
            b.step(chunk)
        snaps.append(_snapshot(b, len(cells)))
        b.close()
        say(f"  refilled lanes, two steps per launch: run {run} done")
    _assert_same(snaps[0], snaps[1], "run to run")
    for c, st in enumerate(snaps[0][1]):
        R = st["RyR"]
        assert np.all(R.sum(axis=2) == 5), c
        assert np.all(R[..., 0] >= 0) and np.all(R[..., 1] >= 0), c
        assert np.all((st["LCC"] >= 1) & (st["LCC"] <= 12)), c
    ref = _batch(gwcapi, gpu_ctx, cells, (1, 64, 4), 1)
    ref.step(10000)
    g1 = np.array([g.as_array() for g in ref.get_globals()])
    ref.close()
    g2 = np.array(snaps[0][0])
    # the same dynamics: only the rounding of the FP64 flux sums differs, and through it
    # the odd FP32 event decision
    dv = np.abs(g1[:, 0] - g2[:, 0])
    dca = np.abs(g1[:, 3] - g2[:, 3]) / g1[:, 3]
    say(f"  against one step per launch at 10 ms: |dV| <= {dv.max():.3g} mV, |dCai|/Cai <= {dca.max():.3g}")
    assert np.all(dv < 1.0) and np.all(dca < 0.05)


def test_two_steps_per_launch_applies_only_where_built(gwcapi, gpu_ctx):
    cells = [gwcapi.make_cell(500, 3, None, _pace(gwcapi))]
    b = gwcapi.Batch(gpu_ctx, cells)            # FP64: one step per launch
    b.set_mode(frozen=False, digest=False)
    b.set_steps_per_launch(2)
    assert b.steps_per_launch() == 1
    b.close()
    b = _batch(gwcapi, gpu_ctx, cells, (1, 64, 4), 2)
    b.set_launch(1, 128, 4)                     # 128-thread shapes: one step per launch
    assert b.steps_per_launch() == 1
    b.set_launch(1, 64, 4)
    b.set_graph(16)                             # replayed graphs: one step per launch
    assert b.steps_per_launch() == 1
    b.set_graph(0)
    assert b.steps_per_launch() == 2
    with pytest.raises(gwcapi.GwError):
        b.set_steps_per_launch(3)
    b.close()


def test_autotune_restores_the_state_and_picks_the_steps_per_launch(gwcapi, gpu_ctx):
    """The autotuner times every FP32-local shape (the one-lane shapes of 64
    threads with two steps per launch), then its pick with the other number of
    steps per launch where it has both, and leaves the batch's state as it
    found it."""
    cells = [gwcapi.make_cell(2000, 10 + c, None, _pace(gwcapi)) for c in range(4)]
    a = gwcapi.Batch(gpu_ctx, cells, precision=gwcapi.PREC_FP32_LOCAL)
    a.set_mode(frozen=False, digest=False)
    a.step(300)
    before = _snapshot(a, len(cells))
    shape, ms = a.autotune(40)
    spl = a.steps_per_launch()
    say(f"  autotune chose {shape}, {spl} steps per launch: {ms}")
    assert a.launch() == shape
    assert spl == 1 or shape[:2] == (1, 64)
    assert all(ms[k] is not None and ms[k] > 0 for k in ms if k.startswith("1x"))
    _assert_same(_snapshot(a, len(cells)), before, "autotune")
    a.close()


def test_two_steps_per_launch_at_the_extremes_of_the_cell_size(gwcapi, gpu_ctx):
    """Cells of 1, 65 and 2^20 CRUs in one batch: with one CRU per lane two
    steps per launch give the same bits as one; with two and eight CRUs per
    lane they are deterministic and keep every channel count valid."""
    cells = [gwcapi.make_cell(n, 300 + k, None, _pace(gwcapi)) for k, n in enumerate((1, 65, 1 << 20))]
    one = _batch(gwcapi, gpu_ctx, cells, (1, 64, 1), 1)
    two = _batch(gwcapi, gpu_ctx, cells, (1, 64, 1), 2)
    for chunk in (3, 500):
        one.step(chunk)
        two.step(chunk)
    say("  extremes, one CRU per lane: 503 steps")
    _assert_same(_snapshot(one, len(cells)), _snapshot(two, len(cells)), "extremes")
    one.close()
    two.close()
    for cpl in (2, 8):
        snaps = []
        for run in range(2):
            b = _batch(gwcapi, gpu_ctx, cells, (1, 64, cpl), 2)
            b.step(1003)
            snaps.append(_snapshot(b, len(cells)))
            b.close()
        _assert_same(snaps[0], snaps[1], f"run to run, {cpl} CRUs per lane")
        for c, st in enumerate(snaps[0][1]):
            R = st["RyR"]
            assert np.all(R.sum(axis=2) == 5), (cpl, c)
            assert np.all((st["LCC"] >= 1) & (st["LCC"] <= 12)), (cpl, c)
        say(f"  extremes, {cpl} CRUs per lane: deterministic")


def test_steps_per_launch_automatic(gwcapi, gpu_ctx):
    """By default (0, automatic) a batch runs two steps per launch when its SSA
    blocks are more than the GPU holds at once, one otherwise."""
    small = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(12500, 7 + c, None, _pace(gwcapi)) for c in range(4)],
                         precision=gwcapi.PREC_FP32_LOCAL)
    small.set_launch(1, 64, 4)
    assert small.steps_per_launch() == 1          # 196 blocks: one round
    small.close()
    big = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(12500, 7 + c, None, _pace(gwcapi)) for c in range(100)],
                       precision=gwcapi.PREC_FP32_LOCAL)
    big.set_mode(frozen=False, digest=False)
    big.set_launch(1, 64, 4)
    assert big.steps_per_launch() == 2            # 4,900 blocks of 256 CRUs: more than one round
    big.step(3)
    big.set_steps_per_launch(1)
    assert big.steps_per_launch() == 1
    big.set_steps_per_launch(0)
    assert big.steps_per_launch() == 2
    big.close()
