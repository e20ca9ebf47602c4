"""Per-cell overrides of a running batch, and what the host reads back when a
cell has stopped.

After a batch has been stepped, gw_batch_set_globals / set_cru_state /
set_stimulus_table change only the named cell: it keeps its clock and step
index and continues exactly as the oracle does when the same change is made to
it between two Euler steps (the oracle keeps its Jxfer/Jtr arrays, as the
reference's model does, GW.tpp:47-62); every other cell continues
bit-identical to a batch that was never touched.
"""
import importlib
import math
import os
import sys

import numpy as np
import pytest

from oracle.oracle import OracleModel

from .test_gpu_parity import assert_globals_equal, assert_states_equal, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = (40, 96, 64)
SEEDS = (21, 22, 23)
FIRST, SECOND = 1200, 900


def _pace(gwcapi):
    return {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}


def _paced_table(steps, t0=0.0):
    out = np.zeros(steps)
    t = t0
    for k in range(steps):
        out[k] = 100.0 if math.fmod(t, 1000.0) < 1.0 else 0.0
        t += 1e-3
    return out


def _batch(gwcapi, ctx):
    b = gwcapi.Batch(ctx, [gwcapi.make_cell(n, s, None, _pace(gwcapi)) for n, s in zip(N, SEEDS)])
    b.set_mode(frozen=False, digest=True)
    return b


def _untouched(gwcapi, ctx):
    b = _batch(gwcapi, ctx)
    b.step(FIRST + SECOND)
    out = ([g.as_array() for g in b.get_globals()], [b.get_cru_state(c) for c in range(3)])
    b.close()
    return out


def _oracle_first_leg(cell):
    m = OracleModel(N[cell], seed=SEEDS[cell])
    m.sample_rest_state()
    m.run(1e-3, FIRST, _paced_table(FIRST))
    return m


def _second_leg_table():
    # the clock after FIRST accumulated steps, as the batch and the oracle accumulate it
    t = 0.0
    for _ in range(FIRST):
        t += 1e-3
    return _paced_table(SECOND, t)


def _check_others(b, want_g, want_s, touched):
    got = b.get_globals()
    for c in range(3):
        if c == touched:
            continue
        assert np.array_equal(got[c].as_array(), want_g[c]), c
        assert_states_equal(b.get_cru_state(c), want_s[c])


def test_mid_run_globals_change_only_that_cell(gwcapi, gpu_ctx):
    say("mid-run set_globals")
    want_g, want_s = _untouched(gwcapi, gpu_ctx)
    b = _batch(gwcapi, gpu_ctx)
    b.step(FIRST)
    g = b.get_globals()[1]
    g.Cai *= 1.5
    g.V += 3.0
    b.set_globals(1, g)
    t_mid, s_mid = b.time()
    assert s_mid == FIRST
    b.step(SECOND)
    _check_others(b, want_g, want_s, 1)
    m = _oracle_first_leg(1)
    go = m.globals()
    go.Cai *= 1.5
    go.V += 3.0
    m.set_state(globals_=go)
    m.run(1e-3, SECOND, _second_leg_table())
    assert_globals_equal(b.get_globals()[1], m.globals())
    assert_states_equal(b.get_cru_state(1), m.state())
    assert b.time()[1] == FIRST + SECOND
    b.close()


def test_mid_run_cru_state_changes_only_that_cell(gwcapi, gpu_ctx):
    say("mid-run set_cru_state")
    want_g, want_s = _untouched(gwcapi, gpu_ctx)
    b = _batch(gwcapi, gpu_ctx)
    b.step(FIRST)
    st = b.get_cru_state(1)
    st["RyR"][:, 0, :] = [2, 0, 3, 0, 0, 0]   # open RyRs in subunit 0 of every CRU
    st["CaSS"][:, 0] *= 2.0
    b.set_cru_state(1, st["CaSS"], st["CaJSR"], st["LCC"], st["LCC_inactivation"], st["RyR"], st["ClCh"], st["JLCC"])
    b.step(SECOND)
    _check_others(b, want_g, want_s, 1)
    m = _oracle_first_leg(1)
    m.set_state(CaSS=st["CaSS"], CaJSR=st["CaJSR"], LCC=st["LCC"], LCC_inactivation=st["LCC_inactivation"],
                RyR=st["RyR"], ClCh=st["ClCh"], JLCC=st["JLCC"])
    m.run(1e-3, SECOND, _second_leg_table())
    assert_globals_equal(b.get_globals()[1], m.globals())
    assert_states_equal(b.get_cru_state(1), m.state())
    b.close()


def test_mid_run_stimulus_table_changes_only_that_cell(gwcapi, gpu_ctx):
    say("mid-run set_stimulus_table")
    want_g, want_s = _untouched(gwcapi, gpu_ctx)
    b = _batch(gwcapi, gpu_ctx)
    b.step(FIRST)
    table = np.zeros(FIRST + SECOND)
    table[FIRST:FIRST + 700] = 60.0    # entry k is the stimulus of step k
    b.set_stimulus_table(2, table)
    b.step(SECOND)
    _check_others(b, want_g, want_s, 2)
    m = _oracle_first_leg(2)
    m.run(1e-3, SECOND, table[FIRST:])
    assert_globals_equal(b.get_globals()[2], m.globals())
    assert_states_equal(b.get_cru_state(2), m.state())
    b.close()


def test_overrides_before_the_first_step_set_the_starting_point(gwcapi, gpu_ctx):
    """Before any step an override is a new starting point at step 0, as the
    reference's set_initial_value before its first euler_step."""
    say("set_globals before stepping")
    b = _batch(gwcapi, gpu_ctx)
    g = b.get_globals()[0]
    g.V = -60.0
    b.set_globals(0, g)
    assert b.time() == (0.0, 0)
    b.step(500)
    m = OracleModel(N[0], seed=SEEDS[0])
    m.sample_rest_state()
    go = m.globals()
    go.V = -60.0
    m.set_state(globals_=go)
    m.run(1e-3, 500, _paced_table(500))
    assert_globals_equal(b.get_globals()[0], m.globals())
    b.close()


def test_clamp_cell_refuses_a_stimulus_table(gwcapi, gpu_ctx):
    clamp = {"kind": gwcapi.PROTO_CLAMP, "hold": -80.0, "step": 0.0, "t_on": 1.0, "t_off": 5.0}
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(16, 1, None, clamp), gwcapi.make_cell(16, 2, None, _pace(gwcapi))])
    with pytest.raises(gwcapi.GwError) as e:
        b.set_stimulus_table(0, np.zeros(10))
    assert gwcapi.STATUS[e.value.code] == "GW_E_INVALID_ARG" and "clamp" in str(e.value)
    b.set_stimulus_table(1, np.zeros(10))   # a paced cell may take one
    b.close()


def test_trace_and_clock_when_a_cell_stops(gwcapi, gpu_ctx):
    """Cell 0 stops mid-run on a non-positive concentration: the clock read
    back is the running cells' one, the trace holds every row of the running
    cells, and cell 0's rows after its stop read NaN."""
    say("trace with a stopped cell")
    b = _batch(gwcapi, gpu_ctx)
    b.set_mode(frozen=False, digest=False)
    b.enable_trace(10, 400)
    b.step(500)
    g = b.get_globals()[0]
    g.Nai = -1.0
    b.set_globals(0, g)
    with pytest.raises(gwcapi.GwError) as e:
        b.step(1000)
    assert "cell 0" in str(e.value) and gwcapi.STATUS[e.value.code] == "GW_E_NONPOSITIVE_CONC"
    t, steps = b.time()
    assert steps == 1500 and abs(t - 1.5) < 1e-9
    tr = b.get_trace()
    assert tr.shape == (150, 3, 4)
    assert np.all(np.isfinite(tr[:, 1:, :]))
    assert np.all(np.isfinite(tr[:50, 0, :])) and np.all(np.isnan(tr[50:, 0, :]))
    b.close()


@pytest.fixture(scope="module")
def gw():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    importlib.import_module("gw-model_amd")
    import build.GreensteinWinslow as m  # resolved through the package's sys.path entry

    return m


def test_pybind_run_in_the_fp32_local_form(gw):
    """GreensteinWinslow.run(precision="fp32_local") runs the variant from the
    drop-in surface; its channel states stay valid and its voltage tracks the
    FP64 run's through the upstroke."""
    n, steps = 200, 3000
    c = gw.rest_state(n, 9)
    stim = lambda t: 100.0 if t < 1.0 else 0.0  # noqa: E731
    a = gw.run(gw.GWParameters(), n, 1e-3, steps, stim, 500, c, seed=9)
    f = gw.run(gw.GWParameters(), n, 1e-3, steps, stim, 500, c, seed=9, precision="fp32_local")
    assert np.all((f.LCC >= 1) & (f.LCC <= 12))
    ryr = np.asarray(f.RyR)
    assert np.all(ryr.sum(axis=-1) == 5)
    assert np.all(ryr[..., :2] >= 0)   # states 1 and 2 are never in a lumped pair
    assert np.max(np.abs(np.asarray(a.V) - np.asarray(f.V))) < 15.0
    assert np.max(np.asarray(f.V)) > 0.0   # depolarised by the stimulus
