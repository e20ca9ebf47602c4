"""Observables formed on the device, against the oracle and the post-hoc
definitions of the reference.

* The APD90 / alternans tracker of the close (gw_cell_obs_t) against the same
  definition applied on the host to the oracle's own voltage trace, step by
  step, over two beats at 2 Hz: the device and the oracle follow the same
  trajectory bit for bit, so the tracker's upstroke and repolarisation times,
  APD90 of both beats and the alternans magnitude agree to rounding.
* gw_batch_get_currents (the post-hoc ICaL and Ito2 of src/model.py:414-481,
  formed on the device) against the package's host restatement of those
  functions on the same state.
* The integrator's own I_CaL (each channel's flux frozen at its opening,
  src/includes/SSA.tpp:141) against the post-hoc one: with the RyRs off and a
  tiny L-type permeability, CaSS stays at its value of every channel's opening
  under a voltage clamp, and the two definitions reduce to the same closed form.
"""
import ctypes as C
import importlib
import os
import sys

import numpy as np
import pytest

from oracle.oracle import GwRecord, OracleModel

from .test_gpu_parity import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def track_apd(t, V):
    """APD90 and alternans by the definition of gw_cell_obs_t: upstroke at the
    upward crossing of -40 mV, end at the downward crossing of
    V_rest + 0.1 (V_peak - V_rest) below V_peak - 10, both interpolated linearly
    within a step; alternans = |APD90_k - APD90_{k-1}| over completed beats."""
    v_rest = v_peak = V[0]
    t_up, last, prev, alt_sum, alt_n, beats = -1.0, -1.0, -1.0, 0.0, 0, 0
    apds = []
    for n in range(1, len(V)):
        v0, v1, t0 = V[n - 1], V[n], t[n - 1]
        dt = t[n] - t0
        if t_up < 0.0:
            v_rest = min(v_rest, v1)
            if v0 < -40.0 <= v1:
                t_up = t0 + dt * (-40.0 - v0) / (v1 - v0)
                v_peak = v1
        else:
            v_peak = max(v_peak, v1)
            thr = v_rest + 0.1 * (v_peak - v_rest)
            if v0 > thr >= v1 and v1 < v_peak - 10.0:
                apd = t0 + dt * (v0 - thr) / (v0 - v1) - t_up
                prev, last = last, apd
                if beats >= 1:
                    alt_sum += abs(apd - prev)
                    alt_n += 1
                beats += 1
                apds.append(apd)
                t_up = -1.0
                v_rest = v1
    return apds, alt_sum, alt_n


def test_apd_tracker_matches_the_oracle_trace(gwcapi, gpu_ctx):
    n, steps, period = 100, 1_000_000, 500.0
    say(f"APD tracker: {n} CRUs, two beats at 2 Hz against the oracle's trace")
    pace = {"kind": gwcapi.PROTO_PACE, "period": period, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 123, None, pace)])
    b.set_mode(frozen=False, digest=False)
    b.step(steps)
    o = b.obs()[0]
    V_end = b.get_globals()[0].V
    b.close()
    # the oracle's voltage after every step, and its accumulated clock
    ist = np.zeros(steps)
    t = 0.0
    for k in range(steps):
        ist[k] = 100.0 if (t % period) < 1.0 else 0.0
        t += 1e-3
    m = OracleModel(n, seed=123, kind="par")
    m.sample_rest_state()
    V0 = m.globals().V
    rec = m.run(1e-3, steps, ist, record_every=1, fields=("t", "V"))
    assert rec["V"][-1] == V_end   # the same trajectory, bit for bit
    tt = np.concatenate([[0.0], rec["t"]])
    VV = np.concatenate([[V0], rec["V"]])
    apds, alt_sum, alt_n = track_apd(tt, VV)
    say(f"  device APD90 {o['apd90_prev']:.4f}, {o['apd90_last']:.4f} ms, alternans {o['alt_sum']:.4f} ms; "
        f"oracle trace {apds}")
    assert len(apds) == 2 and o["beats_done"] == 2
    assert o["apd90_prev"] == pytest.approx(apds[0], abs=1e-9)
    assert o["apd90_last"] == pytest.approx(apds[1], abs=1e-9)
    assert o["alt_count"] == alt_n == 1 and o["alt_sum"] == pytest.approx(alt_sum, abs=1e-9)
    assert o["peak_V"] == np.max(VV[1:]) and o["min_V"] == np.min(VV[1:])


@pytest.fixture(scope="module")
def pkg():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    return importlib.import_module("gw-model_amd")


def _host_currents(pkg, st, V, n):
    p = pkg.GWParameters(NCaRU_sim=n)
    ical = pkg.ICaL(np.array([V]), st["LCC"][None], st["LCC_inactivation"][None], st["CaSS"][None], p)[0]
    ito2 = pkg.Ito2(np.array([V]), st["ClCh"][None], p)[0]
    return ical, ito2


def test_device_posthoc_currents_match_the_host_definitions(gwcapi, gpu_ctx, pkg):
    cells = [gwcapi.make_cell(n, 40 + n, None, {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0,
                                                 "amplitude": 100.0, "start": 0.0}) for n in (300, 2000)]
    b = gwcapi.Batch(gpu_ctx, cells)
    b.set_mode(frozen=False, digest=False)
    checked = 0
    for steps in (1500, 3000, 20000):   # the upstroke, the release, the plateau
        b.step(steps)
        ical, ito2 = b.currents()
        g = b.get_globals()
        for c, cell in enumerate(cells):
            hi, ht = _host_currents(pkg, b.get_cru_state(c), g[c].V, cell.n_cru)
            say(f"  t {b.time()[0]:.1f} ms cell {c}: ICaL {ical[c]:.6g} (host {hi:.6g}), Ito2 {ito2[c]:.6g} (host {ht:.6g})")
            assert ical[c] == pytest.approx(hi, rel=1e-12, abs=1e-15)
            assert ito2[c] == pytest.approx(ht, rel=1e-12, abs=1e-15)
            checked += ical[c] != 0.0
    assert checked >= 3
    b.close()


def test_in_loop_and_posthoc_ical_agree_at_constant_subspace_calcium(gwcapi, gpu_ctx):
    n = 60
    p = gwcapi.Params.default()
    p.rRyR = 0.0          # no release
    p.PCaL *= 1e-6        # a vanishing L-type flux: CaSS stays where it started
    clamp = {"kind": gwcapi.PROTO_CLAMP, "hold": 10.0, "step": 10.0, "t_on": 0.0, "t_off": 1e9}
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 9, p, clamp)])
    b.set_mode(frozen=False, digest=False)
    st = b.get_cru_state(0)
    cai = b.get_globals()[0].Cai
    b.set_cru_state(0, np.full((n, 4), cai), st["CaJSR"], st["LCC"], st["LCC_inactivation"], st["RyR"], st["ClCh"])
    b.enable_trace(1, 8)
    compared = 0
    for k in range(3000):
        b.step(1)
        in_loop = b.get_trace()[-1, 0, 2]
        post, _ = b.currents()
        if post[0] != 0.0:
            assert in_loop == pytest.approx(post[0], rel=1e-6), k
            compared += 1
        else:
            assert in_loop == 0.0
    say(f"  in-loop vs post-hoc I_CaL compared at {compared} steps with open channels")
    assert compared >= 200
    b.close()
