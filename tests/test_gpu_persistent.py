"""The persistent step kernel: one launch per gw_batch_step call for batches
whose blocks can all be resident, with the CRUs held in registers and the
per-step exchange between each cell's SSA blocks and its close block done
through per-cell counters. Its results are bit-identical to one launch per
step and to the oracle; it is refused for a batch that does not fit; and a
bounded wait that expires (forced from the host by a tiny poll limit) ends the
call with GW_E_TIMEOUT while every block exits normally."""
import numpy as np
import pytest

from oracle.oracle import OracleModel

from .test_gpu_parity import assert_globals_equal, assert_states_equal, say

pytestmark = pytest.mark.gpu


def _pace(gwcapi):
    return {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}


def _run(gwcapi, ctx, cells, precision, persistent, steps, digest):
    b = gwcapi.Batch(ctx, cells, precision=precision)
    b.set_mode(frozen=False, digest=digest)
    b.set_persistent(persistent)
    assert b.persistent() == (persistent != 0)
    for chunk in (1, 999, steps - 1000):   # several calls: the kernel resumes from a closed state
        b.step(chunk)
    out = ([g.as_array() for g in b.get_globals()], [b.get_cru_state(c) for c in range(len(cells))],
           [b.digest(c) for c in range(len(cells))] if digest else None, b.obs(), b.time(), b.iterations())
    b.close()
    return out


@pytest.mark.parametrize("precision,digest", [(0, True), (0, False), (1, False)], ids=["fp64-digest", "fp64", "fp32-local"])
def test_persistent_matches_launch_per_step(gwcapi, gpu_ctx, precision, digest):
    cells = [gwcapi.make_cell(n, 60 + n, None, _pace(gwcapi)) for n in (37, 300, 1250)]
    say(f"persistent vs per-step, precision {precision}, digest {digest}")
    a = _run(gwcapi, gpu_ctx, cells, precision, 1, 4000, digest)
    b = _run(gwcapi, gpu_ctx, cells, precision, 0, 4000, digest)
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)
    for x, y in zip(a[1], b[1]):
        assert_states_equal(x, y)
    if digest:
        for (c1, h1), (c2, h2) in zip(a[2], b[2]):
            assert np.array_equal(c1, c2) and np.array_equal(h1, h2)
    for o1, o2 in zip(a[3], b[3]):
        assert o1 == o2
    assert a[4] == b[4] and a[5] == b[5]


def test_persistent_full_size_cell_matches_the_oracle(gwcapi, gpu_ctx):
    n, steps = 12500, 3000
    say(f"persistent {n}-CRU cell vs oracle")
    b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(n, 77, None, _pace(gwcapi))])
    b.set_mode(frozen=False, digest=False)
    b.set_persistent(1)
    b.step(steps)
    g, st = b.get_globals()[0], b.get_cru_state(0)
    b.close()
    ist = np.zeros(steps)
    t = 0.0
    for k in range(steps):
        ist[k] = 100.0 if t < 1.0 else 0.0
        t += 1e-3
    m = OracleModel(n, seed=77, kind="par")
    m.sample_rest_state()
    m.run(1e-3, steps, ist)
    assert_globals_equal(g, m.globals())
    assert_states_equal(st, m.state())


def test_persistent_is_refused_for_a_batch_that_does_not_fit(gwcapi, gpu_ctx):
    cells = [gwcapi.make_cell(12500, s, None, _pace(gwcapi)) for s in range(64)]
    b = gwcapi.Batch(gpu_ctx, cells)
    b.set_mode(frozen=False, digest=False)
    assert not b.persistent()     # off by default
    b.set_persistent(2)
    assert not b.persistent()     # "when it fits" leaves it off
    with pytest.raises(gwcapi.GwError) as e:
        b.set_persistent(1)
    assert gwcapi.STATUS[e.value.code] == "GW_E_INVALID_ARG"
    b.close()


def test_an_expired_wait_ends_the_call_with_timeout(gwcapi, gpu_ctx):
    cells = [gwcapi.make_cell(2000, 5, None, _pace(gwcapi))]
    b = gwcapi.Batch(gpu_ctx, cells)
    b.set_mode(frozen=False, digest=False)
    b.set_persistent(1)
    b.set_spin_limit(1)   # the close block's first wait for the SSA blocks gives up at once
    with pytest.raises(gwcapi.GwError) as e:
        b.step(500)
    assert gwcapi.STATUS[e.value.code] == "GW_E_TIMEOUT"
    b.close()
    # the device and the library carry on
    c = gwcapi.Batch(gpu_ctx, cells)
    c.set_mode(frozen=False, digest=False)
    c.step(100)
    assert np.isfinite(c.get_globals()[0].V)
    c.close()
