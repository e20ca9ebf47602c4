"""Records in the reduced encoding of src/serialization.py:11-39, formed on the
device, against the same encoding applied on the host to full records."""
import importlib

import numpy as np
import pytest

from .test_gpu_parity import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return importlib.import_module("gw-model_amd")


def test_simulate_reduced_equals_reduced_simulate(pkg):
    """GWModel.simulate_reduced (GreensteinWinslow.run_reduced -> gw_run_reduced)
    returns exactly sim_dict_reduced_memory(GWModel.simulate(...)) for the same
    seed: keys, dtypes, shapes and every value, through an upstroke."""
    say("reduced records via GWModel")
    p = pkg.GWParameters(NCaRU_sim=400)

    def model():
        m = pkg.GWModel(parameters=p)
        m.set_stimulus_fn(lambda t: 100.0 if t < 1.0 else 0.0)
        return m

    full = pkg.sim_dict_reduced_memory(model().simulate(1e-3, 6000, record_every=500, PRNG="xoshiro256**", seed=11))
    red = model().simulate_reduced(1e-3, 6000, record_every=500, PRNG="xoshiro256**", seed=11)
    assert set(red) == set(full)
    for k in full:
        assert red[k].dtype == full[k].dtype, k
        assert red[k].shape == full[k].shape, k
        assert np.array_equal(red[k], full[k]), k
    assert red["RyR"].shape == (12, 400, 4)


def test_reduced_cru_state_at_scale(gwcapi, gpu_ctx, pkg):
    """gw_batch_get_cru_state_reduced of two paced 4,000-CRU cells after 20 ms
    (the release, with lumped RyR pairs) equals the host encoding of
    gw_batch_get_cru_state, in both forms."""
    pace = {"kind": gwcapi.PROTO_PACE, "period": 1000.0, "duration": 1.0, "amplitude": 100.0, "start": 0.0}
    for prec in (gwcapi.PREC_FP64, gwcapi.PREC_FP32_LOCAL):
        say(f"reduced CRU state, precision {prec}")
        b = gwcapi.Batch(gpu_ctx, [gwcapi.make_cell(4000, 21 + c, None, pace) for c in range(2)], precision=prec)
        b.set_mode(frozen=False, digest=False)
        b.step(20_000)
        for c in range(2):
            f = b.get_cru_state(c)
            r = b.get_cru_state_reduced(c)
            assert np.array_equal(r["CaSS"], f["CaSS"].astype(np.float32))
            assert np.array_equal(r["CaJSR"], f["CaJSR"].astype(np.float32))
            assert np.array_equal(r["LCC"], f["LCC"].astype(np.int8))
            assert np.array_equal(r["LCC_inactivation"], f["LCC_inactivation"].astype(np.bool_))
            assert np.array_equal(r["ClCh"], f["ClCh"].astype(np.bool_))
            assert np.array_equal(r["RyR"], pkg.RyR_state_to_int16(f["RyR"]))
            R = f["RyR"]
            neg = int(np.sum((R[..., 2] < 0) | (R[..., 3] < 0) | (R[..., 4] < 0) | (R[..., 5] < 0)))
            say(f"  cell {c}: subunits with a negative lumped member: {neg}")
        b.close()


def test_capi_run_reduced_matches_run(gwcapi, gpu_ctx, pkg):
    """gw_run_reduced through ctypes against gw_run plus the host encoding."""
    say("gw_run_reduced")
    p = gwcapi.Params.default()
    n, steps, every = 256, 3000, 700
    ist = np.where(np.cumsum(np.full(steps, 1e-3)) - 1e-3 < 1.0, 100.0, 0.0)
    full = gwcapi.run(gpu_ctx, p, n, 1e-3, steps, ist, every, None, None, seed=5)
    red = gwcapi.run_reduced(gpu_ctx, p, n, 1e-3, steps, ist, every, None, None, seed=5)

    class Sim:
        def to_dict(self):
            return {"vars": dict(full)}

    want = pkg.sim_dict_reduced_memory(Sim())
    for k in want:
        assert red[k].dtype == want[k].dtype and np.array_equal(red[k], want[k]), k


def test_gw_run_fp32_local_records(gwcapi, gpu_ctx, pkg):
    """gw_run and gw_run_reduced in the FP32-local form: the same seed gives
    the same run, so the reduced records formed on the device equal the host
    encoding of the full records; every recorded channel state is valid."""
    n, steps, rec = 1000, 20_000, 1000
    say("gw_run FP32-local records")
    istim = np.where(np.arange(steps) * 1e-3 < 1.0, 100.0, 0.0)
    args = (gpu_ctx, gwcapi.Params.default(), n, 1e-3, steps, istim, rec, gwcapi.default_globals(), None)
    full = gwcapi.run(*args, seed=31, precision=gwcapi.PREC_FP32_LOCAL)
    red = gwcapi.run_reduced(*args, seed=31, precision=gwcapi.PREC_FP32_LOCAL)
    R = full["RyR"]
    assert R.shape == (steps // rec, n, 4, 6)
    assert np.all(R.sum(axis=3) == 5)
    assert np.all((full["LCC"] >= 1) & (full["LCC"] <= 12))
    assert np.all(np.isfinite(full["CaSS"])) and np.all(full["CaSS"] > 0)
    assert full["V"].max() > 0.0   # the stimulus produced an upstroke

    class _Sim:
        def to_dict(self):
            return {"vars": {k: np.copy(v) for k, v in full.items()}}

    host = pkg.sim_dict_reduced_memory(_Sim())
    assert set(red) == set(host)
    for k in host:
        assert red[k].dtype == host[k].dtype, k
        assert np.array_equal(red[k], host[k]), k
