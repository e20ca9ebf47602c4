"""gw_run_noalloc: gw_run with no device allocation (device memory from the
library's static scratch array), called through the C-ABI.

Two calls in a row on one GPU, the second with a smaller cell and fewer steps
(the array keeps the first call's contents), each give records identical to
gw_run's, in both precisions; a run larger than the array is refused with
GW_E_INVALID_ARG and a message that names the limit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _istim(steps):
    out = np.zeros(steps)
    t = 0.0
    for k in range(steps):
        out[k] = 100.0 if t < 1.0 else 0.0
        t += 1e-3
    return out


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32_local"])
def test_noalloc_run_matches_gw_run_twice(gwcapi, gpu_ctx, precision):
    p = gwcapi.Params.default()
    for n, steps, every, seed in ((700, 2500, 500, 31), (90, 1200, 300, 32)):
        print(f"gw_run_noalloc precision {precision}: {n} CRUs x {steps} steps", flush=True)
        crus = gwcapi.rest_state(n, seed)
        ist = _istim(steps)
        want = gwcapi.run(gpu_ctx, p, n, 1e-3, steps, ist, every, None, crus, seed, 0, precision)
        again = gwcapi.run(gpu_ctx, p, n, 1e-3, steps, ist, every, None, crus, seed, 0, precision)
        got = gwcapi.run(gpu_ctx, p, n, 1e-3, steps, ist, every, None, crus, seed, 0, precision, noalloc=True)
        for name, other in (("gw_run again", again), ("gw_run_noalloc", got)):
            for k, v in want.items():
                if not np.array_equal(other[k], v):
                    bad = np.argwhere(other[k] != v)
                    print(f"  {name} {k}: {len(bad)} differ, first {bad[0].tolist()}: {other[k][tuple(bad[0])]!r} vs "
                          f"{v[tuple(bad[0])]!r}", flush=True)
        for k, v in want.items():
            assert got[k].shape == v.shape, k
            assert np.array_equal(again[k], v), ("gw_run is not reproducible", k)
            assert np.array_equal(got[k], v), k
        assert np.max(want["V"]) > 0.0   # the stimulus fired the cell: the records are not trivial


def test_noalloc_run_refuses_a_run_beyond_the_scratch(gwcapi, gpu_ctx):
    p = gwcapi.Params.default()
    n = 400_000   # about 160 MB of device state, beyond the 64 MiB array
    with pytest.raises(gwcapi.GwError) as e:
        gwcapi.run(gpu_ctx, p, n, 1e-3, 10, None, 10, None, None, 1, 0, 0, noalloc=True)
    assert gwcapi.STATUS[e.value.code] == "GW_E_INVALID_ARG"
    assert "MiB" in str(e.value)
    # and the array is still usable afterwards
    crus = gwcapi.rest_state(50, 3)
    a = gwcapi.run(gpu_ctx, p, 50, 1e-3, 200, _istim(200), 100, None, crus, 3, 0, 0)
    b = gwcapi.run(gpu_ctx, p, 50, 1e-3, 200, _istim(200), 100, None, crus, 3, 0, 0, noalloc=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
