"""Pin the oracle's own exp/log/pow against numpy/glibc and exact identities.

The oracle (and the GPU kernels) use their own libm so that CPU and GPU agree
bit for bit; here it is held to within 2 ulp of the platform libm."""
import math

import numpy as np
import pytest


def _ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return np.where(a == b, 0.0, d)


def test_exp_sweep(oracle_lib):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-745, 709.7, 4000), rng.uniform(-1, 1, 4000), rng.uniform(-1e-8, 1e-8, 500),
                         np.array([0.0, -0.0, 1.0, -1.0, 0.5, math.log(2), -math.log(2), 700.0, -700.0, -740.0])])
    got = np.array([oracle_lib.gwo_exp(float(x)) for x in xs])
    want = np.exp(xs)
    normal = want > 2.3e-308
    assert _ulps(got[normal], want[normal]).max() <= 2.0
    # subnormal results: absolute error below the smallest subnormal spacing times 2
    assert np.all(np.abs(got[~normal] - want[~normal]) <= 2 * 5e-324)


def test_exp_special(oracle_lib):
    assert oracle_lib.gwo_exp(0.0) == 1.0
    assert oracle_lib.gwo_exp(1000.0) == math.inf
    assert oracle_lib.gwo_exp(-1000.0) == 0.0
    assert math.isnan(oracle_lib.gwo_exp(math.nan))


def test_log_sweep(oracle_lib):
    rng = np.random.default_rng(2)
    xs = np.concatenate([10 ** rng.uniform(-300, 300, 4000), rng.uniform(0.5, 2.0, 4000),
                         1 + rng.uniform(-1e-6, 1e-6, 500), rng.uniform(0, 1, 2000),
                         np.array([5e-324, 1e-310, 2.2250738585072014e-308, 1.0, 2.0, 0.5, math.e])])
    got = np.array([oracle_lib.gwo_log(float(x)) for x in xs])
    want = np.log(xs)
    assert _ulps(got, want).max() <= 2.0


def test_log_special_and_identity(oracle_lib):
    assert oracle_lib.gwo_log(1.0) == 0.0
    assert oracle_lib.gwo_log(0.0) == -math.inf
    assert math.isnan(oracle_lib.gwo_log(-1.0))
    # log of powers of two is exact up to the rounding of k*ln2
    for k in (-1000, -52, -1, 1, 10, 1000):
        assert _ulps(oracle_lib.gwo_log(2.0**k), k * math.log(2.0)) <= 1
    # -log(u) of the uniform map's extremes (waiting times of the SSA)
    umax = float.fromhex("0x1.fffffffffffffp-1")
    assert -oracle_lib.gwo_log(umax) == pytest.approx(2.0**-53, rel=1e-15)


def test_pow_uses(oracle_lib):
    # exponents the model uses: 1.5 (INaK) and Hf = Hr = 0.75 (Jup). pow is
    # exp(y*log(x)); the rounding of y*log(x) is amplified by |y ln x|, so the
    # bound is 2 + 2|y ln x| ulp.
    rng = np.random.default_rng(3)
    for y in (1.5, 0.75):
        xs = rng.uniform(1e-3, 10, 2000)
        got = np.array([oracle_lib.gwo_pow(float(x), y) for x in xs])
        bound = 2.0 + 2.0 * np.abs(y * np.log(xs))
        assert np.all(_ulps(got, xs**y) <= bound)
    assert oracle_lib.gwo_pow(4.0, 1.5) == 8.0 or _ulps(oracle_lib.gwo_pow(4.0, 1.5), 8.0) <= 2
    assert oracle_lib.gwo_pow(0.0, 0.75) == 0.0
    assert math.isnan(oracle_lib.gwo_pow(-1.0, 0.75))


def test_fixed_point_sum_is_exact_and_order_free(oracle_lib):
    import ctypes as C

    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(0, 1, 3000) * 10 ** rng.uniform(-8, 3, 3000), [1e-30, -1e-30, 0.0]])

    def fx_sum(vals):
        acc = np.zeros(2, np.uint64)
        flag = C.c_int(0)
        for v in vals:
            oracle_lib.gwo_fx_add(acc.ctypes.data, float(v), C.byref(flag))
        assert flag.value == 0
        return oracle_lib.gwo_fx_to_double(acc.ctypes.data)

    s1 = fx_sum(x)
    s2 = fx_sum(x[::-1])
    s3 = fx_sum(rng.permutation(x))
    assert s1 == s2 == s3
    assert s1 == math.fsum(x)  # correctly rounded exact sum (all terms are multiples of 2^-80 here but the tiny ones)
    # out-of-range term raises the flag
    acc = np.zeros(2, np.uint64)
    flag = C.c_int(0)
    oracle_lib.gwo_fx_add(acc.ctypes.data, 2.0**30, C.byref(flag))
    assert flag.value != 0
