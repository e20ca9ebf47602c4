"""Pin the oracle's PRNG, uniform mapping, samplers and common.hpp functions
against golden vectors produced by the reference's own xoshiro.hpp and
common.hpp (tests/golden/make_golden.py via oracle/_ref)."""
import ctypes as C

import numpy as np
import pytest


def _u64(h):
    return int(h, 16)


def test_splitmix64(oracle_lib, golden):
    for seed_hex, outs in golden["splitmix64"].items():
        st = C.c_uint64(_u64(seed_hex))
        got = [oracle_lib.gwo_splitmix64_next(C.byref(st)) for _ in outs]
        assert got == [_u64(h) for h in outs], seed_hex


@pytest.mark.parametrize("prng", range(6))
def test_engines_raw_outputs(oracle_lib, golden, prng):
    for seed_hex, outs in golden["raw"][str(prng)].items():
        s = np.zeros(4, np.uint64)
        oracle_lib.gwo_rng_seed(prng, s.ctypes.data, _u64(seed_hex))
        got = [oracle_lib.gwo_rng_next(prng, s.ctypes.data) for _ in outs]
        assert got == [_u64(h) for h in outs], (prng, seed_hex)


def test_known_answer_xoshiro256ss_seed42(oracle_lib):
    # Xoshiro256StarStar(42), first three outputs (SplitMix64 seeding of xoshiro.hpp:684-703).
    s = np.zeros(4, np.uint64)
    oracle_lib.gwo_rng_seed(0, s.ctypes.data, 42)
    got = [oracle_lib.gwo_rng_next(0, s.ctypes.data) for _ in range(3)]
    assert got == [0x15780B2E0C2EC716, 0x6104D9866D113A7E, 0xAE17533239E499A1]


@pytest.mark.parametrize("prng", [0, 1])
def test_uniform_stream(oracle_lib, golden, prng):
    for seed_hex, outs in golden["uniform"][str(prng)].items():
        s = np.zeros(4, np.uint64)
        oracle_lib.gwo_rng_seed(prng, s.ctypes.data, _u64(seed_hex))
        got = [oracle_lib.gwo_u01(oracle_lib.gwo_rng_next(prng, s.ctypes.data)) for _ in outs]
        assert got == [float.fromhex(h) for h in outs]


def test_uniform_map_special_values(oracle_lib, golden):
    raw = golden["urand"]["raw"]
    want = [float.fromhex(h) for h in golden["urand"]["u"]]
    got = [oracle_lib.gwo_u01(_u64(h)) for h in raw]
    assert got == want
    # closed-form checks of the mapping: u = round(x)/2^64, clamped below 1
    assert oracle_lib.gwo_u01(0) == 0.0
    assert oracle_lib.gwo_u01(2**64 - 1) == float.fromhex("0x1.fffffffffffffp-1")
    assert all(0.0 <= u < 1.0 for u in got)


def test_sample_weights(oracle_lib, golden):
    for case in golden["sample_weights"]:
        w = np.array([float.fromhex(h) for h in case["w"]])
        total = float.fromhex(case["total"])
        for raw_h, idx in zip(case["raw"], case["idx"]):
            u = oracle_lib.gwo_u01(_u64(raw_h))
            assert oracle_lib.gwo_sample_weights(w.ctypes.data, total, len(w), u) == idx


def test_sample_binomial(oracle_lib, golden):
    for case in golden["sample_binomial"]:
        raw = np.array([_u64(h) for h in case["raw"]], dtype=np.uint64)
        X = C.c_int(-1)
        used = oracle_lib.gwo_sample_binomial_replay(float.fromhex(case["p"]), case["N"], raw.ctypes.data, C.byref(X))
        assert (X.value, used) == (case["X"], case["used"])


def _ulps(a, b):
    a, b = float(a), float(b)
    if a == b:
        return 0.0
    return abs(a - b) / np.spacing(max(abs(a), abs(b)))


# common.hpp functions: the oracle uses its own libm (exp/log/pow within 2 ulp of
# glibc), so agreement is to a few ulp, not bitwise. A dropped term, a wrong sign
# or a swapped argument moves these by many orders of magnitude more.
GATE_ULPS = 8


@pytest.mark.parametrize("name", ["alpha_m", "beta_m", "alpha_h", "beta_h", "alpha_j", "beta_j"])
def test_gates_match_reference(oracle_lib, golden, name):
    fn = getattr(oracle_lib, "gwo_" + name)
    for row in golden["common_V"]:
        V = float.fromhex(row["V"])
        want = float.fromhex(row[name])
        got = fn(V)
        if want == 0.0:
            assert got == 0.0
        else:
            assert _ulps(got, want) <= GATE_ULPS, (name, V, got, want)


def test_pumps_exchanger_INa_match_reference(oracle_lib, golden):
    F_RT = 96.5 / (8.314 * 310.0)
    for row in golden["common_V"]:
        V = float.fromhex(row["V"])
        vf = V * F_RT
        e = float.fromhex(row["e"])
        assert _ulps(oracle_lib.gwo_fNaK(vf, e, 0.9), float.fromhex(row["fNaK3"])) <= GATE_ULPS
        assert _ulps(oracle_lib.gwo_INaK(vf, e, 10.3, 0.9, 10.0, 0.65), float.fromhex(row["INaK6"])) <= GATE_ULPS
        got = oracle_lib.gwo_INaCa(vf, e, 10.3, 1.6e-4, 138.0**3, 2.0, 0.35, 0.2, 3.7e-4)
        want = float.fromhex(row["INaCa9"])
        # the numerator is a difference of two terms: compare relative to their scale
        scale = abs(np.exp(0.35 * vf) * 10.3**3 * 2.0) * 3.7e-4
        assert abs(got - want) <= 1e-14 * scale, (V, got, want)
        assert got == pytest.approx(want, rel=1e-12, abs=1e-16)
        assert oracle_lib.gwo_INa(V, 0.3, 0.5, 0.6, 12.8, 70.0) == float.fromhex(row["INa"])
        assert oracle_lib.gwo_Ib(V, 0.00264, 70.0) == float.fromhex(row["Ib"])


def test_nernst_and_ipca_match_reference(oracle_lib, golden):
    RT_F = 8.314 * 310.0 / 96.5
    ok = C.c_int(1)
    for row in golden["common_conc"]:
        Xi = float.fromhex(row["Xi"])
        for key, Xo, z in (("Nernst1", 138.0, 1.0), ("Nernst2", 2.0, 2.0)):
            got = oracle_lib.gwo_Nernst(Xi, Xo, RT_F, z, C.byref(ok))
            want = float.fromhex(row[key])
            if abs(want) < 1e-9:
                assert abs(got) < 1e-12
            else:
                assert _ulps(got, want) <= GATE_ULPS
        assert oracle_lib.gwo_IpCa(Xi * 1e-4, 0.03, 0.0005) == float.fromhex(row["IpCa"])
    # Nernst with a non-positive concentration is an error status, not exit()
    bad = C.c_int(1)
    oracle_lib.gwo_Nernst(0.0, 138.0, RT_F, 1.0, C.byref(bad))
    assert bad.value == 0
