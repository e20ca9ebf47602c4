"""Write golden vectors from the reference's own code.

Loads oracle/_ref/libgwref.so, which `make -C oracle ref` compiles from the
reference's src/includes/xoshiro.hpp and src/includes/common.hpp, and writes
ref_vectors.json next to this script. Run in a container that has the
reference tree:  python tests/golden/make_golden.py
The JSON holds inputs and outputs only.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle.oracle import ref_lib  # noqa: E402


def hexs(a):
    return [f"{int(x):016x}" for x in a]


def main():
    L = ref_lib()
    if L is None:
        raise SystemExit("oracle/_ref/libgwref.so missing: run `make -C oracle ref`")
    out = {"source": "reference src/includes/xoshiro.hpp + src/includes/common.hpp via oracle/ref_harness.cpp"}

    seeds = [0, 1, 42, 105, 0x8000000000000005, 0xFFFFFFFFFFFFFFFF, 42 + (7 << 32), 12345 + (12499 << 32)]
    n = 24
    sm = {}
    for s in seeds:
        a = np.zeros(n, np.uint64)
        L.ref_splitmix64(s, n, a.ctypes.data)
        sm[f"{s:016x}"] = hexs(a)
    out["splitmix64"] = sm

    raw = {}
    for prng in range(6):
        raw[str(prng)] = {}
        for s in seeds:
            a = np.zeros(n, np.uint64)
            L.ref_raw(prng, s, n, a.ctypes.data)
            raw[str(prng)][f"{s:016x}"] = hexs(a)
    out["raw"] = raw

    uni = {}
    for prng in (0, 1):
        uni[str(prng)] = {}
        for s in seeds:
            a = np.zeros(n)
            L.ref_uniform(prng, s, n, a.ctypes.data)
            uni[str(prng)][f"{s:016x}"] = [float(x).hex() for x in a]
    out["uniform"] = uni

    # urand over hand-picked raw values (rounding and the clamp below 1)
    special = [0, 1, (1 << 11) - 1, 1 << 11, (1 << 11) + 1, (1 << 53) - 1, 1 << 53, (1 << 63) - 1, 1 << 63,
               (1 << 64) - 1, (1 << 64) - (1 << 10), (1 << 64) - (1 << 11), (1 << 64) - (1 << 11) - 1,
               (1 << 64) - (1 << 11) + 1, 0x9E3779B97F4A7C15, 0x123456789ABCDEF0]
    rng = np.random.default_rng(20261202)
    special += [int(x) for x in rng.integers(0, 2**63, 48, dtype=np.uint64)]
    special += [int(x) | (1 << 63) for x in rng.integers(0, 2**63, 48, dtype=np.uint64)]
    r = np.array(special, dtype=np.uint64)
    u = np.zeros(len(r))
    L.ref_urand_replay(r.ctypes.data, len(r), u.ctypes.data)
    out["urand"] = {"raw": hexs(r), "u": [float(x).hex() for x in u]}

    # sample_weights: weight tables of the reference (GW_utils.tpp:7,25,37,56), subunit-like
    # rate vectors with zeros, and totals that differ from the exact sum.
    cases = []
    tables = [
        ([0.958, 0.038, 0.004], 1.0),
        ([0.9425, 0.0575], 1.0),
        ([0.609, 0.5 * 0.391, 0.5 * 0.391], 1.0),
        ([0.998, 0.002], 1.0),
        ([3.0, 0.0, 1.0e-3, 5.5], 8.501),
        ([0.0, 0.0, 2.0], 2.0),
        ([1.0, 2.0, 3.0, 0.0, 0.0, 4.0, 0.0, 1e-9, 0.0, 0.5, 0.25, 0.0], 10.75),
        ([1.0, 2.0, 3.0, 0.0, 0.0, 4.0, 0.0, 1e-9, 0.0, 0.5, 0.25, 0.0], 10.75 * (1 + 1e-15)),
        ([2.5, 2.5], 4.0),
    ]
    for w, total in tables:
        w = np.array(w)
        rr = np.array(list(rng.integers(0, 2**63, 40, dtype=np.uint64))
                      + [0, (1 << 64) - 1, 1 << 63, (1 << 64) - (1 << 11)], dtype=np.uint64)
        res = np.zeros(len(rr), np.int32)
        L.ref_sample_weights(w.ctypes.data, total, len(w), rr.ctypes.data, len(rr), res.ctypes.data)
        cases.append({"w": [float(x).hex() for x in w], "total": float(total).hex(), "raw": hexs(rr), "idx": res.tolist()})
    out["sample_weights"] = cases

    binom = []
    for p, N in [(0.5, 5), (0.1, 3), (0.999, 5), (1e-6, 2), (0.3, 0), (0.7, 1)]:
        rr = np.array(rng.integers(0, 2**64 - 1, max(N, 1), dtype=np.uint64), dtype=np.uint64)
        res = C.c_int(0)
        used = L.ref_sample_binomial(p, N, rr.ctypes.data, C.byref(res))
        binom.append({"p": float(p).hex(), "N": N, "raw": hexs(rr), "X": res.value, "used": used})
    out["sample_binomial"] = binom

    # common.hpp point functions over voltage / concentration grids
    Vs = sorted(set([-100.0, -91.382, -80.0, -60.0, -47.13, -47.0, -40.0000001, -40.0, -39.9, -20.0, -10.0,
                     -1e-3, 1e-3, 5.0, 10.0, 20.0, 40.0, 60.0] + list(np.linspace(-95, 55, 31))))
    F_RT = 96.5 / (8.314 * 310.0)
    pts = []
    for V in Vs:
        vf = V * F_RT
        e = float(np.exp(-vf))
        row = {"V": float(V).hex()}
        for name in ("alpha_m", "beta_m", "alpha_h", "beta_h", "alpha_j", "beta_j"):
            row[name] = getattr(L, "ref_" + name)(V).hex()
        row["e"] = e.hex()
        row["fNaK3"] = L.ref_fNaK3(vf, e, 0.0).hex() if False else L.ref_fNaK3(vf, e, 0.9).hex()
        row["INaK6"] = L.ref_INaK6(vf, e, 10.3, 0.9, 10.0, 0.65).hex()
        row["INaCa9"] = L.ref_INaCa9(vf, e, 10.3, 1.6e-4, 138.0 ** 3, 2.0, 0.35, 0.2, 3.7e-4).hex()
        row["INa"] = L.ref_INa(V, 0.3, 0.5, 0.6, 12.8, 70.0).hex()
        row["Ib"] = L.ref_Ib(V, 0.00264, 70.0).hex()
        pts.append(row)
    out["common_V"] = pts
    conc = []
    for Xi in [1e-5, 1.45e-4, 0.5, 4.0, 10.0, 131.84, 138.0, 150.0]:
        conc.append({
            "Xi": float(Xi).hex(),
            "Nernst1": L.ref_Nernst(Xi, 138.0, 8.314 * 310.0 / 96.5, 1.0).hex(),
            "Nernst2": L.ref_Nernst(Xi, 2.0, 8.314 * 310.0 / 96.5, 2.0).hex(),
            "IpCa": L.ref_IpCa(Xi * 1e-4, 0.03, 0.0005).hex(),
        })
    out["common_conc"] = conc
    with open(os.path.join(HERE, "ref_vectors.json"), "w", encoding="utf-8") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "ref_vectors.json"))


if __name__ == "__main__":
    main()
