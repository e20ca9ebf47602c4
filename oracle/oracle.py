"""ctypes wrapper of the CPU oracle (oracle/gw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
REF = os.path.join(HERE, "_ref")

u64 = C.c_uint64
f64 = C.c_double
i32 = C.c_int32
P = C.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


# ---------------------------------------------------------------------------
# struct mirrors of include/gw_capi.h
# ---------------------------------------------------------------------------
def _param_fields():
    """(name, ctype, default) of GW_PARAMETER_FIELDS, parsed from the C header."""
    hdr = os.path.join(HERE, "..", "include", "gw_capi.h")
    text = open(hdr, encoding="utf-8").read()
    start = text.index("#define GW_PARAMETER_FIELDS(X)")
    end = text.index("#define GW_DECLARE_FIELD_")
    body = text[start:end]
    import re

    out = []
    for ty, name, dflt in re.findall(r"X\((double|int32_t),\s*(\w+),\s*([-+0-9.eE]+)\)", body):
        out.append((name, C.c_double if ty == "double" else C.c_int32, float(dflt) if ty == "double" else int(dflt)))
    return out


PARAM_FIELDS = _param_fields()


class GwParams(C.Structure):
    _fields_ = [(n, t) for n, t, _ in PARAM_FIELDS]

    @classmethod
    def default(cls):
        p = cls()
        for n, _, d in PARAM_FIELDS:
            setattr(p, n, d)
        return p


class GwGlobals(C.Structure):
    _fields_ = [(n, f64) for n in ("V", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "m", "h", "j", "xKs")] + [
        ("Kr", f64 * 5),
        ("Kv43", f64 * 10),
        ("Kv14", f64 * 10),
    ]

    @classmethod
    def default(cls):
        """GW::GlobalState defaults (src/includes/GW_utils.hpp:231-247)."""
        g = cls()
        g.V, g.Nai, g.Ki, g.Cai, g.CaNSR = -91.382, 10.0, 131.84, 1.45273e-4, 0.908882
        g.CaLTRPN, g.CaHTRPN, g.m, g.h, g.j, g.xKs = 8.98282e-3, 0.137617, 5.33837e-4, 0.996345, 0.997315, 2.04171e-4
        g.Kr[:] = [0.999503, 4.13720e-4, 7.27568e-5, 8.73984e-6, 1.36159e-6]
        g.Kv43[:] = [0.953060, 0.0253906, 2.53848e-4, 1.12796e-6, 1.87950e-9, 0.0151370, 0.00517622, 8.96600e-4, 8.17569e-5, 2.24032e-6]
        g.Kv14[:] = [0.722328, 0.101971, 0.00539932, 1.27081e-4, 1.82742e-6, 0.152769, 0.00962328, 0.00439043, 0.00195348, 0.00143629]
        return g

    def as_array(self):
        return np.frombuffer(bytes(self), dtype=np.float64).copy()


class GwRecord(C.Structure):
    _fields_ = [("num_records", C.c_int64)] + [
        (n, P)
        for n in (
            "t", "V", "m", "h", "j", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "xKs", "XKr", "XKv14",
            "XKv43", "CaJSR", "CaSS", "LCC", "LCC_inactivation", "ClCh", "RyR",
        )
    ]


# ---------------------------------------------------------------------------
# library loading
# ---------------------------------------------------------------------------
def build(quiet: bool = True) -> None:
    subprocess.run(["make", "-C", HERE, "all"], check=True, capture_output=quiet)


_LIBS: dict[str, C.CDLL] = {}


def lib(kind: str = "strict") -> C.CDLL:
    """The oracle library of one build: "strict" (IEEE, the checker), "par"
    (the same with OpenMP over CRUs) or "omp" (the reference's release flags,
    the timed CPU baseline). Loading "omp" sets flush-to-zero and
    denormals-are-zero for the whole process (its -Ofast start-up code, as
    with the reference's own -Ofast module), so checks that need IEEE
    subnormals must not share a process with it."""
    name = {"strict": "liboracle.so", "par": "liboracle_par.so", "omp": "liboracle_omp.so"}[kind]
    if kind in _LIBS:
        return _LIBS[kind]
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
    _LIBS[kind] = load(path)
    return _LIBS[kind]


def load(path: str) -> C.CDLL:
    """Load an oracle build from `path` and declare its signatures (lib() for
    the three builds of oracle/Makefile; tests also load mutated builds)."""
    L = C.CDLL(path)
    L.gwo_splitmix64_next.restype = u64
    L.gwo_splitmix64_next.argtypes = [C.POINTER(u64)]
    L.gwo_rng_seed.argtypes = [C.c_int, P, u64]
    L.gwo_rng_next.restype = u64
    L.gwo_rng_next.argtypes = [C.c_int, P]
    L.gwo_u01.restype = f64
    L.gwo_u01.argtypes = [u64]
    L.gwo_stream_key.restype = u64
    L.gwo_stream_key.argtypes = [u64, C.c_int64]
    for fn in ("gwo_exp", "gwo_log"):
        getattr(L, fn).restype = f64
        getattr(L, fn).argtypes = [f64]
    L.gwo_pow.restype = f64
    L.gwo_pow.argtypes = [f64, f64]
    L.gwo_fx_add.argtypes = [P, f64, C.POINTER(C.c_int)]
    L.gwo_fx_to_double.restype = f64
    L.gwo_fx_to_double.argtypes = [P]
    L.gwo_Nernst.restype = f64
    L.gwo_Nernst.argtypes = [f64, f64, f64, f64, C.POINTER(C.c_int)]
    sigs = {
        "gwo_INa": 6, "gwo_Ib": 3, "gwo_IpCa": 3, "gwo_fNaK": 3, "gwo_INaK": 6, "gwo_INaCa": 9,
        "gwo_alpha_m": 1, "gwo_beta_m": 1, "gwo_alpha_h": 1, "gwo_beta_h": 1, "gwo_alpha_j": 1, "gwo_beta_j": 1,
        "gwo_XKsinf": 1, "gwo_tauXKs_inv": 1, "gwo_alphaLCC": 1, "gwo_betaLCC": 1, "gwo_yinfLCC": 1, "gwo_tauLCC": 1,
    }
    for fn, n in sigs.items():
        getattr(L, fn).restype = f64
        getattr(L, fn).argtypes = [f64] * n
    L.gwo_sample_weights.restype = C.c_int
    L.gwo_sample_weights.argtypes = [P, f64, C.c_int, f64]
    L.gwo_sample_binomial_replay.restype = C.c_int
    L.gwo_sample_binomial_replay.argtypes = [f64, C.c_int, P, C.POINTER(C.c_int)]
    L.gwo_subunit_rates.restype = f64
    L.gwo_subunit_rates.argtypes = [C.POINTER(GwParams), f64, C.c_int, C.c_int, P, C.c_int, f64, P, P, P, P]
    L.gwo_Q_matrices.argtypes = [C.POINTER(GwParams), f64, P, P, P]
    L.gwo_global_rhs.restype = C.c_int
    L.gwo_global_rhs.argtypes = [C.POINTER(GwParams), C.c_int, C.POINTER(GwGlobals), f64, f64, f64, f64, f64, P, P]
    L.gwo_model_create.restype = P
    L.gwo_model_create.argtypes = [C.POINTER(GwParams), C.c_int, C.c_int, u64]
    L.gwo_model_destroy.argtypes = [P]
    L.gwo_model_set_state.restype = C.c_int
    L.gwo_model_set_state.argtypes = [P, C.POINTER(GwGlobals), P, P, P, P, P, P, P]
    L.gwo_model_sample_rest_state.argtypes = [P]
    L.gwo_model_set_clamp.argtypes = [P, C.c_int, f64]
    L.gwo_model_set_frozen.argtypes = [P, C.c_int]
    L.gwo_model_set_event_log.argtypes = [P, C.c_int]
    L.gwo_model_euler_step.restype = C.c_int
    L.gwo_model_euler_step.argtypes = [P, f64, f64]
    L.gwo_model_run.restype = C.c_int
    L.gwo_model_run.argtypes = [P, f64, C.c_int64, P, i32, C.POINTER(GwRecord)]
    L.gwo_model_get_globals.argtypes = [P, C.POINTER(GwGlobals)]
    L.gwo_model_get_state.argtypes = [P, P, P, P, P, P, P, P]
    L.gwo_model_get_digest.argtypes = [P, P, P]
    L.gwo_model_get_event_log.argtypes = [P, P]
    L.gwo_model_get_sums.argtypes = [P, P]
    L.gwo_model_iterations.restype = u64
    L.gwo_model_iterations.argtypes = [P]
    L.gwo_model_threads.restype = C.c_int
    return L


def ref_lib() -> C.CDLL | None:
    """The reference-built harness (oracle/_ref/libgwref.so), if it was built."""
    path = os.path.join(REF, "libgwref.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_splitmix64.argtypes = [u64, C.c_int, P]
    L.ref_raw.argtypes = [C.c_int, u64, C.c_int, P]
    L.ref_uniform.argtypes = [C.c_int, u64, C.c_int, P]
    L.ref_urand_replay.argtypes = [P, C.c_int, P]
    L.ref_sample_weights.argtypes = [P, f64, C.c_int, P, C.c_int, P]
    L.ref_sample_binomial.restype = C.c_int
    L.ref_sample_binomial.argtypes = [f64, C.c_int, P, P]
    for fn, n in {
        "ref_Nernst": 4, "ref_INa": 6, "ref_Ib": 3, "ref_IpCa": 3, "ref_fNaK2": 2, "ref_fNaK3": 3, "ref_INaK5": 5,
        "ref_INaK6": 6, "ref_INaCa8": 8, "ref_INaCa9": 9, "ref_alpha_m": 1, "ref_beta_m": 1, "ref_alpha_h": 1,
        "ref_beta_h": 1, "ref_alpha_j": 1, "ref_beta_j": 1,
    }.items():
        getattr(L, fn).restype = f64
        getattr(L, fn).argtypes = [f64] * n
    return L


# ---------------------------------------------------------------------------
# model wrapper
# ---------------------------------------------------------------------------
class OracleModel:
    """GW_model<double, PRNG> restated on the CPU (src/includes/GW.hpp:45-124)."""

    def __init__(self, n_cru: int, params: GwParams | None = None, seed: int = 42, prng: int = 0, kind: str = "strict"):
        self.L = lib(kind)
        self.n = int(n_cru)
        self.params = params if params is not None else GwParams.default()
        self.h = self.L.gwo_model_create(C.byref(self.params), self.n, prng, seed)
        if not self.h:
            raise ValueError("gwo_model_create failed")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.gwo_model_destroy(h)
            self.h = None

    def sample_rest_state(self):
        self.L.gwo_model_sample_rest_state(self.h)

    def set_state(self, globals_=None, CaSS=None, CaJSR=None, LCC=None, LCC_inactivation=None, RyR=None, ClCh=None, JLCC=None):
        def arr(a, dt):
            return None if a is None else np.ascontiguousarray(a, dtype=dt)

        keep = [arr(CaSS, np.float64), arr(CaJSR, np.float64), arr(LCC, np.int32), arr(LCC_inactivation, np.int32),
                arr(RyR, np.int32), arr(ClCh, np.int32), arr(JLCC, np.float64)]
        g = C.byref(globals_) if globals_ is not None else None
        self.L.gwo_model_set_state(self.h, g, *[_ptr(a) for a in keep])

    def set_clamp(self, on: bool, v: float = 0.0):
        self.L.gwo_model_set_clamp(self.h, 1 if on else 0, v)

    def set_frozen(self, on: bool):
        self.L.gwo_model_set_frozen(self.h, 1 if on else 0)

    def set_event_log(self, cap: int):
        self.cap = cap
        self.L.gwo_model_set_event_log(self.h, cap)

    def step(self, dt: float = 1e-3, istim: float = 0.0) -> int:
        return self.L.gwo_model_euler_step(self.h, dt, istim)

    def run(self, dt: float, num_steps: int, istim=None, record_every: int = 0, fields=None):
        """run<PRNG> (src/pyGW.hpp:102-122). Returns dict of records (or {} if record_every == 0);
        `fields` limits the records to the named fields (all by default)."""
        ist = None if istim is None else np.ascontiguousarray(istim, dtype=np.float64)
        rec = GwRecord()
        out = {}
        if record_every > 0 and fields is not None:
            R = -(-num_steps // record_every)
            out = {k: np.zeros(R) for k in fields}
            rec.num_records = R
            for k, v in out.items():
                setattr(rec, k, v.ctypes.data)
        elif record_every > 0:
            R = -(-num_steps // record_every)
            n = self.n
            out = {
                "t": np.zeros(R), "V": np.zeros(R), "m": np.zeros(R), "h": np.zeros(R), "j": np.zeros(R),
                "Nai": np.zeros(R), "Ki": np.zeros(R), "Cai": np.zeros(R), "CaNSR": np.zeros(R),
                "CaLTRPN": np.zeros(R), "CaHTRPN": np.zeros(R), "xKs": np.zeros(R), "XKr": np.zeros((R, 5)),
                "XKv14": np.zeros((R, 10)), "XKv43": np.zeros((R, 10)), "CaJSR": np.zeros((R, n)),
                "CaSS": np.zeros((R, n, 4)), "LCC": np.zeros((R, n, 4), np.int32),
                "LCC_inactivation": np.zeros((R, n, 4), np.int32), "ClCh": np.zeros((R, n, 4), np.int32),
                "RyR": np.zeros((R, n, 4, 6), np.int32),
            }
            rec.num_records = R
            for k, v in out.items():
                setattr(rec, k, v.ctypes.data)
        rc = self.L.gwo_model_run(self.h, dt, num_steps, _ptr(ist), record_every, C.byref(rec))
        if rc != 0:
            raise RuntimeError(f"oracle run failed with status {rc}")
        return out

    def globals(self) -> GwGlobals:
        g = GwGlobals()
        self.L.gwo_model_get_globals(self.h, C.byref(g))
        return g

    def state(self) -> dict:
        n = self.n
        s = {
            "CaSS": np.zeros((n, 4)), "CaJSR": np.zeros(n), "LCC": np.zeros((n, 4), np.int32),
            "LCC_inactivation": np.zeros((n, 4), np.int32), "RyR": np.zeros((n, 4, 6), np.int32),
            "ClCh": np.zeros((n, 4), np.int32), "JLCC": np.zeros((n, 4)),
        }
        self.L.gwo_model_get_state(self.h, *[_ptr(s[k]) for k in ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "RyR", "ClCh", "JLCC")])
        return s

    def digest(self):
        c = np.zeros(self.n, np.uint64)
        h = np.zeros(self.n, np.uint64)
        self.L.gwo_model_get_digest(self.h, _ptr(c), _ptr(h))
        return c, h

    def event_log(self):
        out = np.zeros((self.n, self.cap), np.uint32)
        self.L.gwo_model_get_event_log(self.h, _ptr(out))
        return out

    def sums(self):
        s = np.zeros(5)
        self.L.gwo_model_get_sums(self.h, _ptr(s))
        return s

    def iterations(self) -> int:
        return int(self.L.gwo_model_iterations(self.h))
