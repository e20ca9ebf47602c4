"""ctypes binding of the C-ABI in include/gw_capi.h (libgwhip.so).

This is the thin Python view of the HIP library that tests and bench.py use;
the drop-in Python surface of the reference is the GreensteinWinslow module
plus model.py / parameters.py. Loading fails loudly if the HIP library has not
been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "_lib", "libgwhip.so")
HEADER = os.path.join(ROOT, "include", "gw_capi.h")

f64, i32, i64, u64, P = C.c_double, C.c_int32, C.c_int64, C.c_uint64, C.c_void_p

GW_OK = 0
STATUS = {0: "GW_OK", -1: "GW_E_INVALID_ARG", -2: "GW_E_NONPOSITIVE_CONC", -3: "GW_E_NAN", -4: "GW_E_HIP",
          -5: "GW_E_RANGE", -6: "GW_E_NO_DEVICE", -7: "GW_E_TIMEOUT"}
PREC_FP64, PREC_FP32_LOCAL = 0, 1
PROTO_TABLE, PROTO_PACE, PROTO_CLAMP = 0, 1, 2
PRNG_NAMES = ["xoshiro256**", "xoshiro256++", "xoshiro256+", "xoroshiro128**", "xoroshiro128++", "xoroshiro128+"]


def _param_fields():
    text = open(HEADER, encoding="utf-8").read()
    body = text[text.index("#define GW_PARAMETER_FIELDS(X)"):text.index("#define GW_DECLARE_FIELD_")]
    out = []
    for ty, name, dflt in re.findall(r"X\((double|int32_t),\s*(\w+),\s*([-+0-9.eE]+)\)", body):
        out.append((name, f64 if ty == "double" else i32, float(dflt) if ty == "double" else int(dflt)))
    return out


PARAM_FIELDS = _param_fields()
PARAM_NAMES = [n for n, _, _ in PARAM_FIELDS if not n.startswith("_")]


class Params(C.Structure):
    """gw_params_t: GW::Parameters (src/includes/GW_utils.hpp:38-160)."""

    _fields_ = [(n, t) for n, t, _ in PARAM_FIELDS]

    @classmethod
    def default(cls):
        p = cls()
        for n, _, d in PARAM_FIELDS:
            setattr(p, n, d)
        return p


GLOBAL_SCALARS = ("V", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "m", "h", "j", "xKs")


class Globals(C.Structure):
    """gw_globals_t: GW::GlobalState (src/includes/GW_utils.hpp:229-252)."""

    _fields_ = [(n, f64) for n in GLOBAL_SCALARS] + [("Kr", f64 * 5), ("Kv43", f64 * 10), ("Kv14", f64 * 10)]

    def as_array(self):
        return np.frombuffer(bytes(self), dtype=np.float64).copy()


class CruState(C.Structure):
    _fields_ = [(n, P) for n in ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "RyR", "ClCh", "JLCC")]


class Record(C.Structure):
    _fields_ = [("num_records", i64)] + [(n, P) for n in (
        "t", "V", "m", "h", "j", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "xKs", "XKr", "XKv14", "XKv43",
        "CaJSR", "CaSS", "LCC", "LCC_inactivation", "ClCh", "RyR")]


class CruStateReduced(C.Structure):
    _fields_ = [(n, P) for n in ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "ClCh", "RyR")]


class RecordReduced(C.Structure):
    _fields_ = [("num_records", i64)] + [(n, P) for n in (
        "t", "V", "m", "h", "j", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "xKs", "XKr", "XKv14", "XKv43",
        "CaJSR", "CaSS", "LCC", "LCC_inactivation", "ClCh", "RyR")]


class Protocol(C.Structure):
    _fields_ = [("kind", i32), ("_pad", i32), ("period", f64), ("duration", f64), ("amplitude", f64), ("start", f64),
                ("hold", f64), ("step", f64), ("t_on", f64), ("t_off", f64)]


class CellDesc(C.Structure):
    _fields_ = [("n_cru", i32), ("_pad", i32), ("seed", u64), ("params", Params), ("protocol", Protocol)]


OBS_FIELDS = ("peak_V", "min_V", "peak_Cai", "peak_ICaL_neg", "sum_open_RyR_frac", "int_Jrel", "n_steps", "v_prev",
              "apd_v_rest", "apd_v_peak", "apd_t_up", "apd90_last", "apd90_prev", "alt_sum", "alt_count", "beats_done")


class CellObs(C.Structure):
    _fields_ = [(n, f64) for n in OBS_FIELDS]


class GwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    E = [C.c_char_p, C.c_size_t]
    sig = {
        "gw_params_default": (None, [C.POINTER(Params)]),
        "gw_globals_default": (None, [C.POINTER(Globals)]),
        "gw_prng_from_name": (C.c_int, [C.c_char_p]),
        "gw_version": (C.c_int, []),
        "gw_ctx_create": (C.c_int, [C.POINTER(P), C.c_int] + E),
        "gw_ctx_destroy": (C.c_int, [P]),
        "gw_run": (C.c_int, [P, C.POINTER(Params), i32, f64, i64, P, i32, C.POINTER(Globals), C.POINTER(CruState),
                             u64, i32, i32, C.POINTER(Record)] + E),
        "gw_run_noalloc": (C.c_int, [P, C.POINTER(Params), i32, f64, i64, P, i32, C.POINTER(Globals),
                                     C.POINTER(CruState), u64, i32, i32, C.POINTER(Record)] + E),
        "gw_batch_create": (C.c_int, [P, i32, P, i32, i32, f64, C.POINTER(Globals), C.POINTER(P)] + E),
        "gw_batch_destroy": (C.c_int, [P]),
        "gw_batch_set_cru_state": (C.c_int, [P, i32, C.POINTER(CruState)] + E),
        "gw_batch_get_cru_state": (C.c_int, [P, i32, C.POINTER(CruState)] + E),
        "gw_batch_get_cru_state_reduced": (C.c_int, [P, i32, C.POINTER(CruStateReduced)] + E),
        "gw_run_reduced": (C.c_int, [P, C.POINTER(Params), i32, f64, i64, P, i32, C.POINTER(Globals),
                                     C.POINTER(CruState), u64, i32, i32, C.POINTER(RecordReduced)] + E),
        "gw_batch_set_globals": (C.c_int, [P, i32, C.POINTER(Globals)] + E),
        "gw_batch_get_globals": (C.c_int, [P, P] + E),
        "gw_batch_set_stimulus_table": (C.c_int, [P, i32, P, i64] + E),
        "gw_batch_step": (C.c_int, [P, i64] + E),
        "gw_batch_get_time": (C.c_int, [P, C.POINTER(f64), C.POINTER(i64)]),
        "gw_batch_get_obs": (C.c_int, [P, P] + E),
        "gw_batch_reset_obs": (C.c_int, [P] + E),
        "gw_batch_enable_trace": (C.c_int, [P, i32, i64] + E),
        "gw_batch_get_trace": (C.c_int, [P, P, i64, C.POINTER(i64)] + E),
        "gw_batch_get_event_digest": (C.c_int, [P, i32, P, P] + E),
        "gw_batch_set_mode": (C.c_int, [P, i32, i32]),
        "gw_batch_get_iterations": (C.c_int, [P, C.POINTER(u64)]),
        "gw_batch_set_launch": (C.c_int, [P, i32, i32, i32] + E),
        "gw_launch_config": (C.c_int, [i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "gw_batch_time_steps": (C.c_int, [P, i64, C.POINTER(f64), C.POINTER(f64), C.POINTER(i64)] + E),
        "gw_batch_get_launch": (C.c_int, [P, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "gw_batch_set_graph": (C.c_int, [P, i32] + E),
        "gw_batch_set_persistent": (C.c_int, [P, i32] + E),
        "gw_batch_get_persistent": (C.c_int, [P, C.POINTER(i32)]),
        "gw_batch_set_steps_per_launch": (C.c_int, [P, i32] + E),
        "gw_batch_get_steps_per_launch": (C.c_int, [P, C.POINTER(i32)]),
        "gw_batch_set_spin_limit": (C.c_int, [P, i64] + E),
        "gw_batch_get_currents": (C.c_int, [P, P, P] + E),
        "gw_batch_autotune": (C.c_int, [P, i64, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                        C.POINTER(f64)] + E),
        "gw_run_batch": (C.c_int, [P, i32, P, i32, f64, i64, P, P] + E),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def exported_symbols():
    """Names of every function include/gw_capi.h declares."""
    text = open(HEADER, encoding="utf-8").read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(gw_\w+)\s*\(", text, flags=re.M)))


def _call(fn, *args):
    buf = C.create_string_buffer(1024)
    rc = fn(*args, buf, len(buf))
    if rc != GW_OK:
        raise GwError(rc, buf.value.decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def default_globals() -> Globals:
    g = Globals()
    lib().gw_globals_default(C.byref(g))
    return g


class Context:
    """A device context: one HIP stream on one gfx950 GPU (gw_ctx_create)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        self.h = P()
        self._batches = weakref.WeakSet()
        _call(self.L.gw_ctx_create, C.byref(self.h), device)

    def close(self):
        # batches live on this context's stream: release them first
        for b in list(self._batches):
            b.close()
        if self.h:
            self.L.gw_ctx_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def launch_configs() -> list[tuple[int, int, int]]:
    """The launch configurations (lanes per CRU, threads per block, CRUs per
    lane) in gw_launch_config order, the order of the autotuner's timings."""
    L, out, k = lib(), [], 0
    lanes, blk, cpl = i32(), i32(), i32()
    while L.gw_launch_config(k, C.byref(lanes), C.byref(blk), C.byref(cpl)) == 0:
        out.append((lanes.value, blk.value, cpl.value))
        k += 1
    return out


def rest_state(n_cru: int, seed: int, prng: int = 0) -> dict:
    """gw_rest_state: the seeded rest state of one cell's CRUs, sampled on the
    host (initialise_LCC / _LCC_i / _RyR / _ClCh, GW_utils.tpp:5-64); a batch
    samples the same state on the device."""
    n = int(n_cru)
    out = {"CaSS": np.zeros((n, 4)), "CaJSR": np.zeros(n), "LCC": np.zeros((n, 4), np.int32),
           "LCC_inactivation": np.zeros((n, 4), np.int32), "RyR": np.zeros((n, 4, 6), np.int32),
           "ClCh": np.zeros((n, 4), np.int32), "JLCC": np.zeros((n, 4))}
    s = CruState(*[_ptr(out[k]) for k in ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "RyR", "ClCh", "JLCC")])
    L = lib()
    L.gw_rest_state.argtypes = [C.c_int32, C.c_uint64, C.c_int32, C.POINTER(CruState)]
    rc = L.gw_rest_state(n, int(seed) & (2**64 - 1), int(prng), C.byref(s))
    if rc != GW_OK:
        raise GwError(rc, f"gw_rest_state: bad arguments (n_cru {n}, prng {prng})")
    return out


def launch_name(lanes: int, block: int, crus_per_lane: int = 1) -> str:
    """'4x256' (lanes per CRU x threads), '1x64x4' with CRUs per lane above 1."""
    return f"{lanes}x{block}" + (f"x{crus_per_lane}" if crus_per_lane > 1 else "")


def make_cell(n_cru: int, seed: int = 42, params: Params | None = None, protocol: dict | None = None) -> CellDesc:
    d = CellDesc()
    d.n_cru = int(n_cru)
    d.seed = int(seed) & (2**64 - 1)
    d.params = params if params is not None else Params.default()
    pr = protocol or {"kind": PROTO_TABLE}
    for k, v in pr.items():
        setattr(d.protocol, k, v)
    return d


class Batch:
    """Many independent cells advanced together (gw_batch_* session API)."""

    def __init__(self, ctx: Context, cells, precision: int = PREC_FP64, prng: int = 0, dt: float = 1e-3,
                 globals_: Globals | None = None):
        self.ctx = ctx
        self.L = ctx.L
        self.cells = list(cells)
        self.n = len(self.cells)
        arr = (CellDesc * self.n)(*self.cells)
        self.h = P()
        _call(self.L.gw_batch_create, ctx.h, self.n, C.cast(arr, P), precision, prng, dt,
              C.byref(globals_) if globals_ is not None else None, C.byref(self.h))
        ctx._batches.add(self)

    def close(self):
        if self.h:
            self.L.gw_batch_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, frozen: bool = False, digest: bool = False):
        rc = self.L.gw_batch_set_mode(self.h, int(frozen), int(digest))
        if rc != 0:
            raise GwError(rc, "set_mode")

    def set_cru_state(self, cell: int, CaSS, CaJSR, LCC, LCC_inactivation, RyR, ClCh, JLCC=None):
        a = [np.ascontiguousarray(CaSS, np.float64), np.ascontiguousarray(CaJSR, np.float64),
             np.ascontiguousarray(LCC, np.int32), np.ascontiguousarray(LCC_inactivation, np.int32),
             np.ascontiguousarray(RyR, np.int32), np.ascontiguousarray(ClCh, np.int32),
             None if JLCC is None else np.ascontiguousarray(JLCC, np.float64)]
        s = CruState(*[_ptr(x) for x in a])
        _call(self.L.gw_batch_set_cru_state, self.h, cell, C.byref(s))

    def get_cru_state(self, cell: int) -> dict:
        n = self.cells[cell].n_cru
        out = {"CaSS": np.zeros((n, 4)), "CaJSR": np.zeros(n), "LCC": np.zeros((n, 4), np.int32),
               "LCC_inactivation": np.zeros((n, 4), np.int32), "RyR": np.zeros((n, 4, 6), np.int32),
               "ClCh": np.zeros((n, 4), np.int32), "JLCC": np.zeros((n, 4))}
        s = CruState(*[_ptr(out[k]) for k in ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "RyR", "ClCh", "JLCC")])
        _call(self.L.gw_batch_get_cru_state, self.h, cell, C.byref(s))
        return out

    def get_cru_state_reduced(self, cell: int) -> dict:
        """One cell's CRU state in the reduced record encoding (gw_cru_state_reduced_t)."""
        n = self.cells[cell].n_cru
        out = {"CaSS": np.zeros((n, 4), np.float32), "CaJSR": np.zeros(n, np.float32),
               "LCC": np.zeros((n, 4), np.int8), "LCC_inactivation": np.zeros((n, 4), np.bool_),
               "ClCh": np.zeros((n, 4), np.bool_), "RyR": np.zeros((n, 4), np.int16)}
        s = CruStateReduced(*[_ptr(out[k]) for k in ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "ClCh", "RyR")])
        _call(self.L.gw_batch_get_cru_state_reduced, self.h, cell, C.byref(s))
        return out

    def set_globals(self, cell: int, g: Globals):
        _call(self.L.gw_batch_set_globals, self.h, cell, C.byref(g))

    def get_globals(self):
        arr = (Globals * self.n)()
        _call(self.L.gw_batch_get_globals, self.h, C.cast(arr, P))
        return list(arr)

    def set_stimulus_table(self, cell: int, istim):
        a = np.ascontiguousarray(istim, np.float64)
        _call(self.L.gw_batch_set_stimulus_table, self.h, cell, _ptr(a), len(a))

    def step(self, nsteps: int):
        _call(self.L.gw_batch_step, self.h, int(nsteps))

    def time(self):
        t, s = f64(), i64()
        self.L.gw_batch_get_time(self.h, C.byref(t), C.byref(s))
        return t.value, s.value

    def obs(self):
        arr = (CellObs * self.n)()
        _call(self.L.gw_batch_get_obs, self.h, C.cast(arr, P))
        return [{k: getattr(o, k) for k in OBS_FIELDS} for o in arr]

    def reset_obs(self):
        _call(self.L.gw_batch_reset_obs, self.h)

    def enable_trace(self, every: int, capacity: int):
        self.trace_cap = capacity
        _call(self.L.gw_batch_enable_trace, self.h, every, capacity)

    def get_trace(self) -> np.ndarray:
        out = np.zeros((self.trace_cap, self.n, 4))
        rows = i64()
        _call(self.L.gw_batch_get_trace, self.h, _ptr(out), self.trace_cap, C.byref(rows))
        return out[: rows.value]

    def digest(self, cell: int):
        n = self.cells[cell].n_cru
        c = np.zeros(n, np.uint64)
        h = np.zeros(n, np.uint64)
        _call(self.L.gw_batch_get_event_digest, self.h, cell, _ptr(c), _ptr(h))
        return c, h

    def iterations(self) -> int:
        v = u64()
        self.L.gw_batch_get_iterations(self.h, C.byref(v))
        return v.value

    def set_launch(self, lanes_per_cru: int = 0, block: int = 0, crus_per_lane: int = 0):
        """gw_batch_set_launch: lanes per CRU (4 or 1), threads per block and,
        with one lane per CRU, CRUs per lane (1, 2, 4 or 8); 0 = default."""
        _call(self.L.gw_batch_set_launch, self.h, lanes_per_cru, block, crus_per_lane)

    def set_graph(self, steps_per_graph: int):
        """Replay step launches as captured HIP graphs of this many steps (0: plain launches)."""
        _call(self.L.gw_batch_set_graph, self.h, int(steps_per_graph))

    def set_persistent(self, mode: int):
        """Persistent stepping: 0 plain launches (the default), 1 one launch per call (refused if
        the blocks cannot all be resident), 2 persistent when they can be."""
        _call(self.L.gw_batch_set_persistent, self.h, int(mode))

    def set_steps_per_launch(self, steps: int):
        """SSA steps per plain launch of the FP32-local one-lane shapes of 64 threads: 0 automatic (the
        default: two when the launch's blocks are more than the GPU holds at once), 1 or 2."""
        _call(self.L.gw_batch_set_steps_per_launch, self.h, int(steps))

    def steps_per_launch(self) -> int:
        """The SSA steps per launch the next plain steps use."""
        v = i32()
        self.L.gw_batch_get_steps_per_launch(self.h, C.byref(v))
        return v.value

    def currents(self) -> tuple[np.ndarray, np.ndarray]:
        """(ICaL, Ito2) of every cell by the post-hoc definitions of src/model.py:414-481, on the device."""
        i, t = np.zeros(self.n), np.zeros(self.n)
        _call(self.L.gw_batch_get_currents, self.h, _ptr(i), _ptr(t))
        return i, t

    def set_spin_limit(self, polls: int):
        """Polls before a wait of the persistent kernel gives up (0: the default)."""
        _call(self.L.gw_batch_set_spin_limit, self.h, int(polls))

    def persistent(self) -> bool:
        """Whether the next steps run in the persistent kernel."""
        v = i32()
        self.L.gw_batch_get_persistent(self.h, C.byref(v))
        return bool(v.value)

    def launch(self) -> tuple[int, int, int]:
        """(lanes per CRU, threads per block, CRUs per lane) of the current SSA launch shape."""
        lanes, blk, cpl = i32(), i32(), i32()
        self.L.gw_batch_get_launch(self.h, C.byref(lanes), C.byref(blk), C.byref(cpl))
        return lanes.value, blk.value, cpl.value

    def autotune(self, trial_steps: int = 200) -> tuple[tuple[int, int, int], dict]:
        """gw_batch_autotune: pick the fastest SSA launch shape, with or without
        graph replay, on this batch (its state is left unchanged). Returns the
        shape and the trial times per step keyed by launch_name(); `graph_steps`
        is the chosen graph size."""
        lanes, blk, cpl, graph = i32(), i32(), i32(), i32()
        shapes = launch_configs()
        ms = (f64 * (len(shapes) + 1))()
        _call(self.L.gw_batch_autotune, self.h, int(trial_steps), C.byref(lanes), C.byref(blk), C.byref(cpl),
              C.byref(graph), ms)
        trials = {launch_name(*sh): (ms[k] if ms[k] >= 0 else None) for k, sh in enumerate(shapes)}
        trials["graph64"] = ms[len(shapes)] if ms[len(shapes)] >= 0 else None
        self.graph_steps = graph.value
        return (lanes.value, blk.value, cpl.value), trials

    def time_steps(self, nsteps: int, per_kernel: bool = False):
        """(total ms, SSA-kernel ms or None, SSA launches) of nsteps Euler steps."""
        ms, ssa, launches = f64(), f64(), i64()
        _call(self.L.gw_batch_time_steps, self.h, int(nsteps), C.byref(ms), C.byref(ssa) if per_kernel else None,
              C.byref(launches))
        return ms.value, (ssa.value if per_kernel else None), launches.value


def _crus_struct(crus: dict | None):
    if crus is None:
        return None, None
    keep = [np.ascontiguousarray(crus["CaSS"], np.float64), np.ascontiguousarray(crus["CaJSR"], np.float64),
            np.ascontiguousarray(crus["LCC"], np.int32), np.ascontiguousarray(crus["LCC_inactivation"], np.int32),
            np.ascontiguousarray(crus["RyR"], np.int32), np.ascontiguousarray(crus["ClCh"], np.int32)]
    return CruState(*[_ptr(x) for x in keep], None), keep


def run(ctx: Context, params: Params, n_cru: int, dt: float, num_steps: int, istim, record_every: int,
        globals_: Globals | None, crus: dict | None, seed: int, prng: int = 0, precision: int = PREC_FP64,
        noalloc: bool = False) -> dict:
    """gw_run: the run<PRNG> loop (src/pyGW.hpp:102-122) with ceil(num_steps / record_every) records.
    noalloc=True calls gw_run_noalloc instead: the same run with no device allocation."""
    R = -(-int(num_steps) // int(record_every))
    n = int(n_cru)
    out = {"t": np.zeros(R), "V": np.zeros(R), "m": np.zeros(R), "h": np.zeros(R), "j": np.zeros(R),
           "Nai": np.zeros(R), "Ki": np.zeros(R), "Cai": np.zeros(R), "CaNSR": np.zeros(R), "CaLTRPN": np.zeros(R),
           "CaHTRPN": np.zeros(R), "xKs": np.zeros(R), "XKr": np.zeros((R, 5)), "XKv14": np.zeros((R, 10)),
           "XKv43": np.zeros((R, 10)), "CaJSR": np.zeros((R, n)), "CaSS": np.zeros((R, n, 4)),
           "LCC": np.zeros((R, n, 4), np.int32), "LCC_inactivation": np.zeros((R, n, 4), np.int32),
           "ClCh": np.zeros((R, n, 4), np.int32), "RyR": np.zeros((R, n, 4, 6), np.int32)}
    rec = Record()
    rec.num_records = R
    for k, v in out.items():
        setattr(rec, k, v.ctypes.data)
    ist = None if istim is None else np.ascontiguousarray(istim, np.float64)
    cs, keep = _crus_struct(crus)
    fn = lib().gw_run_noalloc if noalloc else lib().gw_run
    _call(fn, ctx.h, C.byref(params), n, dt, int(num_steps), _ptr(ist), int(record_every),
          C.byref(globals_) if globals_ is not None else None, C.byref(cs) if cs is not None else None,
          int(seed) & (2**64 - 1), prng, precision, C.byref(rec))
    del keep
    return out


def run_reduced(ctx: Context, params: Params, n_cru: int, dt: float, num_steps: int, istim, record_every: int,
                globals_: Globals | None, crus: dict | None, seed: int, prng: int = 0,
                precision: int = PREC_FP64) -> dict:
    """gw_run_reduced: run() with its records in the reduced encoding of
    sim_dict_reduced_memory (src/serialization.py:11-39), formed on the device."""
    R = -(-int(num_steps) // int(record_every))
    n = int(n_cru)
    out = {k: np.zeros(R, np.float32) for k in ("t", "V", "m", "h", "j", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN",
                                                 "CaHTRPN", "xKs")}
    out.update({"XKr": np.zeros((R, 4), np.float32), "XKv14": np.zeros((R, 9), np.float32),
                "XKv43": np.zeros((R, 9), np.float32), "CaJSR": np.zeros((R, n), np.float32),
                "CaSS": np.zeros((R, n, 4), np.float32), "LCC": np.zeros((R, n, 4), np.int8),
                "LCC_inactivation": np.zeros((R, n, 4), np.bool_), "ClCh": np.zeros((R, n, 4), np.bool_),
                "RyR": np.zeros((R, n, 4), np.int16)})
    rec = RecordReduced()
    rec.num_records = R
    for k, v in out.items():
        setattr(rec, k, v.ctypes.data)
    ist = None if istim is None else np.ascontiguousarray(istim, np.float64)
    cs, keep = _crus_struct(crus)
    _call(lib().gw_run_reduced, ctx.h, C.byref(params), n, dt, int(num_steps), _ptr(ist), int(record_every),
          C.byref(globals_) if globals_ is not None else None, C.byref(cs) if cs is not None else None,
          int(seed) & (2**64 - 1), prng, precision, C.byref(rec))
    del keep
    return out
