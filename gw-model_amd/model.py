"""GWModel / GWSolution: the reference's Python API (src/model.py:552-1035) over
the MI355X integrator.

GWModel.simulate calls GreensteinWinslow.run, which runs the coupled
integrator on the GPU through the C-ABI. GWSolution exposes the recorded
snapshots and the post-hoc ICaL / Ito2 of the reference (src/model.py:414-481),
the definitions the peak-I_CaL metric uses.
"""
from __future__ import annotations

import pickle as pkl
from typing import Callable, List

import numpy as np
import numpy.typing as npt

import build.GreensteinWinslow as gw_cxx  # type: ignore  # noqa: E402

from .parameters import GWParameters
from .utils import assert_gt_numpy, assert_lt_numpy, assert_positive

F = 96.5
R = 8.314
F_R = F / R

IMPLEMENTED_PRNGS = ["mt19937", "mt19937_64", "xoshiro256+", "xoshiro256++", "xoshiro256**", "xoroshiro128+",
                     "xoroshiro128++", "xoroshiro128**"]


def _catch_invalid_state(x, lb: int, ub: int, name: str) -> None:
    try:
        assert_gt_numpy(x, lb, name, strict=False)
        assert_lt_numpy(x, ub, name, strict=False)
    except ValueError as e:
        raise ValueError(f"{name} must consist of an integers in the range (lb, ..., ub).") from e


def ICaL(V: npt.NDArray, LCC: npt.NDArray, LCC_inactivation: npt.NDArray, CaSS: npt.NDArray,
         parameters: GWParameters) -> npt.NDArray:
    """Post-hoc L-type Ca current [pA/pF] from recorded V and CaSS (src/model.py:414-451).

    Open channels are LCC in {6, 12} with LCC_inactivation == 1; the GHK flux
    uses the instantaneous CaSS (the integrator itself freezes each channel's
    flux at opening).
    """
    assert_gt_numpy(CaSS, 0, "CaSS", strict=True)
    _catch_invalid_state(LCC, 1, 12, "LCC")
    _catch_invalid_state(LCC_inactivation, 0, 1, "LCC_inactivation")
    open_lcc = ((LCC == 6) | (LCC == 12)) & (LCC_inactivation == 1)
    VF_RT = np.asarray(V)[..., np.newaxis, np.newaxis] * F_R / parameters.T
    e2 = np.exp(2 * VF_RT)
    mult = (parameters.NCaRU / parameters.NCaRU_sim) * (parameters.PCaL / parameters.CSA) * (4 * F * VF_RT) / (e2 - 1)
    per_channel = mult * open_lcc * (CaSS * e2 - 0.341 * parameters.Cao)
    return per_channel.sum(axis=(1, 2)) * 1e9


def Ito2(V: npt.NDArray, ClCh: npt.NDArray, parameters: GWParameters) -> npt.NDArray:
    """Post-hoc Ca-activated Cl current [pA/pF] (src/model.py:454-481)."""
    _catch_invalid_state(ClCh, 0, 1, "ClCh")
    VF_RT = np.asarray(V)[..., np.newaxis, np.newaxis] * F_R / parameters.T
    em = np.exp(-VF_RT)
    mult = (parameters.NCaRU / parameters.NCaRU_sim) * (parameters.Pto2 / parameters.CSA) * (F * VF_RT) / (em - 1)
    per_channel = mult * ClCh * (parameters.Clcyto * em - parameters.Clo)
    return per_channel.sum(axis=(1, 2)) * 1e9


_VAR_NAMES = ("t", "V", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "m", "h", "j", "xKs", "XKr", "XKv14",
              "XKv43", "CaSS", "CaJSR", "LCC", "LCC_inactivation", "RyR", "ClCh")


class GWSolution:
    """Snapshots of one simulation (src/model.py:552-899)."""

    def __init__(self, gw_cxx_output, gw_parameters: GWParameters):
        self.__vars = gw_cxx_output
        self.__params = gw_parameters
        self.__ICaL = None
        self.__Ito2 = None

    @classmethod
    def from_dict(cls, state_dict: dict):
        v = state_dict["vars"]
        p = state_dict["params"]
        t = v["t"]
        cxx = gw_cxx.GWVariables(p["NCaRU_sim"], t.shape[0], float(t[-1]) if len(t) else 0.0)
        for name in _VAR_NAMES:
            setattr(cxx, name, v[name])
        return cls(cxx, GWParameters.from_dict(p))

    def to_dict(self) -> dict:
        return {"params": self.__params.to_dict(), "vars": {name: getattr(self, name) for name in _VAR_NAMES}}

    def save(self, fname: str) -> None:
        with open(fname, "wb") as f:
            pkl.dump(self.to_dict(), f)

    @classmethod
    def load(cls, fname: str):
        with open(fname, "rb") as f:
            return cls.from_dict(pkl.load(f))

    @property
    def parameters(self) -> GWParameters:
        return self.__params

    def __getattr__(self, name):
        if name in _VAR_NAMES:
            return getattr(self.__vars, name)
        raise AttributeError(name)

    @property
    def ICaL(self) -> npt.NDArray:
        if self.__ICaL is None:
            self.__ICaL = ICaL(self.V, self.LCC, self.LCC_inactivation, self.CaSS, self.parameters)
        return self.__ICaL

    @property
    def Ito2(self) -> npt.NDArray:
        if self.__Ito2 is None:
            self.__Ito2 = Ito2(self.V, self.ClCh, self.parameters)
        return self.__Ito2

    @property
    def RyR_open_fraction(self) -> npt.NDArray:
        """Fraction of RyRs in the open states 3 and 4 at each record."""
        R = self.RyR
        return (R[..., 2] + R[..., 3]).sum(axis=(1, 2)) / (5.0 * R.shape[1] * R.shape[2])


class GWModel:
    """Simulates the Greenstein and Winslow model (src/model.py:902-1001)."""

    @classmethod
    def PRNG_options(cls) -> List[str]:
        return list(IMPLEMENTED_PRNGS)

    def __init__(self, init_state: None | dict = None, parameters: None | GWParameters = None,
                 stimulus_fn: None | Callable[[float], float] = None):
        self.parameters: GWParameters = parameters if parameters is not None else GWParameters()
        self.__stim: Callable[[float], float] = (lambda t: 0) if stimulus_fn is None else stimulus_fn
        if init_state is not None:
            self.__global_state = _unpack_globals(init_state)
            self.__cru_state = _unpack_crus(init_state)
        else:
            self.__global_state = gw_cxx.GWGlobalState()
            self.__cru_state = gw_cxx.GWCRUState(self.parameters.NCaRU_sim)

    def simulate(self, step_size: float, num_steps: int, record_every: int = 1, PRNG: str = "xoshiro256++",
                 seed: int | None = None, precision: str = "fp64") -> GWSolution:
        """Simulate num_steps Euler steps of size step_size [ms], recording every record_every steps.

        seed (additive): seeds the per-CRU streams for a reproducible run; None
        draws a seed from std::random_device like the reference.
        precision (additive): "fp64" (the reference's arithmetic, the default) or
        "fp32_local" (CRU-local Ca and propensities in FP32; globals, channel
        states and flux sums as in fp64).
        """
        assert_positive(step_size, "step_size")
        assert_positive(num_steps, "num_steps")
        assert_positive(record_every, "record_every")
        try:
            sol = gw_cxx.run(self.parameters.cxx_struct, self.parameters.NCaRU_sim, step_size, num_steps, self.__stim,
                             record_every, init_crus=self.__cru_state, init_globals=self.__global_state, PRNG=PRNG,
                             seed=seed, precision=precision)
        except ValueError as e:
            if PRNG not in IMPLEMENTED_PRNGS:
                raise ValueError(f"{PRNG} is an invalid argument for PRNG.\nAvailable options are: "
                                 + ", ".join(self.PRNG_options())) from e
            raise
        return GWSolution(sol, self.parameters)

    def simulate_reduced(self, step_size: float, num_steps: int, record_every: int = 1, PRNG: str = "xoshiro256++",
                         seed: int | None = None, precision: str = "fp64") -> dict:
        """simulate() with its records returned in the reduced encoding of
        sim_dict_reduced_memory (src/serialization.py:11-39), formed on the
        device: the same dictionary as sim_dict_reduced_memory(simulate(...))
        for the same seed, without the full snapshots leaving the GPU."""
        assert_positive(step_size, "step_size")
        assert_positive(num_steps, "num_steps")
        assert_positive(record_every, "record_every")
        if PRNG not in IMPLEMENTED_PRNGS:
            raise ValueError(f"{PRNG} is an invalid argument for PRNG.\nAvailable options are: "
                             + ", ".join(self.PRNG_options()))
        return gw_cxx.run_reduced(self.parameters.cxx_struct, self.parameters.NCaRU_sim, step_size, num_steps,
                                  self.__stim, record_every, init_crus=self.__cru_state,
                                  init_globals=self.__global_state, PRNG=PRNG, seed=seed, precision=precision)

    def stimulus_fn(self, t: float) -> float:
        return self.__stim(t)

    def set_stimulus_fn(self, stimulus_fn: Callable[[float], float]) -> None:
        self.__stim = stimulus_fn


def _unpack_globals(state_dict: dict):
    s = gw_cxx.GWGlobalState()
    for k in ("V", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "m", "h", "j", "xKs", "XKr", "XKv14", "XKv43"):
        setattr(s, k, state_dict[k])
    return s


def _unpack_crus(state_dict: dict):
    n = np.asarray(state_dict["CaSS"]).shape[0]
    s = gw_cxx.GWCRUState(n)
    for k in ("CaSS", "CaJSR", "LCC", "LCC_inactivation", "RyR", "ClCh"):
        setattr(s, k, state_dict[k])
    return s
