"""Reduced-memory records of a simulation (src/serialization.py:11-39).

sim_dict_reduced_memory converts the records of a GWSolution on the host, as
the reference does. GWModel.simulate_reduced returns the same dictionary formed
on the device (GreensteinWinslow.run_reduced, gw_run_reduced in the C-ABI), so
that the full snapshots (184 bytes per CRU and record) never cross PCIe; the
reduced ones take 40.

The reference's HDF5 writers (save_sim, save_sim_group,
src/serialization.py:42-79) need h5py, which this environment does not have;
they are not provided.
"""
from __future__ import annotations

from typing import Any, Dict

import numpy as np

from .utils import RyR_state_to_int16

# records kept as FP32 (src/serialization.py:14-30)
_FP32 = ("t", "V", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "m", "h", "j", "xKs", "CaSS", "CaJSR")


def sim_dict_reduced_memory(sim: Any) -> Dict:
    """The records of `sim` (anything with to_dict()["vars"]) in the reduced
    encoding: FP32 concentrations and globals, the first 4 of 5 IKr states and
    the first 9 of 10 Kv states, LCC as int8, y and ClCh as bool, and the int16
    RyR code (src/utils.py:8-22)."""
    d = sim.to_dict()["vars"]
    for name in _FP32:
        d[name] = np.float32(d[name])
    d["XKr"] = np.float32(d["XKr"][:, :4])
    d["XKv14"] = np.float32(d["XKv14"][:, :9])
    d["XKv43"] = np.float32(d["XKv43"][:, :9])
    d["LCC"] = np.int8(d["LCC"])
    d["LCC_inactivation"] = np.bool_(d["LCC_inactivation"])
    d["ClCh"] = np.bool_(d["ClCh"])
    d["RyR"] = RyR_state_to_int16(d["RyR"])
    return d
