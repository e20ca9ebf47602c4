"""GWParameters: the reference's validated parameter wrapper (src/parameters.py:446-1660).

Wraps the `GWParameters` struct of the GreensteinWinslow module. Every field
is a property whose setter validates like the reference: "positive" fields
must be > 0, "nonneg" fields >= 0, NCaRU and NCaRU_sim must be int and > 0.
NCaRU_sim (the number of CRUs actually simulated) defaults to 1250
(src/parameters.py:459).
"""
from __future__ import annotations

import build.GreensteinWinslow as gw  # type: ignore  # noqa: E402

from .utils import assert_nonnegative, assert_positive, assert_type

# (name, constraint, description). Order and constraints follow the reference
# (src/parameters.py:9-440 and the setters at :500-1660); defaults come from
# the C struct (include/gw_capi.h, GW_utils.hpp:40-156).
_NN = "nonneg"
_P = "positive"
PARAMETER_TABLE = [
    ("T", _P, "Temperature [K]"), ("CSA", _P, "Membrane capacitance [pF]"), ("Vcyto", _P, "Cytosol volume [pL]"),
    ("VNSR", _P, "Network SR volume [pL]"), ("VJSR", _P, "Junctional SR volume [pL]"),
    ("VSS", _P, "Dyadic subspace volume [pL]"), ("NCaRU", _P, "Number of calcium release units in the cell"),
    ("Ko", _P, "Extracellular K [mM]"), ("Nao", _P, "Extracellular Na [mM]"), ("Cao", _P, "Extracellular Ca [mM]"),
    ("Clo", _P, "Extracellular Cl [mM]"), ("Clcyto", _P, "Cytosolic Cl [mM]"),
    ("f", _NN, "LCC 5 -> 6 (opening) rate [1/ms]"), ("g", _NN, "LCC 6 -> 5 (closing) rate [1/ms]"),
    ("f1", _NN, "LCC 11 -> 12 (opening, Ca mode) rate [1/ms]"), ("g1", _NN, "LCC 12 -> 11 (closing, Ca mode) rate [1/ms]"),
    ("a", _NN, "LCC mode-switch factor a"), ("b", _P, "LCC mode-switch factor b"),
    ("gamma0", _NN, "LCC Ca-dependent inactivation rate [1/(mM ms)]"), ("omega", _NN, "LCC recovery rate [1/ms]"),
    ("PCaL", _P, "LCC permeability [cm^3/s]"), ("kfClCh", _NN, "ClCh opening rate [1/(mM ms)]"),
    ("kbClCh", _NN, "ClCh closing rate [1/ms]"), ("Pto2", _P, "ClCh permeability [cm^3/s]"),
    ("k12", _NN, "RyR 1 -> 2 rate [1/(mM^2 ms)]"), ("k21", _NN, "RyR 2 -> 1 rate [1/ms]"),
    ("k23", _NN, "RyR 2 -> 3 rate [1/(mM^2 ms)]"), ("k32", _NN, "RyR 3 -> 2 rate [1/ms]"),
    ("k34", _NN, "RyR 3 -> 4 rate [1/(mM^2 ms)]"), ("k43", _NN, "RyR 4 -> 3 rate [1/ms]"),
    ("k45", _NN, "RyR 4 -> 5 rate [1/ms]"), ("k54", _NN, "RyR 5 -> 4 rate [1/(mM^2 ms)]"),
    ("k56", _NN, "RyR 5 -> 6 rate [1/(mM^2 ms)]"), ("k65", _NN, "RyR 6 -> 5 rate [1/ms]"),
    ("k25", _NN, "RyR 2 -> 5 rate [1/(mM^2 ms)]"), ("k52", _NN, "RyR 5 -> 2 rate [1/ms]"),
    ("rRyR", _P, "RyR release rate [1/ms]"), ("rxfer", _P, "Subspace -> cytosol transfer rate [1/ms]"),
    ("rtr", _P, "NSR -> JSR transfer rate [1/ms]"), ("riss", _P, "Inter-subspace transfer rate [1/ms]"),
    ("BSRT", _P, "Subspace SR-membrane buffer total [mM]"), ("KBSR", _P, "SR-membrane buffer affinity [mM]"),
    ("BSLT", _P, "Subspace sarcolemmal buffer total [mM]"), ("KBSL", _P, "Sarcolemmal buffer affinity [mM]"),
    ("CSQNT", _P, "Calsequestrin total [mM]"), ("KCSQN", _P, "Calsequestrin affinity [mM]"),
    ("CMDNT", _P, "Calmodulin total [mM]"), ("KCMDN", _P, "Calmodulin affinity [mM]"),
    ("GNa", _P, "INa conductance [mS/uF]"), ("GKr", _P, "IKr conductance [mS/uF]"),
    ("Kf", _NN, "IKr chain forward rate [1/ms]"), ("Kb", _NN, "IKr chain backward rate [1/ms]"),
    ("GKs", _P, "IKs conductance [mS/uF]"), ("GKv43", _P, "Kv4.3 conductance [mS/uF]"),
    ("alphaa0Kv43", _NN, "Kv4.3 activation rate scale [1/ms]"), ("aaKv43", _NN, "Kv4.3 activation exponent [1/mV]"),
    ("betaa0Kv43", _NN, "Kv4.3 deactivation rate scale [1/ms]"), ("baKv43", _NN, "Kv4.3 deactivation exponent [1/mV]"),
    ("alphai0Kv43", _NN, "Kv4.3 recovery rate scale [1/ms]"), ("aiKv43", _NN, "Kv4.3 recovery exponent [1/mV]"),
    ("betai0Kv43", _NN, "Kv4.3 inactivation rate scale [1/ms]"), ("biKv43", _NN, "Kv4.3 inactivation exponent [1/mV]"),
    ("f1Kv43", _P, "Kv4.3 allosteric factor f1"), ("f2Kv43", _P, "Kv4.3 allosteric factor f2"),
    ("f3Kv43", _P, "Kv4.3 allosteric factor f3"), ("f4Kv43", _P, "Kv4.3 allosteric factor f4"),
    ("b1Kv43", _P, "Kv4.3 allosteric factor b1"), ("b2Kv43", _P, "Kv4.3 allosteric factor b2"),
    ("b3Kv43", _P, "Kv4.3 allosteric factor b3"), ("b4Kv43", _P, "Kv4.3 allosteric factor b4"),
    ("PKv14", _P, "Kv1.4 permeability [cm/s]"),
    ("alphaa0Kv14", _NN, "Kv1.4 activation rate scale [1/ms]"), ("aaKv14", _NN, "Kv1.4 activation exponent [1/mV]"),
    ("betaa0Kv14", _NN, "Kv1.4 deactivation rate scale [1/ms]"), ("baKv14", _NN, "Kv1.4 deactivation exponent [1/mV]"),
    ("alphai0Kv14", _NN, "Kv1.4 recovery rate scale [1/ms]"), ("aiKv14", _NN, "Kv1.4 recovery exponent [1/mV]"),
    ("betai0Kv14", _NN, "Kv1.4 inactivation rate scale [1/ms]"), ("biKv14", _NN, "Kv1.4 inactivation exponent [1/mV]"),
    ("f1Kv14", _P, "Kv1.4 allosteric factor f1"), ("f2Kv14", _P, "Kv1.4 allosteric factor f2"),
    ("f3Kv14", _P, "Kv1.4 allosteric factor f3"), ("f4Kv14", _P, "Kv1.4 allosteric factor f4"),
    ("b1Kv14", _P, "Kv1.4 allosteric factor b1"), ("b2Kv14", _P, "Kv1.4 allosteric factor b2"),
    ("b3Kv14", _P, "Kv1.4 allosteric factor b3"), ("b4Kv14", _P, "Kv1.4 allosteric factor b4"),
    ("Csc", _P, "Specific membrane capacitance [uF/cm^2]"), ("GK1", _P, "IK1 conductance [mS/uF]"),
    ("KmK1", _P, "IK1 K half-saturation [mM]"), ("GKp", _P, "IKp conductance [mS/uF]"),
    ("kNaCa", _P, "Na/Ca exchanger scale [pA/pF]"), ("KmNa", _P, "Exchanger Na half-saturation [mM]"),
    ("KmCa", _P, "Exchanger Ca half-saturation [mM]"), ("ksat", _P, "Exchanger saturation factor"),
    ("eta", _P, "Exchanger voltage-dependence position"), ("INaKmax", _P, "Na/K pump maximum [pA/pF]"),
    ("KmNai", _P, "Na/K pump Na half-saturation [mM]"), ("KmKo", _P, "Na/K pump K half-saturation [mM]"),
    ("IpCamax", _P, "Sarcolemmal Ca pump maximum [pA/pF]"), ("KmpCa", _P, "Sarcolemmal Ca pump half-saturation [mM]"),
    ("GCab", _P, "Background Ca conductance [mS/uF]"), ("GNab", _P, "Background Na conductance [mS/uF]"),
    ("kHTRPNp", _P, "High-affinity troponin on-rate [1/(mM ms)]"), ("kHTRPNm", _P, "High-affinity troponin off-rate [1/ms]"),
    ("kLTRPNp", _P, "Low-affinity troponin on-rate [1/(mM ms)]"), ("kLTRPNm", _P, "Low-affinity troponin off-rate [1/ms]"),
    ("HTRPNtot", _P, "High-affinity troponin total [mM]"), ("LTRPNtot", _P, "Low-affinity troponin total [mM]"),
    ("Vmaxf", _P, "SERCA forward maximum [mM/ms]"), ("Vmaxr", _P, "SERCA reverse maximum [mM/ms]"),
    ("Kmf", _P, "SERCA forward half-saturation [mM]"), ("Kmr", _P, "SERCA reverse half-saturation [mM]"),
    ("Hf", _P, "SERCA forward Hill coefficient"), ("Hr", _P, "SERCA reverse Hill coefficient"),
]
PARAMETER_NAMES = [n for n, _, _ in PARAMETER_TABLE]


class GWParameters:
    """Greenstein and Winslow model parameters."""

    def __init__(self, cxx_struct=None, **kwargs):
        self.__cxx_struct = cxx_struct if cxx_struct is not None else gw.GWParameters()
        self.__NCaRU_sim = 1250
        if "NCaRU_sim" in kwargs:
            self.NCaRU_sim = kwargs.pop("NCaRU_sim")
        for name, value in kwargs.items():
            if name in PARAMETER_NAMES:
                setattr(self, name, value)
            else:
                raise ValueError(f"{name} not a valid argument to GWParameters constructor. "
                                 "Call help(GWParameters) to see a list of acceptable parameter names")

    @classmethod
    def from_dict(cls, state_dict: dict):
        cxx = gw.GWParameters()
        for name in PARAMETER_NAMES:
            setattr(cxx, name, state_dict[name])
        obj = cls(cxx)
        if "NCaRU_sim" in state_dict:
            obj.NCaRU_sim = int(state_dict["NCaRU_sim"])
        return obj

    def to_dict(self) -> dict:
        d = {"NCaRU_sim": self.NCaRU_sim}
        for name in PARAMETER_NAMES:
            d[name] = getattr(self, name)
        return d

    @property
    def cxx_struct(self):
        """The C++ struct passed to GreensteinWinslow.run."""
        return self.__cxx_struct

    @property
    def NCaRU_sim(self) -> int:
        """int: Number of calcium release units to simulate. Value must be > 0. Defaults to 1250."""
        return self.__NCaRU_sim

    @NCaRU_sim.setter
    def NCaRU_sim(self, value: int) -> None:
        assert_type(value, int, "NCaRU_sim")
        assert_positive(value, "NCaRU_sim")
        self.__NCaRU_sim = value


def _make_property(name: str, constraint: str, doc: str):
    def getter(self):
        return getattr(self.cxx_struct, name)

    if name == "NCaRU":
        def setter(self, value):
            assert_type(value, int, name)
            assert_positive(value, name)
            setattr(self.cxx_struct, name, value)
    elif constraint == _P:
        def setter(self, value):
            assert_positive(value, name)
            setattr(self.cxx_struct, name, value)
    else:
        def setter(self, value):
            assert_nonnegative(value, name)
            setattr(self.cxx_struct, name, value)

    dflt = getattr(gw.GWParameters(), name)
    bound = "> 0" if constraint == _P else ">= 0"
    return property(getter, setter, doc=f"{'int' if name == 'NCaRU' else 'float'}: {doc}. Value must be {bound}. "
                                        f"Defaults to {dflt}.")


for _name, _c, _doc in PARAMETER_TABLE:
    setattr(GWParameters, _name, _make_property(_name, _c, _doc))
del _name, _c, _doc
