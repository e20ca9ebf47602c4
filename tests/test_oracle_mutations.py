"""The oracle pins fail when a plausible transcription error is planted in the oracle.

Each case copies oracle/gw_oracle.c, changes one token (a coefficient, an
exponent sign, a stoichiometric factor, an index), builds the copy with the
checker's flags into a temporary directory, and runs every check of
test_oracle_pins.py against it: at least one must fail. The unmodified
source, built the same way, must pass them all. Mistakes are planted only in
this temporary copy of the oracle, on the CPU.
"""
import inspect
import os
import subprocess

import pytest

import tests.test_oracle_pins as pins
from oracle import oracle as O

SRC = os.path.join(O.HERE, "gw_oracle.c")
INCLUDE = os.path.join(os.path.dirname(O.HERE), "include")

# (what the mistake is, text in gw_oracle.c, replacement)
MUTATIONS = [
    ("alpha_LCC slope 0.012 -> 0.12", "2.0 * gwo_exp(0.012 * (V - 35.0))", "2.0 * gwo_exp(0.12 * (V - 35.0))"),
    ("beta_LCC exponent sign", "0.0882 * gwo_exp(-0.05 * (V - 35.0))", "0.0882 * gwo_exp(0.05 * (V - 35.0))"),
    ("y_inf midpoint sign", "gwo_exp((V + 12.5) / 5.0)", "gwo_exp((V - 12.5) / 5.0)"),
    ("tau_LCC amplitude", "340.0 / (1.0 + gwo_exp((V + 30.0) / 12.0))", "34.0 / (1.0 + gwo_exp((V + 30.0) / 12.0))"),
    ("XKs_inf slope", "(V - 24.7) / 13.6", "(V - 24.7) / 1.36"),
    ("tau_XKs second rate", "gwo_exp(0.0687 * (V - 10.0)) - 1.0", "gwo_exp(0.0678 * (V - 10.0)) - 1.0"),
    ("IKr factor 0.5 dropped", "(V - c->EK) * 0.5;", "(V - c->EK);"),
    ("IKr sqrt(Ko) replaced by Ko", "c->sqrtKo = sqrt(p->Ko);", "c->sqrtKo = p->Ko;"),
    ("IKr rectification constant", "1.0 / (1.0 + 1.4945 * gwo_exp(0.0446 * V))", "1.0 / (1.0 + 1.4945 * gwo_exp(0.446 * V))"),
    ("IKp midpoint", "gwo_exp((7.488 - V) / 5.98)", "gwo_exp((7.488 + V) / 5.98)"),
    ("Jup Hill coefficient Hr for Hf", "gwo_pow(g->Cai / p->Kmf, p->Hf)", "gwo_pow(g->Cai / p->Kmf, p->Hr)"),
    ("Kv1.4 activation exponent sign", "p->alphaa0Kv14 * gwo_exp(p->aaKv14 * V)", "p->alphaa0Kv14 * gwo_exp(-p->aaKv14 * V)"),
    ("Kr 3->4 exponent sign", "KR(3, 4) = 0.0622 * gwo_exp(0.0120 * V)", "KR(3, 4) = 0.0622 * gwo_exp(-0.0120 * V)"),
    ("Kr master equation index", "KR(3, 2) * x[3] + KR(4, 2)", "KR(2, 3) * x[3] + KR(4, 2)"),
    ("Na/K pump Na stoichiometry", "3 * INaCa + 3 * INaK)", "3 * INaCa + 2 * INaK)"),
    ("Na/K pump K stoichiometry", "IKp - 2 * INaK)", "IKp - 3 * INaK)"),
    ("exchanger Ca stoichiometry", "(ICab + IpCa - 2 * INaCa)", "(ICab + IpCa - INaCa)"),
    ("Ca charge factor", "-0.5 * c->CSA_FVcyto * (ICab", "-1.0 * c->CSA_FVcyto * (ICab"),
    ("uptake sign in dCai", "c->VSS_Vcyto * Jxfer_tot - Jup", "c->VSS_Vcyto * Jxfer_tot + Jup"),
    ("NSR volume ratio", "c->VJSR_VNSR * Jtr_tot)", "c->VSS_Vcyto * Jtr_tot)"),
]


def _build(tmp_path, text, name):
    src = tmp_path / f"{name}.c"
    src.write_text(text)
    so = tmp_path / f"{name}.so"
    subprocess.run(["gcc", "-std=c11", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                    f"-I{INCLUDE}", f"-I{O.HERE}", str(src), "-o", str(so), "-lm"], check=True, capture_output=True)
    return O.load(str(so))


def _pin_failures(L):
    """Run every check of test_oracle_pins.py on library L; return the names that fail."""
    failed = []
    for name, fn in inspect.getmembers(pins, inspect.isfunction):
        if not name.startswith("test_"):
            continue
        params = list(inspect.signature(fn).parameters)
        marks = getattr(fn, "pytestmark", [])
        cases = [{}]
        for m in marks:
            if m.name == "parametrize":
                argname, values = m.args[0], m.args[1]
                cases = [{argname: v} for v in values]
        for kw in cases:
            try:
                fn(L, **kw) if params and params[0] == "oracle_lib" else fn(**kw)
            except (AssertionError, ArithmeticError):
                failed.append(f"{name}{kw or ''}")
    return failed


def test_unmodified_oracle_passes_the_pins(tmp_path):
    with open(SRC, encoding="utf-8") as f:
        text = f.read()
    assert _pin_failures(_build(tmp_path, text, "unmodified")) == []


@pytest.mark.parametrize("what,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_fails_a_pin(tmp_path, what, old, new):
    with open(SRC, encoding="utf-8") as f:
        text = f.read()
    assert text.count(old) >= 1, f"mutation site not found: {old}"
    failed = _pin_failures(_build(tmp_path, text.replace(old, new), "mutant"))
    assert failed, f"no pin caught: {what}"
