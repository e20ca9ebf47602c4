"""Shared pytest configuration: the `gpu` marker and oracle/library fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O

    if not os.path.exists(os.path.join(O.BUILD, "liboracle.so")):
        O.build()
    return O.lib("strict")


@pytest.fixture(scope="session")
def golden():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json"), encoding="utf-8") as f:
        return json.load(f)


def _load_capi():
    import importlib.util

    spec = importlib.util.spec_from_file_location("gw_capi_binding", os.path.join(ROOT, "gw-model_amd", "capi.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def gwcapi():
    return _load_capi()


@pytest.fixture(scope="session")
def gpu_ctx(gwcapi):
    ctx = gwcapi.Context(0)
    yield ctx
    ctx.close()
