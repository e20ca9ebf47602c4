"""MI355X-native Greenstein-Winslow coupled integrator.

Mirrors the reference package layout (src/__init__.py:3-9): this directory is
appended to sys.path so that `build.GreensteinWinslow` resolves to the
pybind11 module built into ./build, and the Python API is re-exported.
"""
import os
import sys

_HERE = os.path.dirname(os.path.realpath(__file__))
if _HERE not in sys.path:
    sys.path.append(_HERE)

from . import capi  # noqa: E402,F401
from .utils import *  # noqa: E402,F401,F403
from .parameters import GWParameters  # noqa: E402,F401
from .model import ICaL, Ito2, GWModel, GWSolution  # noqa: E402,F401
from .serialization import sim_dict_reduced_memory  # noqa: E402,F401
from .utils import assert_positive, assert_nonnegative, assert_type  # noqa: E402,F401
import build.GreensteinWinslow as gw  # noqa: E402,F401
