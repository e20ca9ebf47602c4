"""Validators and encodings of the reference's src/utils.py.

The validators keep the reference's behaviour, including that the numpy
variants raise only when every element violates the bound
(src/utils.py:167,173,195,201).
"""
from typing import Any

import numpy as np
import numpy.typing as npt


def RyR_state_to_int16(RyRs: npt.NDArray) -> npt.NDArray:
    """Encode per-subunit RyR counts as one int16 (src/utils.py:8-22).

    A negative member of a lumped pair (3/4 or 5/6) is folded into the first
    member of the pair before encoding; state 6 is implied by the other counts.
    """
    RyRs = np.asarray(RyRs)
    n34 = RyRs[..., 2] + RyRs[..., 3]
    neg34 = (RyRs[..., 2] < 0) | (RyRs[..., 3] < 0)
    s3 = np.where(neg34, n34, RyRs[..., 2])
    s4 = np.where(neg34, 0, RyRs[..., 3])
    n56 = RyRs[..., 4] + RyRs[..., 5]
    neg56 = (RyRs[..., 4] < 0) | (RyRs[..., 5] < 0)
    s5 = np.where(neg56, n56, RyRs[..., 4])
    code = RyRs[..., 0] + 6 * RyRs[..., 1] + 36 * s3 + 216 * s4 + 1296 * s5
    return np.int16(code)


def assert_gt(x: Any, lb: Any, name: str, strict: bool = False) -> None:
    if strict:
        if x <= lb:
            raise ValueError(f"{name} must be > {lb}")
    elif x < lb:
        raise ValueError(f"{name} must be >= {lb}")


def assert_lt(x: Any, ub: Any, name: str, strict: bool = False) -> None:
    if strict:
        if x >= ub:
            raise ValueError(f"{name} must be < {ub}")
    elif x > ub:
        raise ValueError(f"{name} must be <= {ub}")


def assert_positive(x: Any, name: str) -> None:
    assert_gt(x, 0, name, strict=True)


def assert_nonnegative(x: Any, name: str) -> None:
    assert_gt(x, 0, name, strict=False)


def assert_type(x: Any, tp: Any, name: str) -> None:
    if not isinstance(x, tp):
        raise TypeError(f"{name} must be of type {tp}. Got type {type(x)}")


def assert_gt_numpy(x: npt.NDArray, lb, name: str, strict: bool = False) -> None:
    x = np.asarray(x)
    bad = (x <= lb) if strict else (x < lb)
    if np.all(bad):
        raise ValueError(f"All values of {name} must be {'>' if strict else '>='} {lb}. Violated at indices {bad.nonzero()}.")


def assert_lt_numpy(x: npt.NDArray, ub, name: str, strict: bool = False) -> None:
    x = np.asarray(x)
    bad = (x >= ub) if strict else (x > ub)
    if np.all(bad):
        raise ValueError(f"All values of {name} must be {'<' if strict else '<='} {ub}. Violated at indices {bad.nonzero()}.")
