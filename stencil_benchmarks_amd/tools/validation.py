"""Result verification with the reference's tolerances.

Follows stencil_benchmarks/tools/validation.py:37-116: ``np.isclose`` with
``rtol=1e-4, atol=1e-5`` for float32 and numpy's defaults otherwise; a
failure raises :class:`ValidationError` listing the failing points.
:func:`check_close` adds an explicit-tolerance form for the stricter checks
this backend states (bit-exact FP64, FP32 storage bounds).
"""

import numpy as np


class ValidationError(RuntimeError):
    """Verification of a benchmark result failed (CLI exit code 2)."""


def _tolerances(dtype):
    if dtype == np.float32:
        return dict(rtol=1e-4, atol=1e-5)
    return dict()


def _report(result, expected, failure, limit=20):
    where = np.argwhere(failure)
    lines = []
    for index in where[:limit]:
        idx = tuple(int(v) for v in index)
        lines.append(
            "failed at "
            + ", ".join(f"{v:3}" for v in idx)
            + f": {result[idx]:12.7f}    != {expected[idx]:12.7f}"
        )
    if len(where) > limit:
        lines.append(f"omitted further {len(where) - limit} failures.")
    return "\n".join(lines)


def _raise(name, result, expected, failure):
    raise ValidationError(
        f"validation of field {name} failed at "
        f"{int(np.count_nonzero(failure))} points:\n"
        + _report(result, expected, failure)
    )


def check_equality(name, result, expected):
    """Reference check: elementwise isclose with dtype-dependent tolerance."""
    assert result.dtype == expected.dtype
    failure = ~np.isclose(result, expected, **_tolerances(result.dtype))
    if np.any(failure):
        _raise(name, result, expected, failure)


def check_close(name, result, expected, rtol, atol=0.0):
    """Elementwise check with explicit tolerances (rtol=atol=0: bit-exact)."""
    if rtol == 0 and atol == 0:
        failure = ~((result == expected) | (np.isnan(result) & np.isnan(expected)))
    else:
        failure = ~np.isclose(result, expected, rtol=rtol, atol=atol)
    if np.any(failure):
        _raise(name, result, expected, failure)


def error_stats(result, expected, floor=1e-3):
    """Max-abs, normwise and floored pointwise relative error of `result`.

    ``pointwise_rel`` is max |r-e|/|e| over points with |e| > `floor`.
    """
    r = np.asarray(result, dtype=np.float64)
    e = np.asarray(expected, dtype=np.float64)
    diff = np.abs(r - e)
    max_abs = float(diff.max()) if diff.size else 0.0
    norm_e = float(np.abs(e).max()) if e.size else 0.0
    mask = np.abs(e) > floor
    pointwise = float((diff[mask] / np.abs(e[mask])).max()) if mask.any() else 0.0
    return dict(
        max_abs=max_abs,
        normwise_rel=max_abs / norm_e if norm_e else 0.0,
        pointwise_rel=pointwise,
        rel_floor=floor,
    )
