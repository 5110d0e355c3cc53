"""Result verification with the reference's tolerances.

The acceptance rule is the reference's (stencil_benchmarks/tools/validation.py:
98-116): ``np.isclose`` with ``rtol=1e-4, atol=1e-5`` for float32 and numpy's
defaults otherwise.  A failure raises :class:`ValidationError` naming the
field, the count of failing points, the largest absolute and relative
differences and the worst points.
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


def _raise(name, result, expected, failure, shown=10):
    """ValidationError naming the field, the number of failing points, the
    largest differences and the `shown` worst points (index: got / expected)."""
    where = np.argwhere(failure)
    got = np.asarray(result, dtype=np.float64)[failure]
    want = np.asarray(expected, dtype=np.float64)[failure]
    diff = np.abs(got - want)
    diff = np.where(np.isnan(diff), np.inf, diff)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = diff / np.abs(want)
    worst = np.argsort(-diff, kind="stable")[:shown]
    lines = [f"{name}: {len(where)} of {failure.size} points differ "
             f"(largest |difference| {diff.max():.4g}, largest relative {np.nanmax(rel):.4g})"]
    for n in worst:
        index = ", ".join(str(int(v)) for v in where[n])
        lines.append(f"  [{index}]: {float(got[n])!r} / expected {float(want[n])!r}")
    if len(where) > shown:
        lines.append(f"  ... and {len(where) - shown} more")
    raise ValidationError("\n".join(lines))


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
