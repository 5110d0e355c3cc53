"""Error of the float32 forms against the FP64 oracle - CPU only.

The GPU kernels equal two oracle computations bit for bit (asserted by
test_gpu_parity.py on the same inputs):
  * float32 storage, float32 arithmetic  == the float32 oracle
    (every operation rounded to float32, the reference's numpy order);
  * float32 storage, FP64 arithmetic     == the FP64 oracle on the stored
    float32 inputs, rounded once to float32.
So their error against the FP64 oracle (on the same float32-representable
inputs: only arithmetic error is measured) follows from the oracle alone,
at the benchmark sizes.  Metrics: max-abs, normwise = max-abs / max|ref|,
pointwise max-rel over points with |ref| >= 1e-3.

Bounds (the float32 form must stay within 1e-5 relative):
  FP64 arithmetic, any stencil:  normwise <= 6e-8 (one float32 rounding)
  float32 arithmetic, copy / averages / Laplacian / hdiff: normwise <= 1e-6
  float32 arithmetic, vadv: exceeds 1e-5 on the weakly diagonally dominant
    columns of U[0,1) inputs - why the Thomas solve keeps FP64 arithmetic.
"""

import numpy as np
import pytest

from conftest import inner, make_fields
from oracle import oracle

FLOOR = 1e-3


def _widen(a):
    out = np.empty(a.shape, np.float64)
    out[...] = a
    return out


def stats(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    diff = np.abs(got - ref)
    mask = np.abs(ref) >= FLOOR
    pointwise = float((diff[mask] / np.abs(ref[mask])).max()) if mask.any() else 0.0
    return {"max_abs": float(diff.max()), "normwise": float(diff.max() / np.abs(ref).max()),
            "pointwise": pointwise}


def _clone64(f32):
    from stencil_benchmarks_amd.tools import array

    c = array.alloc_array(f32.shape, np.float64, (2, 1, 0), 0, (0, 0, 0))
    c[...] = f32
    return c


def forms(func, fields32):
    """(float32-arithmetic, FP64-arithmetic rounded once) stats vs FP64 oracle."""
    ref = func(*[_clone64(f) for f in fields32])
    f32 = func(*fields32)
    f64c = ref.astype(np.float32)
    return stats(f32, ref), stats(f64c, ref)


CASES = {
    "copy": ((64, 64, 40), (1, 1, 1), 1, lambda d, h: lambda a: oracle.copy(a, d, h)),
    "onesided_average": ((64, 64, 40), (1, 1, 1), 1,
                         lambda d, h: lambda a: oracle.onesided_average(a, 0, d, h)),
    "symmetric_average": ((64, 64, 40), (1, 1, 1), 1,
                          lambda d, h: lambda a: oracle.symmetric_average(a, 1, d, h)),
    "laplacian": ((128, 128, 40), (1, 1, 1), 1,
                  lambda d, h: lambda a: oracle.laplacian(a, (1, 1, 0), d, h)),
    "hdiff": ((128, 128, 80), (3, 3, 3), 2,
              lambda d, h: lambda a, b: oracle.hdiff(a, b, d, h)),
    "vadv": ((256, 256, 80), (3, 3, 3), 5,
             lambda d, h: lambda *f: oracle.vadv(*f, d, h)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_float32_forms_error(name):
    domain, halo, nfields, make = CASES[name]
    fields = make_fields(domain, halo, np.float32, count=nfields, seed=11)
    f32, f64c = forms(make(domain, halo), fields)
    assert f64c["normwise"] <= 6e-8, f64c
    if name == "copy":
        assert f32["max_abs"] == 0.0 and f64c["max_abs"] == 0.0
    elif name == "vadv":
        assert f32["normwise"] > 1e-5, f32  # documented: FP32 Thomas breaks the bound
    else:
        assert f32["normwise"] <= 1e-6, f32


def table():
    """Markdown table of the error statistics (used to write the design notes)."""
    rows = ["| stencil | size | float32 arithmetic: max-abs / normwise / pointwise | "
            "FP64 arithmetic: max-abs / normwise / pointwise |", "|---|---|---|---|"]
    for name in sorted(CASES):
        domain, halo, nfields, make = CASES[name]
        fields = make_fields(domain, halo, np.float32, count=nfields, seed=11)
        f32, f64c = forms(make(domain, halo), fields)
        fmt = lambda s: f"{s['max_abs']:.2e} / {s['normwise']:.2e} / {s['pointwise']:.2e}"
        rows.append(f"| {name} | {'x'.join(map(str, domain))} | {fmt(f32)} | {fmt(f64c)} |")
    return "\n".join(rows)


if __name__ == "__main__":
    print(table())
