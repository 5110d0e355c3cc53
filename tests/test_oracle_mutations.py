"""The oracle pins catch planted mistakes - CPU only.

Each case compiles oracle/oracle.c with one plausible mistake (a flipped
sign, a dropped term, a wrong neighbour or coefficient) into a temporary
library, swaps it in for liboracle.so and checks that at least one pinning
test of test_oracle.py fails on it.
"""

import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

import test_oracle
from oracle import oracle

SOURCE = Path(__file__).resolve().parents[1] / "oracle" / "oracle.c"

# name: (original text, mutated text, pinning tests that must fail)
MUTATIONS = {
    "limiter_sign": ("(T)(flux * gradient) > (T)0", "(T)(flux * gradient) < (T)0",
                     [test_oracle.test_hdiff_delta_spike]),
    "limiter_off": ("(T)(flux * gradient) > (T)0 ? (T)0 : flux", "flux",
                    [test_oracle.test_hdiff_delta_spike]),
    "flux_neighbour": ("(T)(lc - lap_##SUF(p - si, si, sj))", "(T)(lc - lap_##SUF(p - sj, si, sj))",
                       [lambda: test_oracle.test_hdiff_quartic_closed_form(0)]),
    "laplacian_weight": ("(T)((T)4 * p[0] -", "(T)((T)2 * p[0] -",
                         [lambda: test_oracle.test_hdiff_quartic_closed_form(0)]),
    "update_sign": ("(T)(p[0] - (T)(AT(coeff, i, j, k) * sum))",
                    "(T)(p[0] + (T)(AT(coeff, i, j, k) * sum))",
                    [test_oracle.test_hdiff_delta_spike]),
    "laplacian_neighbour_sign": (
        "const T term = (T)((T)((T)2 * p[0] - p[st[a]]) -",
        "const T term = (T)((T)((T)2 * p[0] + p[st[a]]) -",
        [lambda: test_oracle.test_laplacian_known_answers((2, 1, 0))]),
    "backward_sign": ("(dd[k] - (T)(cc[k] * x))", "(dd[k] + (T)(cc[k] * x))",
                      [lambda: test_oracle.test_vadv_matches_dense_solve((1, 0))]),
    "dropped_correction": (
        "dd[k] = (T)((T)((T)(d0 + corr) - (T)(dd[k - 1] * a)) * r);",
        "dd[k] = (T)((T)((T)(d0) - (T)(dd[k - 1] * a)) * r);",
        [lambda: test_oracle.test_vadv_matches_dense_solve((1, 0))]),
    "average_neighbour": ("(T)((p[e] + p[-e]) / (T)2)", "(T)((p[e] + p[0]) / (T)2)",
                          [lambda: test_oracle.test_symmetric_average_of_quadratic_field(0)]),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(name, tmp_path):
    original, mutated, pins = MUTATIONS[name]
    text = SOURCE.read_text()
    assert original in text, f"mutation anchor for {name} not found"
    source = tmp_path / f"{name}.c"
    source.write_text(text.replace(original, mutated, 1))
    library = tmp_path / f"{name}.so"
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
                    "-shared", "-o", str(library), str(source)], check=True)
    saved = oracle.lib()
    oracle._lib = ctypes.CDLL(str(library))
    try:
        caught = 0
        for pin in pins:
            try:
                pin()
            except AssertionError:
                caught += 1
        assert caught == len(pins), f"{name}: the pins did not catch the mistake"
    finally:
        oracle._lib = saved
