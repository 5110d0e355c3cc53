"""The C oracle and the product's numpy restatements agree bit for bit.

The oracle (oracle/oracle.c) and the numpy forms inside the product
(stencils/base.py: ``horizontal_diffusion``, ``vertical_advection`` and the
``expected`` methods of the basic stencils, plus the numpy ``OnTheFly``
backend) restate the reference's verification code
(stencil_benchmarks/benchmarks_collection/stencils/base.py:180-254,
:276-311, :415-473) independently: one as scalar C loops, the other as
whole-array numpy expressions in the reference's own array form.  Closed
forms and dense solves pin what each computes but not the order of its
floating-point operations; FP64 equality of the two on random U[0,1) fields
pins that order too (a reordered sum or a fused product differs in the last
bits somewhere among the 245,760 points).  CPU only.
"""

import numpy as np
import pytest

from oracle import oracle
from stencil_benchmarks_amd.benchmarks_collection.stencils import base
from stencil_benchmarks_amd.benchmarks_collection.stencils.numpy import (
    basic as numpy_basic, horizontal_diffusion as numpy_hdiff)
from stencil_benchmarks_amd.tools import array

DOMAIN, HALO = (64, 48, 80), (3, 3, 3)
LAYOUTS = [(2, 1, 0), (0, 1, 2)]


def _fields(layout, count, seed):
    rng = np.random.default_rng(seed)
    shape = tuple(d + 2 * h for d, h in zip(DOMAIN, HALO))
    out = []
    for _ in range(count):
        a = array.alloc_array(shape, np.float64, layout, 0, HALO)
        a[...] = rng.random(shape)
        out.append(a)
    return out


def _inner():
    return tuple(slice(h, h + d) for d, h in zip(DOMAIN, HALO))


@pytest.mark.parametrize("layout", LAYOUTS)
def test_hdiff_numpy_form_and_on_the_fly_equal_the_oracle(layout):
    inp, coeff = _fields(layout, 2, seed=1)
    expected = oracle.hdiff(inp, coeff, DOMAIN, HALO)
    numpy_form = base.horizontal_diffusion(inp, coeff)[_inner()]
    assert np.array_equal(numpy_form, expected)
    bench = numpy_hdiff.OnTheFly(domain=DOMAIN, halo=HALO, layout=layout, verify=False)
    data = bench._data[0]
    data.inp[...] = inp
    data.coeff[...] = coeff
    bench.run()
    assert np.array_equal(data.out[_inner()], expected)


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("component,shift", [("u", (1, 0)), ("v", (0, 1)), ("w", (0, 0))])
def test_vadv_numpy_form_equals_the_oracle(layout, component, shift):
    stage, pos, tens, tensstage, wcon = _fields(layout, 5, seed=2)
    expected = oracle.vadv(stage, pos, tens, tensstage, wcon, DOMAIN, HALO, shift=shift)
    fields = {"stage": stage, "pos": pos, "tens": tens, "tensstage": tensstage, "wcon": wcon}
    numpy_form = base.vertical_advection(fields, DOMAIN, HALO, shift)
    assert np.array_equal(numpy_form, expected), component


def _basic(cls, layout, **params):
    bench = cls(domain=DOMAIN, halo=HALO, layout=layout, verify=False, **params)
    (inp,) = _fields(layout, 1, seed=3)
    return bench, inp


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_average_expectations_equal_the_oracle(layout, axis):
    bench, inp = _basic(numpy_basic.OnesidedAverage, layout, axis=axis)
    assert np.array_equal(bench.expected(inp), oracle.onesided_average(inp, axis, DOMAIN, HALO))
    bench, inp = _basic(numpy_basic.SymmetricAverage, layout, axis=axis)
    assert np.array_equal(bench.expected(inp), oracle.symmetric_average(inp, axis, DOMAIN, HALO))


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("along", [(True, True, False), (True, False, True), (True, True, True)])
def test_laplacian_expectation_equals_the_oracle(layout, along):
    bench, inp = _basic(numpy_basic.Laplacian, layout, along_x=along[0], along_y=along[1],
                        along_z=along[2])
    assert np.array_equal(bench.expected(inp), oracle.laplacian(inp, along, DOMAIN, HALO))


def test_a_reordered_sum_is_caught():
    """The comparison above can see an operation-order change: the
    Laplacian summed as 2n c - sum(neighbours) (the numpy backend's form)
    differs from the oracle's per-axis form on these inputs."""
    bench, inp = _basic(numpy_basic.Laplacian, (2, 1, 0))
    inner = bench.inner_slice()
    reordered = 4 * inp[inner]
    for axis in (0, 1):
        up = np.eye(3, dtype=int)[axis]
        reordered = reordered - inp[bench.inner_slice(up)] - inp[bench.inner_slice(-up)]
    expected = oracle.laplacian(inp, (True, True, False), DOMAIN, HALO)
    assert np.allclose(reordered, expected, rtol=1e-12, atol=1e-12)
    assert not np.array_equal(reordered, expected)
