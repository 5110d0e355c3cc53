"""Seeded uniform [0, 1) field fill.

The reference fills every field with unseeded U[0,1) values from one
``minstd_rand`` per thread (stencil_benchmarks/tools/parallel.cpp:51-96), so
its inputs are not reproducible.  Here every field draws from
``numpy.random.default_rng`` over its full halo-inclusive logical extent,
in logical (i, j, k) C order, so a (seed, data set, field) triple always
yields the same values whatever the layout, padding or alignment.
"""

import numpy as np


def field_rng(seed: int, data_set: int, field_index: int) -> np.random.Generator:
    return np.random.default_rng([int(seed), int(data_set), int(field_index)])


def random_fill(array: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """Fill the logical extent of `array` with U[0,1) (or 0/1 for ints)."""
    if array.dtype == np.float32:
        array[...] = rng.random(array.shape, dtype=np.float32)
        return array
    values = rng.random(array.shape, dtype=np.float64)
    if array.dtype.kind == "f":
        array[...] = values.astype(array.dtype)
    elif array.dtype.kind == "b":
        array[...] = values < 0.5
    else:
        array[...] = (values * 2).astype(array.dtype)
    return array
