"""numpy basic stencils (stencils/numpy/basic.py:44-87 of the reference)."""

import numpy as np

from ....tools import timing
from ..base import CopyStencil, LaplacianStencil, OnesidedAverageStencil, SymmetricAverageStencil


def _unit(axis):
    shift = np.zeros(3, dtype=int)
    shift[axis] = 1
    return shift


class Copy1D(CopyStencil):
    """Copies the whole buffer, halo included."""

    @timing.return_time
    def run_stencil(self, data):
        data.out[:] = data.inp[:]


class Copy(CopyStencil):
    @timing.return_time
    def run_stencil(self, data):
        inner = self.inner_slice()
        data.out[inner] = data.inp[inner]


class OnesidedAverage(OnesidedAverageStencil):
    @timing.return_time
    def run_stencil(self, data):
        up = _unit(self.axis)
        data.out[self.inner_slice()] = (
            data.inp[self.inner_slice(up)] + data.inp[self.inner_slice()]
        ) / 2


class SymmetricAverage(SymmetricAverageStencil):
    @timing.return_time
    def run_stencil(self, data):
        up = _unit(self.axis)
        data.out[self.inner_slice()] = (
            data.inp[self.inner_slice(up)] + data.inp[self.inner_slice(-up)]
        ) / 2


class Laplacian(LaplacianStencil):
    """Sequential-subtraction form ``2n inp - in[+e] - in[-e] - ...``."""

    @timing.return_time
    def run_stencil(self, data):
        inner = self.inner_slice()
        out = data.out
        out[inner] = 2 * sum(self.along_axes) * data.inp[inner]
        for axis, enabled in enumerate(self.along_axes):
            if enabled:
                up = _unit(axis)
                out[inner] -= data.inp[self.inner_slice(up)]
                out[inner] -= data.inp[self.inner_slice(-up)]
