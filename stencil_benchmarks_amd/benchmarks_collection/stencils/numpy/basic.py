"""numpy basic stencils: the reference's numpy backend for copy, averages and
the Laplacian (stencils/numpy/basic.py:44-87 there).

Each kernel works on shifted views of the input (``_shifted(data.inp, axis,
offset)``, the interior moved by `offset` points along `axis`) and writes the
interior of the output.  The operation order is the reference numpy
backend's: averages as (a + b) / 2, the Laplacian as 2n inp minus each
neighbour in turn (axes in order, +1 then -1), so results equal that
backend's bit for bit.
"""

import numpy as np

from ....tools import timing
from ..base import CopyStencil, LaplacianStencil, OnesidedAverageStencil, SymmetricAverageStencil


def _shifted(bench, field, axis=None, offset=0):
    """The interior of `field`, moved `offset` points along `axis`."""
    move = np.zeros(len(bench.domain), dtype=int)
    if axis is not None:
        move[axis] = offset
    return field[bench.inner_slice(move)]


class Copy1D(CopyStencil):
    """Copies the whole buffer, halo included."""

    @timing.return_time
    def run_stencil(self, data):
        np.copyto(data.out, data.inp)


class Copy(CopyStencil):
    @timing.return_time
    def run_stencil(self, data):
        data.out[self.inner_slice()] = _shifted(self, data.inp)


class OnesidedAverage(OnesidedAverageStencil):
    @timing.return_time
    def run_stencil(self, data):
        ahead = _shifted(self, data.inp, self.axis, +1)
        data.out[self.inner_slice()] = (ahead + _shifted(self, data.inp)) / 2


class SymmetricAverage(SymmetricAverageStencil):
    @timing.return_time
    def run_stencil(self, data):
        ahead = _shifted(self, data.inp, self.axis, +1)
        behind = _shifted(self, data.inp, self.axis, -1)
        data.out[self.inner_slice()] = (ahead + behind) / 2


class Laplacian(LaplacianStencil):
    """2n inp - in[+e] - in[-e] - ... over the n enabled axes."""

    @timing.return_time
    def run_stencil(self, data):
        axes = [axis for axis, on in enumerate(self.along_axes) if on]
        result = (2 * len(axes)) * _shifted(self, data.inp)
        for axis in axes:
            result -= _shifted(self, data.inp, axis, +1)
            result -= _shifted(self, data.inp, axis, -1)
        data.out[self.inner_slice()] = result
