"""Basic stencils on the MI355X (stencils/cuda_hip/basic.py:101-132).

Copy, one-sided and symmetric averages and the Laplacian, computed only on
the interior with the verification oracle's operation order (bit-exact in
FP64).  The kernels vectorise along the unit-stride axis (dwordx4 loads
when the pitches allow) or march along the second axis; the FP64 Laplacian
along x and y on whole 64-column tiles takes a tile staged into LDS by
global-to-LDS loads (``variant`` 3, the library's default there), see
csrc/basic.hip.
"""

import numpy as np

from .... import native
from ..base import (
    CopyStencil,
    EmptyStencil,
    LaplacianStencil,
    OnesidedAverageStencil,
    SymmetricAverageStencil,
)
from .mixin import StencilMixin


class BasicStencilMixin(StencilMixin):
    @property
    def written_fields(self):
        return ("out",)

    def compulsory_bytes(self):
        return 2 * int(np.prod(self.domain)) * self.dtype_size


class Empty(BasicStencilMixin, EmptyStencil):
    """Times an empty launch (no kernel work); verifies nothing."""

    @property
    def written_fields(self):
        return ()

    def launch(self, pointers, options, seconds):
        geom = native.Geom(0, 0, 0, *self.strides, *self.halo)
        native.call("sb_copy", self.precision, pointers[0], pointers[1], geom, options, seconds)


class Copy(BasicStencilMixin, CopyStencil):
    def launch(self, pointers, options, seconds):
        native.call("sb_copy", self.precision, pointers[0], pointers[1], self.geom, options, seconds)


class OnesidedAverage(BasicStencilMixin, OnesidedAverageStencil):
    def launch(self, pointers, options, seconds):
        native.call(
            "sb_onesided_average", self.precision, self.axis, pointers[0], pointers[1],
            self.geom, options, seconds,
        )

    def compulsory_bytes(self):
        grown = list(self.domain)
        grown[self.axis] += 1
        return (int(np.prod(grown)) + int(np.prod(self.domain))) * self.dtype_size


class SymmetricAverage(BasicStencilMixin, SymmetricAverageStencil):
    def launch(self, pointers, options, seconds):
        native.call(
            "sb_symmetric_average", self.precision, self.axis, pointers[0], pointers[1],
            self.geom, options, seconds,
        )

    def compulsory_bytes(self):
        grown = list(self.domain)
        grown[self.axis] += 2
        return (int(np.prod(grown)) + int(np.prod(self.domain))) * self.dtype_size


class Laplacian(BasicStencilMixin, LaplacianStencil):
    def launch(self, pointers, options, seconds):
        native.call(
            "sb_laplacian", self.precision, int(self.along_x), int(self.along_y),
            int(self.along_z), pointers[0], pointers[1], self.geom, options, seconds,
        )

    def compulsory_bytes(self):
        grown = [d + 2 * int(a) for d, a in zip(self.domain, self.along_axes)]
        return (int(np.prod(grown)) + int(np.prod(self.domain))) * self.dtype_size
