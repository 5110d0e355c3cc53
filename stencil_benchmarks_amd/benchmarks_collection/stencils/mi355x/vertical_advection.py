"""Vertical advection on the MI355X.

One lane per (i, j) column with contiguous i, forward sweep and fused
backward substitution in one kernel (csrc/vadv.hip).  ``Fused`` picks the
solver: one wave per 64-column group with the lowest levels' Thomas
coefficients in registers where every group of the domain fits on the chip
at once (nz <= 96), else the warp-specialised one (four waves per group)
for nz >= 96 (nz >= 48 with float32 storage) and the single-wave pipeline
below that, both with the Thomas coefficients in LDS (nz <= 160 in FP64
arithmetic), and a global FP64 scratch beyond.  ``LdsColumn``,
``GlobalScratch``, ``Pipelined``, ``WarpSpecialised`` and
``RegisterColumns`` pin one solver (``WarpSpecialised`` with
``levels_per_step`` 4 or 8 per workgroup barrier).  Arithmetic follows
stencils/base.py:415-473 (bit-exact in FP64; with float32 storage the
default arithmetic is FP64 throughout).
"""

import ctypes

import numpy as np

from .... import native
from ....benchmark import Parameter
from ..base import VerticalAdvectionStencil
from .mixin import StencilMixin


class VerticalAdvectionMixin(StencilMixin):
    VARIANT = 0
    written_fields_are_read = True

    @property
    def written_fields(self):
        return tuple(f"{c}tensstage" for c, _ in self.components)

    def launch_options(self):
        options = super().launch_options()
        options.variant = self.variant or self.VARIANT
        return options

    def launch(self, pointers, options, seconds):
        array = (ctypes.c_void_p * len(pointers))(*[p.value for p in pointers])
        native.call(
            "sb_vadv", self.precision, int(self.all_components), array, self.geom, options, seconds
        )

    def compulsory_bytes(self):
        nx, ny, nz = self.domain
        n = nx * ny * nz
        ncomp = len(self.components)
        # per component: stage, pos, tens, tensstage read + tensstage write;
        # wcon (with its i+1 / j+1 neighbour column) read once
        wcon = (nx + 1) * (ny + (1 if self.all_components else 0)) * nz
        return (5 * ncomp * n + wcon) * self.dtype_size


class Fused(VerticalAdvectionMixin, VerticalAdvectionStencil):
    VARIANT = 0


class LdsColumn(VerticalAdvectionMixin, VerticalAdvectionStencil):
    VARIANT = 1


class GlobalScratch(VerticalAdvectionMixin, VerticalAdvectionStencil):
    VARIANT = 2


class Pipelined(VerticalAdvectionMixin, VerticalAdvectionStencil):
    VARIANT = 3


class WarpSpecialised(VerticalAdvectionMixin, VerticalAdvectionStencil):
    VARIANT = 4
    levels_per_step = Parameter("levels per workgroup barrier step", 4, choices=[4, 8])

    def launch_options(self):
        options = super().launch_options()
        options.unroll = self.levels_per_step
        return options


class RegisterColumns(VerticalAdvectionMixin, VerticalAdvectionStencil):
    VARIANT = 5
