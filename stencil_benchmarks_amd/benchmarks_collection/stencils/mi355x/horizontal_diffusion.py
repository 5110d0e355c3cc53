"""Horizontal diffusion on the MI355X.

``Fused`` is the LDS-tiled fused kernel (csrc/hdiff.hip): the input tile
and its 2-point halo are staged in LDS and the Laplacian, flux limiter and
update never round-trip HBM.  With the default block size the library picks
the tile: for FP64 domains of whole 64 x 8 tiles and float32 domains of
whole 64 x 16 tiles the one staged by global-to-LDS loads (variant 2, also
``LdsDma``), else the register-staged tile
(variant 1), which an explicit tile height or plane count selects.
``OnTheFly`` recomputes every neighbour Laplacian from cached loads, like the
reference's OnTheFly schedule (stencils/cuda_hip/horizontal_diffusion.py:
50-96).  All use the oracle's operation order (stencils/base.py:284-307),
bit-exact in FP64.
"""

import numpy as np

from .... import native
from ....benchmark import Parameter, ParameterError
from ..base import HorizontalDiffusionStencil
from .mixin import StencilMixin


class HorizontalDiffusionMixin(StencilMixin):
    block_size = Parameter(
        "tile size: i (64), j (8, 16, 32; LDS-DMA tile 8, 16), k-planes per workgroup "
        "(1, 2, 4, 8, 16; LDS-DMA tile 1); 0: measured default",
        (64, 0, 0),
    )

    VARIANT = 0

    def setup(self):
        super().setup()
        bx, by, bz = self.block_size
        variant = self.variant or self.VARIANT
        if variant in (0, 1) and (bx != 64 or by not in (0, 8, 16, 32) or
                                  bz not in (0, 1, 2, 4, 8, 16)):
            raise ParameterError(f"unsupported tile size {self.block_size}")
        if variant == 2 and (bx != 64 or by not in (0, 8, 16) or bz not in (0, 1)):
            raise ParameterError(f"unsupported LDS-DMA tile size {self.block_size}")

    @property
    def written_fields(self):
        return ("out",)

    def launch_options(self):
        options = super().launch_options()
        options.variant = self.variant or self.VARIANT
        options.bx, options.by, options.bz = self.block_size
        return options

    def launch(self, pointers, options, seconds):
        native.call("sb_hdiff", self.precision, *pointers, self.geom, options, seconds)

    def compulsory_bytes(self):
        nx, ny, nz = self.domain
        return ((nx + 4) * (ny + 4) * nz + 2 * nx * ny * nz) * self.dtype_size


class Fused(HorizontalDiffusionMixin, HorizontalDiffusionStencil):
    VARIANT = 0


class LdsDma(HorizontalDiffusionMixin, HorizontalDiffusionStencil):
    VARIANT = 2


class OnTheFly(HorizontalDiffusionMixin, HorizontalDiffusionStencil):
    VARIANT = 3
