"""MI355X (gfx950) backend: hand-written HIP kernels behind libsbmi355x.so."""

from . import basic, horizontal_diffusion, vertical_advection

__all__ = ["basic", "horizontal_diffusion", "vertical_advection"]
