"""numpy CPU backend (stencils/numpy of the reference)."""

from . import basic, horizontal_diffusion

__all__ = ["basic", "horizontal_diffusion"]
