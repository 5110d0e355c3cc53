"""Benchmark collections; importing registers every benchmark class."""

from . import stencils, stream

__all__ = ["stencils", "stream"]
