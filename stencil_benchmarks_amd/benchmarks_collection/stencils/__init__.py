"""Structured-grid stencil benchmarks: numpy (CPU) and mi355x (GPU) backends."""

from . import mi355x, numpy

__all__ = ["mi355x", "numpy"]
