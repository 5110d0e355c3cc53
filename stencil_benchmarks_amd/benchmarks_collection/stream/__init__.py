"""STREAM benchmarks: the MI355X GPU STREAM and the host (CPU) STREAM."""

from . import mc_calpin, mi355x

__all__ = ["mc_calpin", "mi355x"]
