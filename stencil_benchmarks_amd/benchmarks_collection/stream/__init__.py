"""STREAM benchmarks: the MI355X GPU STREAM."""

from . import mi355x

__all__ = ["mi355x"]
