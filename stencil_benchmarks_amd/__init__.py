"""MI355X-native stencil and STREAM benchmark backend.

Drop-in for the ``stencils`` and ``stream`` benchmark collections of
GridTools/stencil_benchmarks (Benchmark/Parameter API, ``sbench`` CLI,
numpy verification), with the device work done by hand-written gfx950 HIP
kernels behind the C-ABI library ``libsbmi355x.so``.
"""

__version__ = "0.1.0"
