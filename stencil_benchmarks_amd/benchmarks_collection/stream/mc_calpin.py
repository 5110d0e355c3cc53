"""Host (CPU) STREAM, McCalpin's benchmark (stream/mc_calpin.py of the
reference, which renders and compiles stream.c at run time: original.j2 and
the vectorised native.j2).

Here the kernels are prebuilt in libsbhost.so (host/stream_host.cpp, OpenMP):
``Original`` runs plain loops vectorised by the compiler, ``Native`` explicit
vector loads with non-temporal stores (AVX-512 where the CPU has it, else
AVX2).  Each of the ``ntimes`` iterations runs copy (c = a), scale (b = 3 c),
add (c = a + b) and triad (a = b + 3 c); results follow McCalpin's table: per
kernel the best rate in MB/s (1e6 B/s, bytes 2N/2N/3N/3N times the element
size) and the average, minimum and maximum time, skipping iteration 0.  The
arrays start at a = 1, b = 2, c = 0 (a doubled before the timed loop) and are
checked against the closed form of the kernel sequence.  Parameters of the
reference's code generation (compiler, flags, platform preset, architecture,
vector size, unroll factor, fma, offset) do not apply to the prebuilt
kernels.
"""

import ctypes
from pathlib import Path

import numpy as np

from ...benchmark import Benchmark, ExecutionError, Parameter, ParameterError

KERNELS = ("copy", "scale", "add", "triad")
WORDS = {"copy": 2, "scale": 2, "add": 3, "triad": 3}
LIBRARY = Path(__file__).resolve().parents[2] / "libsbhost.so"
# McCalpin's validation thresholds (relative error) per element type
EPSILON = {4: 1e-6, 8: 1e-13}

_lib = None


class HostLibraryMissing(RuntimeError):
    """libsbhost.so is not built (make -C stencil_benchmarks_amd/host)."""


def host_lib():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise HostLibraryMissing(f"{LIBRARY} not found: run make -C stencil_benchmarks_amd/host")
        lib = ctypes.CDLL(str(LIBRARY))
        lib.sbh_stream.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        lib.sbh_stream.restype = ctypes.c_int
        lib.sbh_max_threads.restype = ctypes.c_int
        lib.sbh_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


class Original(Benchmark):
    array_size = Parameter("number of elements in arrays", 10000000)
    ntimes = Parameter("number of runs", 10)
    dtype = Parameter("data type in NumPy format, e.g. float32 or float64", "float64")
    threads = Parameter("OpenMP threads (0: the OpenMP default)", 0)

    MODE = 0

    def setup(self):
        super().setup()
        if np.dtype(self.dtype).itemsize not in EPSILON:
            raise ParameterError(f"unsupported dtype {self.dtype}")
        if self.array_size < 1 or self.ntimes < 2:
            raise ParameterError("array_size must be >= 1 and ntimes >= 2")
        host_lib()

    def mode(self):
        return self.MODE

    def run(self):
        itemsize = np.dtype(self.dtype).itemsize
        times = (ctypes.c_double * (4 * self.ntimes))()
        error = ctypes.c_double(0.0)
        isa = ctypes.c_int(0)
        lib = host_lib()
        if lib.sbh_stream(itemsize, self.array_size, self.ntimes, self.threads, self.mode(),
                          times, ctypes.byref(error), ctypes.byref(isa)):
            raise ExecutionError(lib.sbh_last_error().decode())
        if not error.value <= EPSILON[itemsize]:
            raise ExecutionError(f"Failed Validation: max relative error {error.value:.3e} "
                                 f"exceeds {EPSILON[itemsize]:g}")
        results = []
        for k, name in enumerate(KERNELS):
            measured = list(times[k * self.ntimes + 1:(k + 1) * self.ntimes])
            best = min(measured)
            results.append({
                "name": name,
                "bandwidth": 1e-6 * WORDS[name] * self.array_size * itemsize / best,
                "avg-time": float(np.mean(measured)),
                "time": best,
                "max-time": max(measured),
                "vector-bits": isa.value,
                "threads": self.threads or lib.sbh_max_threads(),
            })
        return results


class Native(Original):
    streaming_stores = Parameter("bypass load of destination array", True)

    def mode(self):
        # the explicit kernels always store non-temporally; without
        # streaming stores the compiler's loops are the plain form
        return 1 if self.streaming_stores else 0
