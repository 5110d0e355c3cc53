"""GPU STREAM on the MI355X (stream/cuda_hip.py:41-144 of the reference).

Three device arrays a, b, c of ``array_size`` elements; each of the
``ntimes`` iterations runs copy (c = a), scale (b = s c), add (c = a + b)
and triad (a = b + s c) with s = 3, each timed with HIP events.  Results
follow the reference's table: per kernel the best rate in MB/s (1e6 B/s,
bytes 2N/2N/3N/3N times the element size) and avg/min/max time, skipping
iteration 0 (stream/cuda_hip.j2:244-276).

``init='constant'`` starts from a = 1, b = 2, c = 0 and checks the closed
form of stream/cuda_hip.j2:329-345 on the device; ``init='random'`` fills
a, b, c with U[0,1) from a counter hash of (seed, index) and checks that
every element follows the recurrence to within the reference's per-type
epsilon.

Launch shape, as the reference's options: ``block_size`` threads per
workgroup, ``vector_size`` elements per lane and load (0: 16 bytes, i.e. 2
float64 or 4 float32; 1 or 2 for float64, 1, 2 or 4 for float32),
``unroll_factor`` vectors per thread, ``axis`` the grid dimension the
threads and workgroups are laid along, non-temporal loads and stores, and
``workgroups_per_cu``: 0 launches one pass over the arrays, k > 0 a
persistent grid of k workgroups per CU that strides over them.
"""

import ctypes
import warnings

import numpy as np

from ... import native
from ...benchmark import Benchmark, ExecutionError, Parameter, ParameterError
from ...device import DeviceBuffer

KERNELS = ("copy", "scale", "add", "triad")


def closed_form(ntimes, a=1.0, b=2.0, c=0.0, scalar=3.0, dtype=np.float64):
    """Values of a, b, c after `ntimes` iterations (elementwise recurrence)."""
    a, b, c, s = (np.asarray(v, dtype=dtype) for v in (a, b, c, scalar))
    for _ in range(ntimes):
        c = a.copy()
        b = s * c
        c = a + b
        a = b + s * c
    return a, b, c


class Native(Benchmark):
    array_size = Parameter("number of elements in arrays", 10000000)
    ntimes = Parameter("number of runs", 10)
    block_size = Parameter("threads per block", 1024)
    dtype = Parameter("data type in NumPy format, e.g. float32 or float64", "float64")
    axis = Parameter("compute grid dimension to use", "x", choices=["x", "y", "z"])
    vector_size = Parameter("vector size in elements (0: 16 bytes)", 0)
    unroll_factor = Parameter("loop unroll factor (vectors per thread)", 1)
    workgroups_per_cu = Parameter("persistent grid: workgroups per CU (0: one pass)", 0)
    streaming_stores = Parameter("use non-temporal store instructions", False)
    streaming_loads = Parameter("use non-temporal load instructions", False)
    init = Parameter("initial values", "constant", choices=["constant", "random"])
    seed = Parameter("seed for init=random", 0)
    device = Parameter("GPU device index", 0)
    verify = Parameter("verify results", True)

    def setup(self):
        super().setup()
        if self.dtype not in ("float32", "float64"):
            raise ParameterError(f"unsupported dtype {self.dtype}")
        if self.ntimes < 2:
            raise ParameterError("ntimes must be at least 2 (iteration 0 is not reported)")
        if self.block_size % 64 or not 64 <= self.block_size <= 1024:
            raise ParameterError("block size must be a multiple of 64 in [64, 1024]")
        if self.unroll_factor not in (1, 2, 4, 8):
            raise ParameterError("unroll factor must be one of 1, 2, 4, 8")
        wide = 16 // np.dtype(self.dtype).itemsize
        vector = self.vector_size or wide
        if vector not in (1, 2, wide):
            raise ParameterError(f"vector size {self.vector_size} not supported for "
                                 f"{self.dtype} (1, 2 or {wide} elements)")
        if not 0 <= self.workgroups_per_cu <= 64:
            raise ParameterError("workgroups per CU must be in 0..64")
        per_block = self.block_size * vector * self.unroll_factor
        if self.array_size % per_block:
            warnings.warn(f"array size rounded up to a multiple of {per_block} elements "
                          "(block x vector x unroll)")
            self.array_size = -(-self.array_size // per_block) * per_block
        try:
            native.init(self.device)
        except ExecutionError as error:
            raise ParameterError(f"GPU device {self.device} unavailable: {error}") from None
        self.precision = native.precision_code(self.dtype)
        nbytes = self.array_size * np.dtype(self.dtype).itemsize
        self.arrays = [DeviceBuffer(nbytes) for _ in range(3)]

    def _initialise(self):
        n = self.array_size
        a, b, c = (ctypes.c_void_p(x.ptr) for x in self.arrays)
        if self.init == "constant":
            for ptr, value in zip((a, b, c), (1.0, 2.0, 0.0)):
                native.call("sb_stream_fill", self.precision, ptr, n, value)
        else:
            for offset, ptr in enumerate((a, b, c)):
                native.call("sb_stream_fill_random", self.precision, ptr, n, self.seed * 3 + offset)

    def launch(self, kernel, options, seconds=None):
        a, b, c = (ctypes.c_void_p(x.ptr) for x in self.arrays)
        native.call(
            "sb_stream_kernel", self.precision, native.STREAM_OPS[kernel], a, b, c,
            self.array_size, 3.0, options, seconds,
        )

    def options(self):
        variant = (1 if self.streaming_loads else 0) | (2 if self.streaming_stores else 0)
        variant |= "xyz".index(self.axis) << 2
        return native.launch_options(
            variant=variant, block=(self.block_size, self.vector_size, self.workgroups_per_cu),
            unroll=self.unroll_factor,
        )

    def run(self):
        self._initialise()
        options = self.options()
        times = {k: [] for k in KERNELS}
        for _ in range(self.ntimes):
            for kernel in KERNELS:
                seconds = ctypes.c_double(0.0)
                self.launch(kernel, options, ctypes.byref(seconds))
                times[kernel].append(seconds.value)
        if self.verify:
            self.check()
        itemsize = np.dtype(self.dtype).itemsize
        words = {"copy": 2, "scale": 2, "add": 3, "triad": 3}
        results = []
        for kernel in KERNELS:
            measured = times[kernel][1:]
            best = min(measured)
            results.append(
                {
                    "name": kernel,
                    "bandwidth": 1e-6 * words[kernel] * self.array_size * itemsize / best,
                    "avg-time": float(np.mean(measured)),
                    "time": best,
                    "max-time": max(measured),
                }
            )
        return results

    def check(self):
        """Closed-form (constant init) or sampled recurrence (random init) check."""
        eps = 1e-6 if self.dtype == "float32" else 1e-13
        n = self.array_size
        if self.init == "constant":
            expected = closed_form(self.ntimes, dtype=np.dtype(self.dtype))
            for name, buffer, value in zip("abc", self.arrays, expected):
                worst = ctypes.c_double(0.0)
                native.call(
                    "sb_stream_check", self.precision, ctypes.c_void_p(buffer.ptr), n,
                    float(value), ctypes.byref(worst),
                )
                if worst.value > eps:
                    raise ExecutionError(
                        f"failed validation on array {name}[]: max rel. error {worst.value:.3e}"
                    )
            return
        # random init: recompute the recurrence on the host for a sample of
        # indices from the same counter hash the device fill used
        from ...tools.hashfill import uniform

        sample = np.unique(np.linspace(0, n - 1, num=min(n, 4096)).astype(np.int64))
        dtype = np.dtype(self.dtype)
        initial = [uniform(self.seed * 3 + o, sample, dtype) for o in range(3)]
        expected = closed_form(self.ntimes, *initial, dtype=dtype)
        itemsize = dtype.itemsize
        for name, buffer, value in zip("abc", self.arrays, expected):
            got = np.empty(len(sample), dtype=dtype)
            for pos, index in enumerate(sample):
                one = np.empty(1, dtype=dtype)
                native.call(
                    "sb_memcpy_d2h", ctypes.c_void_p(one.ctypes.data),
                    ctypes.c_void_p(buffer.ptr + int(index) * itemsize), itemsize,
                )
                got[pos] = one[0]
            if not np.array_equal(got, value):
                bad = int(np.count_nonzero(got != value))
                raise ExecutionError(f"failed validation on array {name}[]: {bad} sampled errors")
