"""Common driver of the MI355X stencil benchmarks.

Plays the role of the reference's cuda_hip ``StencilMixin``
(stencils/cuda_hip/mixin.py:50-175) without run-time code generation: the
kernels are prebuilt in libsbmi355x.so and take the geometry at run time.

* ``setup`` initialises the device, allocates a device mirror of every
  field of every data set (same strides as the host field, interior origin
  on a 256-byte boundary) and uploads it once; fields stay resident in HBM.
* ``run_stencil`` launches the kernel through the C ABI, timed with HIP
  events on the library stream (``timers='gpu'``, the reference default),
  with events the kernel launches record themselves (``timers='hip-ext'``,
  hipExtLaunchKernelGGL) or with a host clock around launch and
  synchronisation (``timers='wall'``),
  after ``dry_runs`` untimed launches and, with ``flush_caches``, after
  streaming 1 GiB through L2 and the Infinity Cache; then it downloads the
  fields the kernel wrote so that ``verify_stencil`` sees them.
* ``compute`` selects the arithmetic type for float32 storage: ``float64``
  (default; all arithmetic in FP64 registers, results rounded once, and the
  verification oracle runs in FP64 on the stored inputs) or ``storage``
  (FP32 arithmetic in the oracle's operation order).
"""

import abc
import ctypes
import time
import warnings

import numpy as np

from .... import native
from ....benchmark import Benchmark, ExecutionError, Parameter, ParameterError
from ....device import DeviceField, geometry


class StencilMixin(Benchmark):
    device = Parameter("GPU device index", 0)
    dry_runs = Parameter("kernel dry-runs before the measurement", 0)
    timers = Parameter("timer type", default="gpu", choices=["gpu", "wall", "hip-ext"])
    flush_caches = Parameter(
        "evict L2 and Infinity Cache before each timed launch (cold-cache timing)", True
    )
    compute = Parameter(
        "arithmetic type for float32 storage", "float64", choices=["float64", "storage"]
    )
    variant = Parameter("kernel variant (0: default)", 0)

    def setup(self):
        super().setup()
        if self.dtype not in ("float32", "float64"):
            raise ParameterError(f"unsupported dtype {self.dtype}")
        try:
            native.init(self.device)
        except native.NativeLibraryMissing:
            raise
        except ExecutionError as error:
            raise ParameterError(f"GPU device {self.device} unavailable: {error}") from None
        self.oracle_dtype = (
            np.float64 if (self.dtype == "float32" and self.compute == "float64") else None
        )
        self.precision = native.precision_code(
            self.dtype, np.float64 if self.compute == "float64" else self.dtype
        )
        self._device_data = [
            [DeviceField(field, self.halo) for field in data_set] for data_set in self._data
        ]
        for host_set, device_set in zip(self._data, self._device_data):
            for host, dev in zip(host_set, device_set):
                dev.upload(host)
        self.geom = geometry(self.domain, self.strides, self.halo)
        if self.verify and self.dry_runs and self.written_fields_are_read:
            warnings.warn(
                "using --dry-runs together with verification might lead to "
                "false negatives for stencils with read-write fields"
            )

    written_fields_are_read = False

    @property
    @abc.abstractmethod
    def written_fields(self):
        """Names of the arguments the kernel writes."""

    @abc.abstractmethod
    def launch(self, pointers, options, seconds):
        """Call the C-ABI entry point with the device `pointers` (args order)."""

    def launch_options(self):
        timer = native.SB_TIMER_HIP_EXT if self.timers == "hip-ext" else native.SB_TIMER_EVENTS
        return native.launch_options(variant=self.variant, dry_runs=self.dry_runs,
                                     flush=self.flush_caches, timer=timer)

    def run_stencil(self, data):
        device_set = self._device_data[self.data_set_index]
        pointers = [ctypes.c_void_p(f.pointer()) for f in device_set]
        options = self.launch_options()
        if self.timers in ("gpu", "hip-ext"):
            seconds = ctypes.c_double(0.0)
            self.launch(pointers, options, ctypes.byref(seconds))
            elapsed = seconds.value
        else:
            # the same launch shape, without the library's own dry runs and flush
            plain = native.Launch.from_buffer_copy(options)
            plain.dry_runs, plain.flush = 0, 0
            for _ in range(options.dry_runs):
                self.launch(pointers, plain, None)
            if options.flush:
                native.call("sb_flush_caches")
            native.call("sb_synchronize")
            start = time.perf_counter()
            self.launch(pointers, plain, None)
            native.call("sb_synchronize")
            elapsed = time.perf_counter() - start
        for name in self.written_fields:
            position = self.args.index(name)
            device_set[position].download(data[position])
        if elapsed <= 0:
            elapsed = 1e-9
        return dict(time=elapsed, **self.extra_results(elapsed))

    def compulsory_bytes(self):
        return self.data_size

    def extra_results(self, elapsed):
        points = int(np.prod(self.domain))
        return {
            "compulsory-bandwidth": self.compulsory_bytes() / elapsed / 1e9,
            "gridpoints-per-s": points / elapsed,
        }
