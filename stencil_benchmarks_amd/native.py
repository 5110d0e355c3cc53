"""ctypes binding of libsbmi355x.so, the C ABI of the HIP kernels.

This is the drop-in replacement of the reference's device boundary: the
ctypes ``api.Runtime`` over libamdhip64 (stencils/cuda_hip/api.py:39-104) and
the per-benchmark generated ``kernel()`` libraries loaded through
``compilation.GnuLibrary`` (tools/compilation.py:89-196).  All entry points
return a status code; a non-zero status raises :class:`ExecutionError`
carrying ``sb_last_error()``.  There is no fallback: if the library is not
built, :func:`lib` raises :class:`NativeLibraryMissing`.
"""

import ctypes
import os
import threading
from pathlib import Path

from .benchmark import ExecutionError

LIBRARY_NAME = "libsbmi355x.so"
LIBRARY_PATH = Path(__file__).resolve().parent / LIBRARY_NAME

SB_F64, SB_F32, SB_F32_F64 = 0, 1, 2
STREAM_OPS = {"copy": 0, "scale": 1, "add": 2, "triad": 3}


class NativeLibraryMissing(ImportError):
    """libsbmi355x.so is not built (run __graft_entry__.build())."""


class Geom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("nx", "ny", "nz", "si", "sj", "sk", "hi", "hj", "hk")]


class Launch(ctypes.Structure):
    _fields_ = [
        (n, ctypes.c_int32)
        for n in ("variant", "bx", "by", "bz", "unroll", "dry_runs", "flush", "timer")
    ]


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int32),
        ("wave_size", ctypes.c_int32),
        ("max_threads_per_block", ctypes.c_int32),
        ("xcds", ctypes.c_int32),
        ("lds_per_block", ctypes.c_int64),
        ("lds_per_cu", ctypes.c_int64),
        ("total_mem", ctypes.c_int64),
        ("l2_bytes", ctypes.c_int64),
        ("clock_khz", ctypes.c_int32),
        ("mem_clock_khz", ctypes.c_int32),
        ("mem_bus_width", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_SZ = ctypes.c_size_t
_D = ctypes.c_double
_GP = ctypes.POINTER(Geom)
_LP = ctypes.POINTER(Launch)
_DP = ctypes.POINTER(ctypes.c_double)

# name -> argtypes; every function returns int
SIGNATURES = {
    "sb_abi_version": [],
    "sb_init": [_I],
    "sb_device_count": [ctypes.POINTER(_I)],
    "sb_get_device_info": [_I, ctypes.POINTER(DeviceInfo)],
    "sb_mem_info": [ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)],
    "sb_malloc": [ctypes.POINTER(_P), _SZ],
    "sb_free": [_P],
    "sb_host_malloc": [ctypes.POINTER(_P), _SZ],
    "sb_host_free": [_P],
    "sb_memcpy_h2d": [_P, _P, _SZ],
    "sb_memcpy_d2h": [_P, _P, _SZ],
    "sb_memcpy_d2d": [_P, _P, _SZ],
    "sb_memset": [_P, _I, _SZ],
    "sb_synchronize": [],
    "sb_event_create": [ctypes.POINTER(_P)],
    "sb_event_record": [_P],
    "sb_event_synchronize": [_P],
    "sb_event_destroy": [_P],
    "sb_flush_caches": [],
    "sb_copy": [_I, _P, _P, _GP, _LP, _DP],
    "sb_onesided_average": [_I, _I, _P, _P, _GP, _LP, _DP],
    "sb_symmetric_average": [_I, _I, _P, _P, _GP, _LP, _DP],
    "sb_laplacian": [_I, _I, _I, _I, _P, _P, _GP, _LP, _DP],
    "sb_hdiff": [_I, _P, _P, _P, _GP, _LP, _DP],
    "sb_vadv": [_I, _I, ctypes.POINTER(_P), _GP, _LP, _DP],
    "sb_vadv_noalloc": [_I, _I, ctypes.POINTER(_P), _GP, _LP, _DP],
    "sb_stream_fill": [_I, _P, _SZ, _D],
    "sb_stream_fill_random": [_I, _P, _SZ, ctypes.c_uint64],
    "sb_fill_random_fields": [_I, ctypes.POINTER(_P), _I, _SZ, ctypes.POINTER(ctypes.c_uint64)],
    "sb_stream_kernel": [_I, _I, _P, _P, _P, _SZ, _D, _LP, _DP],
    "sb_stream_check": [_I, _P, _SZ, _D, _DP],
    "sb_batch_step": [_I, _P, _I, _LP, _LP, _DP],
    "sb_batch_plan_create": [_I, _P, _I, _I, _LP, _LP, ctypes.POINTER(_P)],
    "sb_batch_plan_run": [_P, _I, _DP],
    "sb_batch_plan_destroy": [_P],
}

_lock = threading.Lock()
_lib = None
_initialised_devices = set()


def _bind(library):
    library.sb_last_error.restype = ctypes.c_char_p
    library.sb_last_error.argtypes = []
    for name, argtypes in SIGNATURES.items():
        func = getattr(library, name)
        func.argtypes = argtypes
        func.restype = ctypes.c_int
    return library


def load(path=None):
    """Load (without initialising a device) and bind the library."""
    path = Path(path or os.environ.get("SBMI355X_LIBRARY", LIBRARY_PATH))
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} not found: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (or `make -C stencil_benchmarks_amd/csrc`)"
        )
    return _bind(ctypes.CDLL(str(path)))


def lib():
    """The bound library (loaded on first use)."""
    global _lib
    with _lock:
        if _lib is None:
            _lib = load()
        return _lib


def last_error() -> str:
    msg = lib().sb_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(status: int, what: str = ""):
    if status != 0:
        detail = last_error()
        raise ExecutionError(f"{what}: {detail}" if what else detail)


def call(name: str, *args):
    """Call entry point `name`, raising ExecutionError on a non-zero status."""
    check(getattr(lib(), name)(*args), name)


def init(device: int = 0):
    """Initialise `device` once per process (stream, events)."""
    if device not in _initialised_devices:
        call("sb_init", device)
        _initialised_devices.add(device)
    else:
        call("sb_init", device)


def device_count() -> int:
    count = ctypes.c_int(0)
    call("sb_device_count", ctypes.byref(count))
    return count.value


def device_info(device: int = 0) -> dict:
    info = DeviceInfo()
    call("sb_get_device_info", device, ctypes.byref(info))
    out = {name: getattr(info, name) for name, _ in DeviceInfo._fields_ if name != "reserved"}
    out["name"] = info.name.decode(errors="replace")
    out["arch"] = info.arch.decode(errors="replace")
    return out


def mem_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    call("sb_mem_info", ctypes.byref(free), ctypes.byref(total))
    return free.value, total.value


class Event:
    """A marker on the library stream: record() behind the work submitted so
    far, wait() until that work has completed."""

    def __init__(self):
        handle = ctypes.c_void_p()
        call("sb_event_create", ctypes.byref(handle))
        self.handle = handle

    def record(self):
        call("sb_event_record", self.handle)

    def wait(self):
        call("sb_event_synchronize", self.handle)

    def close(self):
        if self.handle:
            call("sb_event_destroy", self.handle)
            self.handle = None


SB_TIMER_EVENTS, SB_TIMER_HIP_EXT = 0, 1


def launch_options(variant=0, block=(0, 0, 0), unroll=0, dry_runs=0, flush=False,
                   timer=SB_TIMER_EVENTS) -> Launch:
    bx, by, bz = (tuple(block) + (0, 0, 0))[:3]
    return Launch(int(variant), int(bx), int(by), int(bz), int(unroll), int(dry_runs),
                  int(bool(flush)), int(timer))


def precision_code(dtype, compute=None) -> int:
    """C-ABI precision code for storage `dtype` and arithmetic `compute`."""
    import numpy as np

    dtype = np.dtype(dtype)
    compute = dtype if compute is None else np.dtype(compute)
    if dtype == np.float64 and compute == np.float64:
        return SB_F64
    if dtype == np.float32 and compute == np.float32:
        return SB_F32
    if dtype == np.float32 and compute == np.float64:
        return SB_F32_F64
    raise ValueError(f"unsupported storage/compute combination {dtype}/{compute}")
