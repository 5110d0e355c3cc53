"""ORACLE - test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline``
leg import this module, and only as the checker (or, for the CPU baseline,
as the thing timed on the host).  The product (stencil_benchmarks_amd)
never imports, loads or calls anything under oracle/.

Python front end of liboracle.so (oracle.c): the reference's verification
oracles (stencil_benchmarks/benchmarks_collection/stencils/base.py) and
the GPU STREAM recurrence (benchmarks_collection/stream/cuda_hip.j2)
restated in C, operation by operation, so float64 results equal the
reference oracle's bit for bit.  Each function takes full halo-inclusive
numpy arrays (any strides) plus the halo width and writes or returns the
interior result.

Pinning (see tests/test_oracle.py): every function is checked against
closed forms, hand-computed known answers or an independent textbook
computation (dense linear solve) chosen so that a dropped term, a wrong
sign, index or operand fails; there are no golden vectors in the
reference for these paths (its tests recompute the oracle on random
inputs every run).
"""

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None
_cpu = None

_idx = ctypes.c_ssize_t


def _load(name):
    path = HERE / name
    if not path.exists():
        raise ImportError(f"{path} missing: run `make -C oracle`")
    return ctypes.CDLL(str(path))


def lib():
    global _lib
    if _lib is None:
        _lib = _load("liboracle.so")
    return _lib


def cpu_lib():
    global _cpu
    if _cpu is None:
        _cpu = _load("libcpu_baseline.so")
        for name in ("cpu_hdiff_f64", "cpu_hdiff_f32", "cpu_vadv_f64", "cpu_vadv_f32",
                     "cpu_laplacian_f64", "cpu_copy_f64"):
            getattr(_cpu, name).restype = ctypes.c_double
    return _cpu


def _suffix(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise TypeError(f"oracle supports float32/float64, not {dtype}")


def _origin(a, halo):
    return ctypes.c_void_p(a.ctypes.data + sum(s * h for s, h in zip(a.strides, halo)))


def _geom(a, domain):
    strides = [s // a.itemsize for s in a.strides]
    return [_idx(int(d)) for d in domain] + [_idx(int(s)) for s in strides]


def _same_geometry(*arrays):
    first = arrays[0]
    for a in arrays[1:]:
        if a.strides != first.strides or a.dtype != first.dtype or a.shape != first.shape:
            raise ValueError("oracle fields must share shape, strides and dtype")


def _clone(a, copy_values=False):
    """New array with exactly the strides of `a` (padding included)."""
    span = sum((n - 1) * st for n, st in zip(a.shape, a.strides)) + a.itemsize
    buffer = np.zeros(span, dtype=np.uint8)
    out = np.ndarray(a.shape, dtype=a.dtype, buffer=buffer, strides=a.strides)
    if copy_values:
        out[...] = a
    return out


def _interior(a, domain, halo):
    return a[tuple(slice(h, h + d) for d, h in zip(domain, halo))]


def copy(inp, domain, halo):
    out = _clone(inp)
    getattr(lib(), f"oracle_copy_{_suffix(inp.dtype)}")(
        _origin(inp, halo), _origin(out, halo), *_geom(inp, domain))
    return _interior(out, domain, halo).copy()


def onesided_average(inp, axis, domain, halo):
    out = _clone(inp)
    getattr(lib(), f"oracle_onesided_average_{_suffix(inp.dtype)}")(
        _origin(inp, halo), _origin(out, halo), ctypes.c_int(axis), *_geom(inp, domain))
    return _interior(out, domain, halo).copy()


def symmetric_average(inp, axis, domain, halo):
    out = _clone(inp)
    getattr(lib(), f"oracle_symmetric_average_{_suffix(inp.dtype)}")(
        _origin(inp, halo), _origin(out, halo), ctypes.c_int(axis), *_geom(inp, domain))
    return _interior(out, domain, halo).copy()


def laplacian(inp, along, domain, halo):
    out = _clone(inp)
    getattr(lib(), f"oracle_laplacian_{_suffix(inp.dtype)}")(
        _origin(inp, halo), _origin(out, halo), *[ctypes.c_int(int(a)) for a in along],
        *_geom(inp, domain))
    return _interior(out, domain, halo).copy()


def hdiff(inp, coeff, domain, halo):
    _same_geometry(inp, coeff)
    out = _clone(inp)
    getattr(lib(), f"oracle_hdiff_{_suffix(inp.dtype)}")(
        _origin(inp, halo), _origin(coeff, halo), _origin(out, halo), *_geom(inp, domain))
    return _interior(out, domain, halo).copy()


COMPONENT_SHIFTS = {"u": (1, 0), "v": (0, 1), "w": (0, 0)}


def vadv(stage, pos, tens, tensstage, wcon, domain, halo, shift=(1, 0), arithmetic=None):
    """Thomas solve of one component; returns the new tensstage interior.

    `arithmetic` None: the fields' own type; ``'longdouble'``: x87 extended
    precision on float64 fields (error bars only).
    """
    _same_geometry(stage, pos, tens, tensstage, wcon)
    result = _clone(tensstage, copy_values=True)
    if arithmetic == "longdouble":
        if stage.dtype != np.float64:
            raise TypeError("long double oracle takes float64 fields")
        name = "oracle_vadv_ld"
    else:
        name = f"oracle_vadv_{_suffix(stage.dtype)}"
    status = getattr(lib(), name)(
        _origin(stage, halo), _origin(pos, halo), _origin(tens, halo), _origin(result, halo),
        _origin(wcon, halo), ctypes.c_int(shift[0]), ctypes.c_int(shift[1]),
        *_geom(stage, domain))
    if status:
        raise ValueError("oracle vadv failed (nz < 2 or out of memory)")
    return _interior(result, domain, halo).copy()


def stream(a, b, c, ntimes, scalar=3.0):
    """Apply `ntimes` STREAM iterations elementwise; returns new (a, b, c)."""
    a, b, c = (np.ascontiguousarray(x).copy() for x in (a, b, c))
    getattr(lib(), f"oracle_stream_{_suffix(a.dtype)}")(
        ctypes.c_void_p(a.ctypes.data), ctypes.c_void_p(b.ctypes.data),
        ctypes.c_void_p(c.ctypes.data), _idx(a.size), ctypes.c_int(ntimes),
        ctypes.c_double(scalar))
    return a, b, c


def upcast(*arrays):
    """float64 copies keeping logical shape (for FP64 arithmetic on FP32 storage)."""
    return [np.asarray(a, dtype=np.float64).copy() for a in arrays]
