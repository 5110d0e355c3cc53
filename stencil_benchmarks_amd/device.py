"""Device-resident fields with the host field's geometry.

The reference mirrors each host field on the device with the same shape,
strides and alignment and moves it with one linear copy of ``nbytes``
(stencils/cuda_hip/mixin.py:125-160, tools/array.py:180-208).  Here the
device buffer has the same strides, so the copy stays one linear transfer,
but its base is chosen so that the first interior point (the element at
index == halo) sits on a 256-byte boundary: that is free (only the
allocation offset changes) and lets the kernels use aligned dwordx4 loads
whenever the row and plane pitches allow it.
"""

import ctypes
import weakref

import numpy as np

from . import native
from .tools import array as array_tools

DEVICE_ALIGNMENT = 256


class DeviceBuffer:
    """An owning device allocation, freed when garbage collected."""

    def __init__(self, nbytes: int):
        ptr = ctypes.c_void_p()
        native.call("sb_malloc", ctypes.byref(ptr), int(max(nbytes, 1)))
        self.ptr = ptr.value
        self.nbytes = int(nbytes)
        self._finalizer = weakref.finalize(self, _free, self.ptr)

    def free(self):
        self._finalizer()


def _free(ptr):
    try:
        native.lib().sb_free(ctypes.c_void_p(ptr))
    except Exception:  # interpreter shutdown
        pass


def _first_element_offset(host: np.ndarray) -> int:
    """Bytes from the lowest addressed element to element (0, ..., 0)."""
    return int(sum(-(n - 1) * s for n, s in zip(host.shape, host.strides) if s < 0))


class DeviceField:
    """Device mirror of a (possibly padded) host field."""

    def __init__(self, host: np.ndarray, index_to_align=None):
        self.shape = host.shape
        self.dtype = host.dtype
        self.strides = host.strides
        self.span = array_tools.nbytes(host)
        index = tuple(index_to_align) if index_to_align is not None else (0,) * host.ndim
        self.origin = array_tools.origin_offset(host, index)
        pad = (-self.origin) % DEVICE_ALIGNMENT
        self.buffer = DeviceBuffer(self.span + pad + DEVICE_ALIGNMENT)
        base = self.buffer.ptr
        self.base = base + ((-(base + self.origin)) % DEVICE_ALIGNMENT)
        self.index_to_align = index

    def pointer(self, index=None) -> int:
        """Device address of element `index` (default: the aligned index)."""
        index = self.index_to_align if index is None else tuple(index)
        return self.base + sum(s * i for s, i in zip(self.strides, index))

    def upload(self, host: np.ndarray):
        assert host.strides == self.strides and host.dtype == self.dtype
        native.call(
            "sb_memcpy_h2d", ctypes.c_void_p(self.base), ctypes.c_void_p(host.ctypes.data), self.span
        )

    def download(self, host: np.ndarray):
        assert host.strides == self.strides and host.dtype == self.dtype
        native.call(
            "sb_memcpy_d2h", ctypes.c_void_p(host.ctypes.data), ctypes.c_void_p(self.base), self.span
        )

    def fill_from(self, other: "DeviceField"):
        assert other.span == self.span
        native.call("sb_memcpy_d2d", ctypes.c_void_p(self.base), ctypes.c_void_p(other.base), self.span)

    def free(self):
        self.buffer.free()


def geometry(domain, strides_elements, halo) -> native.Geom:
    nx, ny, nz = (int(d) for d in domain)
    si, sj, sk = (int(s) for s in strides_elements)
    hi, hj, hk = (int(h) for h in halo)
    return native.Geom(nx, ny, nz, si, sj, sk, hi, hj, hk)
