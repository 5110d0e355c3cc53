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

import bisect
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


class DeviceArena:
    """One device allocation carved into blocks: first fit over a free list
    sorted by offset, neighbours coalesced on release.

    For callers that admit and retire many fields (the request server):
    taking a block is host bookkeeping only, with no ``hipMalloc`` and no
    ``hipFree`` (which waits for the device) per field.  Blocks are
    `alignment`-aligned.  A block must not be released while kernels that
    use it are in flight.
    """

    def __init__(self, nbytes: int, alignment: int = DEVICE_ALIGNMENT, buffer=None):
        self.alignment = int(alignment)
        self.capacity = int(nbytes) // self.alignment * self.alignment
        # `buffer`: an existing allocation of at least nbytes (ptr, free())
        self.buffer = DeviceBuffer(self.capacity) if buffer is None else buffer
        self._free = [(0, self.capacity)]  # (offset, size), sorted by offset
        self.in_use = 0

    def alloc(self, nbytes: int) -> "ArenaBlock":
        size = max(self.alignment, -(-int(nbytes) // self.alignment) * self.alignment)
        for index, (offset, free) in enumerate(self._free):
            if free >= size:
                if free == size:
                    del self._free[index]
                else:
                    self._free[index] = (offset + size, free - size)
                self.in_use += size
                return ArenaBlock(self, offset, size)
        raise MemoryError(f"device arena: no free block of {size} bytes "
                          f"({self.capacity - self.in_use} of {self.capacity} free, fragmented)")

    def _release(self, offset: int, size: int):
        self.in_use -= size
        index = bisect.bisect_left(self._free, (offset, 0))
        if index < len(self._free) and self._free[index][0] == offset + size:  # merge right
            size += self._free[index][1]
            del self._free[index]
        if index > 0 and sum(self._free[index - 1]) == offset:  # merge left
            offset, size = self._free[index - 1][0], self._free[index - 1][1] + size
            del self._free[index - 1]
            index -= 1
        self._free.insert(index, (offset, size))

    def close(self):
        self.buffer.free()


class ArenaBlock:
    """A block of a DeviceArena (the DeviceBuffer interface: ptr, nbytes, free)."""

    def __init__(self, arena: DeviceArena, offset: int, size: int):
        self.ptr = arena.buffer.ptr + offset
        self.nbytes = size
        self._finalizer = weakref.finalize(self, arena._release, offset, size)

    def free(self):
        self._finalizer()


def _first_element_offset(host: np.ndarray) -> int:
    """Bytes from the lowest addressed element to element (0, ..., 0)."""
    return int(sum(-(n - 1) * s for n, s in zip(host.shape, host.strides) if s < 0))


class DeviceField:
    """Device mirror of a (possibly padded) host field."""

    def __init__(self, host: np.ndarray, index_to_align=None, arena: DeviceArena = None):
        self.shape = host.shape
        self.dtype = host.dtype
        self.strides = host.strides
        self.span = array_tools.nbytes(host)
        index = tuple(index_to_align) if index_to_align is not None else (0,) * host.ndim
        self.origin = array_tools.origin_offset(host, index)
        pad = (-self.origin) % DEVICE_ALIGNMENT
        nbytes = self.span + pad + DEVICE_ALIGNMENT
        self.buffer = arena.alloc(nbytes) if arena is not None else DeviceBuffer(nbytes)
        base = self.buffer.ptr
        self.base = base + ((-(base + self.origin)) % DEVICE_ALIGNMENT)
        self.index_to_align = index

    @classmethod
    def group(cls, shape, dtype, strides, count, index_to_align=None, arena=None):
        """`count` fields of one geometry (`strides` in bytes, non-negative)
        carved from one allocation (a block of `arena` if given): the
        geometry is computed once, and each field's slot keeps the same
        256-byte alignment of `index_to_align`."""
        dtype = np.dtype(dtype)
        shape, strides = tuple(shape), tuple(strides)
        index = tuple(index_to_align) if index_to_align is not None else (0,) * len(shape)
        span = int(sum((n - 1) * st for n, st in zip(shape, strides)) + dtype.itemsize)
        origin = int(sum(st * i for st, i in zip(strides, index)))
        slot = -(-(span + DEVICE_ALIGNMENT) // DEVICE_ALIGNMENT) * DEVICE_ALIGNMENT
        buffer = arena.alloc(count * slot) if arena is not None else DeviceBuffer(count * slot)
        fields = []
        for q in range(count):
            field = cls.__new__(cls)
            field.shape, field.dtype, field.strides = shape, dtype, strides
            field.span, field.origin, field.index_to_align = span, origin, index
            field.buffer = buffer  # shared; freeing it once frees every field
            base = buffer.ptr + q * slot
            field.base = base + ((-(base + origin)) % DEVICE_ALIGNMENT)
            fields.append(field)
        return fields

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
