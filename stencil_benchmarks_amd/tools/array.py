"""Padded, aligned field allocation with a layout permutation.

Geometry contract of stencil_benchmarks/tools/array.py:53-208:

* ``layout`` is a permutation of ``0..ndim-1``; the dimension holding the
  value ``ndim - 1`` gets unit stride, the one holding ``0`` the largest
  (``array.py:137-145``);
* only the innermost extent is padded up to a multiple of ``alignment``
  bytes, so every stride except the unit one is a multiple of it;
* the element at ``index_to_align`` (the first interior point when a
  stencil passes its halo) sits on an ``alignment``-byte boundary
  (``array.py:162-176``);
* ``apply_offset`` staggers successive allocations by multiples of the L1
  line size across the L1 sets (``array.py:147-160``).

>>> x = alloc_array((2, 3), dtype='int32', layout=(1, 0))
>>> x.strides
(4, 8)
>>> alloc_array((2, 3), dtype='int32', layout=(0, 1)).strides
(12, 4)
>>> y = alloc_array((2, 3), dtype='int32', layout=(0, 1), alignment=64)
>>> y.strides, y.ctypes.data % 64
((64, 4), 0)
>>> nbytes(y)
76
"""

import mmap
import warnings
from pathlib import Path
from typing import Any, Callable, Optional, Tuple, Union

import numpy as np

_PAGE = mmap.PAGESIZE
_MAP_HUGETLB = 0x40000
_stagger_count = 0


def _buffer_address(buffer) -> int:
    return np.frombuffer(buffer, dtype=np.uint8).ctypes.data


def alloc_smallpages(nbytes: int):
    """Page-aligned host buffer on normal pages."""
    raw = np.empty(max(int(nbytes), 1) + _PAGE, dtype=np.uint8)
    start = (-raw.ctypes.data) % _PAGE
    return raw[start : start + max(int(nbytes), 1)]


def alloc_hugepages(nbytes: int, transparent: bool = True):
    """Huge-page backed host buffer (THP advice, or MAP_HUGETLB)."""
    size = max(int(nbytes), 1)
    if transparent:
        region = mmap.mmap(-1, size, flags=mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS)
        if hasattr(region, "madvise") and hasattr(mmap, "MADV_HUGEPAGE"):
            region.madvise(mmap.MADV_HUGEPAGE)
        return region
    huge = 2 << 20
    size = (size + huge - 1) // huge * huge
    try:
        return mmap.mmap(
            -1, size, flags=mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS | _MAP_HUGETLB
        )
    except OSError as error:
        raise RuntimeError(f"explicit huge page allocation failed: {error}") from None


def _read_cache_value(name: str) -> int:
    path = Path("/sys/devices/system/cpu/cpu0/cache/index0") / name
    try:
        return int(path.read_text().strip())
    except (OSError, ValueError):
        return 0


def l1_dcache_linesize() -> int:
    return _read_cache_value("coherency_line_size")


def l1_dcache_sets() -> int:
    return _read_cache_value("number_of_sets")


def _floor_log2(value: int) -> int:
    return max(int(value), 1).bit_length() - 1


def compute_strides(shape, itemsize, layout, alignment) -> Tuple[int, ...]:
    """Byte strides for `shape` under `layout`, innermost extent padded."""
    ndim = len(shape)
    strides = [0] * ndim
    running = itemsize
    for rank in reversed(range(ndim)):
        axis = layout.index(rank)
        strides[axis] = running
        running *= shape[axis]
        if rank == ndim - 1 and alignment:
            running = -(-running // alignment) * alignment
    return tuple(strides), running


def alloc_array(
    shape: Tuple[int, ...],
    dtype: Union[str, np.dtype],
    layout: Tuple[int, ...],
    alignment: int = 0,
    index_to_align: Optional[Tuple[int, ...]] = None,
    alloc: Optional[Callable[[int], Any]] = None,
    apply_offset: bool = False,
) -> np.ndarray:
    """Allocate an uninitialised padded array; see the module docstring."""
    global _stagger_count
    shape = tuple(int(s) for s in shape)
    ndim = len(shape)
    dtype = np.dtype(dtype)
    layout = tuple(layout)
    alignment = int(alignment)
    index_to_align = (0,) * ndim if index_to_align is None else tuple(index_to_align)
    alloc = alloc_smallpages if alloc is None else alloc

    if tuple(sorted(layout)) != tuple(range(ndim)):
        raise ValueError(f"layout {layout} is not a permutation of the {ndim} axes")
    if alignment < 0:
        raise ValueError(f"alignment of {alignment} bytes is negative")
    if not ndim == len(layout) == len(index_to_align):
        raise ValueError(f"shape, layout and index_to_align differ in length "
                         f"({ndim}, {len(layout)}, {len(index_to_align)})")

    strides, span = compute_strides(shape, dtype.itemsize, layout, alignment)

    offset = 0
    if apply_offset:
        sets = l1_dcache_sets()
        if not sets:
            warnings.warn("L1 data cache set count unknown on this host; staggering "
                          "allocations over 64 sets")
            sets = 64
        line = l1_dcache_linesize()
        if not line:
            warnings.warn("L1 data cache line size unknown on this host; staggering "
                          "allocations by 64 bytes")
            line = 64
        offset = (_stagger_count % sets) << _floor_log2(max(line, alignment))
        _stagger_count += 1

    buffer = alloc(span + alignment + offset)
    if alignment:
        origin = sum(s * i for s, i in zip(strides, index_to_align))
        target = _buffer_address(buffer) + origin + offset
        offset += (-target) % alignment
    return np.ndarray(
        shape=shape, dtype=dtype, buffer=buffer, offset=offset, strides=strides
    )


def nbytes(data: np.ndarray) -> int:
    """Bytes covered by the array including padding (1-D copy length).

    >>> nbytes(np.zeros((2, 3), dtype='int32'))
    24
    """
    last = sum((n - 1) * s for n, s in zip(data.shape, data.strides))
    return int(last + data.itemsize)


def element_strides(data: np.ndarray) -> Tuple[int, ...]:
    """Strides in elements."""
    return tuple(s // data.itemsize for s in data.strides)


def origin_offset(data: np.ndarray, index) -> int:
    """Byte offset of element `index` from the array's first element."""
    return int(sum(s * i for s, i in zip(data.strides, index)))
