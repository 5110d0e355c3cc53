"""Device arena bookkeeping (first fit, coalescing), grouped field geometry
and the cached subdomain structs - CPU only.

The arena is exercised over a stand-in buffer (no device memory): blocks
never overlap, stay inside the arena and aligned, the in-use count matches
the live blocks, a released arena coalesces back into one free block, and
an unsatisfiable request raises MemoryError.
"""

import random

import pytest

from stencil_benchmarks_amd.device import DeviceArena


class _Buffer:
    def __init__(self, base):
        self.ptr = base
        self.freed = False

    def free(self):
        self.freed = True


def _arena(capacity, alignment=256, base=1 << 20):
    return DeviceArena(capacity, alignment, buffer=_Buffer(base))


def _check(arena, live):
    spans = sorted((b.ptr, b.ptr + b.nbytes) for b in live)
    base = arena.buffer.ptr
    for lo, hi in spans:
        assert base <= lo < hi <= base + arena.capacity
        assert (lo - base) % arena.alignment == 0
    for (_, hi), (lo, _) in zip(spans, spans[1:]):
        assert hi <= lo, "blocks overlap"
    assert arena.in_use == sum(b.nbytes for b in live)
    free = arena._free
    assert free == sorted(free)
    for (o1, s1), (o2, _) in zip(free, free[1:]):
        assert o1 + s1 < o2, "adjacent free blocks were not merged"


def test_random_alloc_free_keeps_blocks_disjoint_and_coalesces():
    rng = random.Random(5)
    arena = _arena(1 << 24)
    live = []
    for _ in range(3000):
        if live and rng.random() < 0.45:
            live.pop(rng.randrange(len(live))).free()
        else:
            try:
                live.append(arena.alloc(rng.choice([1, 255, 256, 257, 4096, 100000, 1 << 20])))
            except MemoryError:
                pass
        _check(arena, live)
    for block in live:
        block.free()
    assert arena.in_use == 0 and arena._free == [(0, arena.capacity)]


def test_first_fit_reuses_the_lowest_hole_and_sizes_round_up():
    arena = _arena(4096)
    a, b, c = arena.alloc(100), arena.alloc(300), arena.alloc(256)
    assert (a.nbytes, b.nbytes, c.nbytes) == (256, 512, 256)
    assert b.ptr == a.ptr + 256 and c.ptr == b.ptr + 512
    b.free()
    d = arena.alloc(200)
    assert d.ptr == b.ptr  # first hole
    b.free()  # releasing twice is a no-op
    assert arena.in_use == 256 * 3


def test_exhausted_or_fragmented_arena_raises():
    arena = _arena(1024)
    blocks = [arena.alloc(256) for _ in range(4)]
    with pytest.raises(MemoryError):
        arena.alloc(1)
    blocks[0].free()
    blocks[2].free()  # 512 bytes free, but not contiguous
    with pytest.raises(MemoryError):
        arena.alloc(512)
    blocks[1].free()
    assert arena.alloc(768).ptr == arena.buffer.ptr


def test_close_frees_the_underlying_allocation():
    arena = _arena(1024)
    arena.close()
    assert arena.buffer.freed


def test_grouped_fields_share_geometry_and_do_not_overlap():
    import numpy as np

    from stencil_benchmarks_amd.device import DeviceField
    from stencil_benchmarks_amd.tools import array

    arena = _arena(1 << 34, base=(1 << 40) + 64)  # base deliberately not 256-aligned
    for domain in [(70, 37, 9), (1, 1, 2), (512, 100, 3)]:
        halo = (3, 3, 3)
        shape = tuple(d + 2 * h for d, h in zip(domain, halo))
        template = array.alloc_array(shape, np.float32, (2, 1, 0), 64, halo)
        single = DeviceField(template, halo, arena=arena)
        group = DeviceField.group(shape, np.float32, template.strides, 7, halo, arena)
        spans = sorted((f.base, f.base + f.span) for f in group)
        for f in group:
            assert (f.strides, f.span, f.origin) == (single.strides, single.span, single.origin)
            assert f.pointer() % 256 == 0 and f.pointer() == f.base + f.origin
        assert all(a[1] <= b[0] for a, b in zip(spans, spans[1:]))
        lo, hi = group[0].buffer.ptr, group[0].buffer.ptr + group[0].buffer.nbytes
        assert all(lo <= f.base and f.base + f.span <= hi for f in group)
        in_use = arena.in_use
        for f in group:
            f.free()  # one shared block, released once
        assert arena.in_use == in_use - group[0].buffer.nbytes
        single.free()


def test_subdomain_struct_cache_matches_fresh_structs():
    import ctypes

    from stencil_benchmarks_amd import batch

    arena = _arena(1 << 34)
    subs = [batch.Subdomain(d, (3, 3, 3), "float32", fill=False, arena=arena)
            for d in [(64, 16, 40), (33, 70, 97), (512, 3, 2)]]
    subs[1].advance(1)
    array = batch._array(subs)
    assert len(array) == 3
    for i, sub in enumerate(subs):
        assert bytes(array[i]) == bytes(sub.struct()) and array[i].parity == sub.parity
    subs[1].advance(1)
    assert batch._array(subs)[1].parity == 0
    assert bytes(batch._array([])[0]) == bytes(ctypes.sizeof(batch.SubdomainStruct))
    nx, ny, nz = subs[0].domain
    assert subs[0].points == nx * ny * nz
    assert subs[0].compulsory_bytes_per_step() == (
        ((nx + 4) * (ny + 4) * nz + 2 * nx * ny * nz) + (5 * nx * ny * nz + (nx + 1) * ny * nz)) * 4
