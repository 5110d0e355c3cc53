"""Host tools: array geometry, validation, fills, ranges, timing."""

import numpy as np
import pytest

from stencil_benchmarks_amd.tools import array, cli, fill, hashfill, timing, validation


def test_alloc_basic():
    data = array.alloc_array((3, 5, 7), "int32", (0, 2, 1))
    assert data.shape == (3, 5, 7) and data.dtype == np.dtype("int32")
    assert data.strides[1] < data.strides[2] < data.strides[0]
    i, j, k = np.meshgrid(range(3), range(5), range(7), indexing="ij")
    data[...] = i + 2 * j + 6 * k
    assert np.array_equal(data, i + 2 * j + 6 * k)


def test_alloc_alignment():
    data = array.alloc_array((3, 5, 7), "int32", (0, 2, 1), alignment=64)
    assert data.ctypes.data % 64 == 0 and data.strides[2] % 64 == 0
    assert data.strides[1] < data.strides[2] < data.strides[0]


def test_index_to_align():
    data = array.alloc_array((3, 5, 7), "int32", (0, 2, 1), alignment=64, index_to_align=(1, 1, 2))
    assert data[1:, 1:, 2:].ctypes.data % 64 == 0 and data.strides[2] % 64 == 0


def test_only_innermost_padded():
    # 518 doubles = 4144 B, padded to 4160 B (65 x 64 B); planes not re-padded
    data = array.alloc_array((518, 518, 166), "float64", (2, 1, 0), alignment=64)
    assert data.strides == (8, 4160, 4160 * 518)
    assert array.nbytes(data) == 4160 * 518 * 165 + 4160 * 517 + 517 * 8 + 8


def test_layout_and_errors():
    assert array.alloc_array((2, 3, 4), "float32", (2, 1, 0)).strides == (4, 8, 24)
    assert array.alloc_array((2, 3, 4), "float32", (0, 1, 2)).strides == (48, 16, 4)
    with pytest.raises(ValueError):
        array.alloc_array((2, 3), "float32", (0, 0))
    with pytest.raises(ValueError):
        array.alloc_array((2, 3), "float32", (0, 1), alignment=-1)


def test_offset_allocations_keep_alignment():
    for _ in range(3):
        data = array.alloc_array((9, 9, 9), "float64", (2, 1, 0), 64, (3, 3, 3), apply_offset=True)
        assert data[3:, 3:, 3:].ctypes.data % 64 == 0


def test_huge_pages_transparent():
    data = array.alloc_array((64, 64, 8), "float64", (2, 1, 0), 64, (1, 1, 1),
                             alloc=lambda n: array.alloc_hugepages(n, True))
    data[...] = 1.5
    assert data.sum() == 1.5 * data.size


def test_validation_tolerances():
    a = np.ones((4, 4, 4))
    validation.check_equality("a", a, a * (1 + 1e-6))
    with pytest.raises(validation.ValidationError, match=r"a: 1 of 64 points differ"):
        b = a.copy()
        b[1, 2, 3] = 2
        validation.check_equality("a", b, a)
    f = np.ones(5, dtype=np.float32)
    validation.check_equality("f", f, f * np.float32(1 + 5e-5))
    with pytest.raises(validation.ValidationError):
        validation.check_close("exact", a, np.nextafter(a, 2), rtol=0)
    stats = validation.error_stats([1.0, 2.0, 0.0], [1.0, 2.002, 1e-9])
    assert stats["max_abs"] == pytest.approx(0.002)
    assert stats["pointwise_rel"] == pytest.approx(0.002 / 2.002)


def test_fill_is_seeded_and_layout_independent():
    a = array.alloc_array((6, 5, 4), "float64", (2, 1, 0), 64, (1, 1, 1))
    b = array.alloc_array((6, 5, 4), "float64", (0, 1, 2))
    fill.random_fill(a, fill.field_rng(3, 0, 1))
    fill.random_fill(b, fill.field_rng(3, 0, 1))
    assert np.array_equal(a, b) and a.min() >= 0 and a.max() < 1
    c = np.empty((100,), np.float32)
    fill.random_fill(c, fill.field_rng(0, 0, 0))
    assert c.min() >= 0 and c.max() < 1


def test_hashfill_known_values():
    # splitmix64 reference values: the first outputs of the generator
    # seeded with 0 are 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4.
    assert int(hashfill.mix64(0, 0)) == 0xE220A8397B1DCDAF
    assert int(hashfill.mix64(0, 1)) == 0x6E789E6AA1B965F4
    u = hashfill.uniform(0, np.arange(1000))
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.05


def test_ranges_and_pretty_parameters():
    assert cli.unpack_ranges(a=1) == [{"a": 1}]
    r = cli.ArgRange("x", (1, 2))
    assert cli.unpack_ranges(a=(r, cli.ArgRange("x", None), 3)) == [
        {"a": (1, 1, 3)}, {"a": (2, 2, 3)}]
    with pytest.raises(ValueError):
        cli.ArgRange("x", (1,)).unique_values(cli.ArgRange("x", (2,)))
    parsed = cli.range_type(int).convert("[1-16:*2]", None, None)
    assert parsed.values == (1, 2, 4, 8, 16)


def test_return_time():
    @timing.return_time
    def work():
        sum(range(1000))

    assert work()["time"] > 0
