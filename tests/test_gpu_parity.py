"""HIP kernels vs the oracle, through the C ABI, on seeded inputs.

Bar: FP64 results equal the oracle's bit for bit (the kernels follow the
reference oracle's operation order with no FMA contraction).  Float32
storage with float32 arithmetic equals the float32 oracle bit for bit.
Float32 storage with float64 arithmetic equals the float64 oracle run on
the same (float32-rounded) inputs, rounded once to float32, bit for bit.
Every case also checks that no point outside the interior of the output
was written (sentinel halo).
"""

import ctypes

import numpy as np
import pytest

from conftest import inner, make_fields
from oracle import oracle

pytestmark = pytest.mark.gpu

SENTINEL = -1234.5
PRECISIONS = ["f64", "f32", "f32_f64"]


def _prec(native, name):
    return {"f64": native.SB_F64, "f32": native.SB_F32, "f32_f64": native.SB_F32_F64}[name]


def _dtype(name):
    return np.float64 if name == "f64" else np.float32


def _expected(func, fields, prec, *args):
    """Oracle result in the storage type for precision `prec`."""
    if prec == "f32_f64":
        widened = [f.astype(np.float64) for f in fields]
        from stencil_benchmarks_amd.tools import array

        clones = []
        for f, w in zip(fields, widened):
            c = array.alloc_array(f.shape, np.float64, _layout_of(f), 0, (0, 0, 0))
            c[...] = w
            clones.append(c)
        return func(*clones, *args).astype(np.float32)
    return func(*fields, *args)


def _layout_of(a):
    order = np.argsort(-np.array(a.strides))
    layout = [0, 0, 0]
    for rank, axis in enumerate(order):
        layout[axis] = rank
    return tuple(layout)


def _run(native, entry, fields, written, domain, halo, *pre_args, launch=None):
    from stencil_benchmarks_amd.device import DeviceField, geometry

    devs = [DeviceField(f, halo) for f in fields]
    for d, f in zip(devs, fields):
        d.upload(f)
    geom = geometry(domain, [s // fields[0].itemsize for s in fields[0].strides], halo)
    ptrs = [ctypes.c_void_p(d.pointer()) for d in devs]
    opts = launch or native.launch_options()
    native.call(entry, *pre_args, *ptrs, geom, opts, None)
    native.call("sb_synchronize")
    results = []
    for w in written:
        out = np.ndarray(fields[w].shape, fields[w].dtype,
                         buffer=np.zeros(devs[w].span, np.uint8), strides=fields[w].strides)
        devs[w].download(out)
        results.append(out)
    return results


def _with_sentinel(field):
    field[...] = SENTINEL
    return field


def _check_halo(out, domain, halo):
    mask = np.ones(out.shape, dtype=bool)
    mask[inner(domain, halo)] = False
    assert np.all(out[mask] == SENTINEL), "kernel wrote outside the interior"


BASIC_DOMAINS = [(64, 16, 4), (37, 13, 5), (1, 1, 1), (130, 3, 2)]
GEOMETRIES = [((2, 1, 0), 0), ((2, 1, 0), 64), ((0, 1, 2), 0), ((1, 2, 0), 16)]


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("layout,alignment", GEOMETRIES)
@pytest.mark.parametrize("domain", BASIC_DOMAINS)
def test_basic_stencils(gpu, prec, layout, alignment, domain):
    halo = (1, 2, 1)
    dtype = _dtype(prec)
    code = _prec(gpu, prec)
    inp, = make_fields(domain, halo, dtype, layout, alignment, seed=hash(domain) % 1000)
    cases = [("sb_copy", (), lambda f: oracle.copy(f, domain, halo))]
    for axis in range(3):
        cases.append(("sb_onesided_average", (axis,),
                      lambda f, a=axis: oracle.onesided_average(f, a, domain, halo)))
        cases.append(("sb_symmetric_average", (axis,),
                      lambda f, a=axis: oracle.symmetric_average(f, a, domain, halo)))
    for mask in [(1, 1, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)]:
        cases.append(("sb_laplacian", mask,
                      lambda f, m=mask: oracle.laplacian(f, m, domain, halo)))
    for entry, extra, func in cases:
        out = _with_sentinel(make_fields(domain, halo, dtype, layout, alignment)[0])
        got, = _run(gpu, entry, [inp, out], [1], domain, halo, code, *extra)
        expected = _expected(func, [inp], prec)
        assert np.array_equal(got[inner(domain, halo)], expected), (entry, extra)
        _check_halo(got, domain, halo)


MARCH_DOMAINS = [(130, 70, 3), (64, 33, 2), (37, 9, 4), (256, 100, 2), (2, 65, 1)]


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("domain", MARCH_DOMAINS)
@pytest.mark.parametrize("variant,rows", [(1, 0), (2, 8), (2, 32), (2, 64)])
def test_basic_stencil_kernels(gpu, prec, domain, variant, rows):
    """Row kernel (variant 1) and marching kernel (variant 2, `rows` per
    thread) on vector-aligned fields: runs shorter than, equal to and
    longer than the rows per thread, ragged ends of the unit-stride axis."""
    halo = (1, 2, 1)
    dtype = _dtype(prec)
    code = _prec(gpu, prec)
    inp, = make_fields(domain, halo, dtype, (2, 1, 0), 64, seed=sum(domain))
    cases = []
    for axis in (0, 1) if variant == 2 else (0, 1, 2):
        cases.append(("sb_onesided_average", (axis,),
                      lambda f, a=axis: oracle.onesided_average(f, a, domain, halo)))
        cases.append(("sb_symmetric_average", (axis,),
                      lambda f, a=axis: oracle.symmetric_average(f, a, domain, halo)))
    masks = [(1, 1, 0), (1, 0, 0), (0, 1, 0)] + ([(1, 1, 1)] if variant == 1 else [])
    for mask in masks:
        cases.append(("sb_laplacian", mask,
                      lambda f, m=mask: oracle.laplacian(f, m, domain, halo)))
    opts = gpu.launch_options()
    opts.variant, opts.bz = variant, rows
    for entry, extra, func in cases:
        out = _with_sentinel(make_fields(domain, halo, dtype, (2, 1, 0), 64)[0])
        got, = _run(gpu, entry, [inp, out], [1], domain, halo, code, *extra, launch=opts)
        expected = _expected(func, [inp], prec)
        assert np.array_equal(got[inner(domain, halo)], expected), (entry, extra)
        _check_halo(got, domain, halo)


def test_marching_kernel_rejects_axis_2_terms(gpu):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    domain, halo = (64, 8, 4), (1, 1, 1)
    inp, out = make_fields(domain, halo, np.float64, (2, 1, 0), 64, count=2)
    opts = gpu.launch_options()
    opts.variant = 2
    with pytest.raises(ExecutionError, match="marching kernel"):
        _run(gpu, "sb_laplacian", [inp, out], [1], domain, halo, gpu.SB_F64, 1, 1, 1, launch=opts)


@pytest.mark.parametrize("tile", [0, 8, 16, 32])
@pytest.mark.parametrize("domain", [(64, 32, 1), (128, 96, 3), (192, 32, 2), (64, 64, 5)])
@pytest.mark.parametrize("halo,alignment", [((1, 1, 0), 0), ((3, 3, 3), 64), ((1, 2, 1), 8)])
def test_laplacian_lds_dma_tile(gpu, tile, domain, halo, alignment):
    # variant 3: FP64 x/y Laplacian on 64 x tile tiles staged by global-to-LDS
    # loads; the tile's corner is 8-byte aligned only (halo 1, no padding)
    inp, = make_fields(domain, halo, np.float64, (2, 1, 0), alignment, seed=sum(domain) + tile)
    out = _with_sentinel(make_fields(domain, halo, np.float64, (2, 1, 0), alignment)[0])
    opts = gpu.launch_options(variant=3, block=(0, tile, 0))
    got, = _run(gpu, "sb_laplacian", [inp, out], [1], domain, halo, gpu.SB_F64, 1, 1, 0,
                launch=opts)
    assert np.array_equal(got[inner(domain, halo)], oracle.laplacian(inp, (1, 1, 0), domain, halo))
    _check_halo(got, domain, halo)


@pytest.mark.parametrize("domain", [(64, 16, 3), (64, 24, 3), (128, 8, 2)])
def test_laplacian_default_tile_rows(gpu, domain):
    # the default takes 16-row tiles where ny allows, else 8-row ones
    halo = (1, 1, 1)
    inp, = make_fields(domain, halo, np.float64, (2, 1, 0), 0, seed=sum(domain))
    out = _with_sentinel(make_fields(domain, halo, np.float64, (2, 1, 0), 0)[0])
    got, = _run(gpu, "sb_laplacian", [inp, out], [1], domain, halo, gpu.SB_F64, 1, 1, 0)
    assert np.array_equal(got[inner(domain, halo)], oracle.laplacian(inp, (1, 1, 0), domain, halo))
    _check_halo(got, domain, halo)


@pytest.mark.parametrize("domain,prec,mask,tile,match", [
    ((70, 16, 2), "f64", (1, 1, 0), 0, "multiple of 64"),
    ((64, 12, 2), "f64", (1, 1, 0), 0, "multiple of 8"),
    ((64, 16, 2), "f64", (1, 1, 0), 32, "multiple of 32"),
    ((64, 16, 2), "f64", (1, 1, 0), 4, "tile rows 4"),
    ((64, 16, 2), "f64", (1, 1, 1), 0, "along i and j"),
    ((64, 16, 2), "f64", (1, 0, 0), 0, "along i and j"),
    ((64, 16, 2), "f32_f64", (1, 1, 0), 0, "FP64")])
def test_laplacian_lds_dma_tile_rejects(gpu, domain, prec, mask, tile, match):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    halo = (1, 1, 1)
    dtype = _dtype(prec)
    inp, out = make_fields(domain, halo, dtype, count=2)
    opts = gpu.launch_options(variant=3, block=(0, tile, 0))
    with pytest.raises(ExecutionError, match=match):
        _run(gpu, "sb_laplacian", [inp, out], [1], domain, halo, _prec(gpu, prec), *mask,
             launch=opts)


@pytest.mark.parametrize("entry", ["sb_hdiff", "sb_laplacian"])
def test_lds_dma_tiles_reject_negative_row_stride(gpu, entry):
    # the tiles address rows by 32-bit offsets from the tile corner: a flipped
    # j axis (negative row stride) is refused, so the default, which asks the
    # same check, takes another kernel for it
    from stencil_benchmarks_amd.benchmark import ExecutionError
    from stencil_benchmarks_amd.device import DeviceField, geometry

    domain, halo = (64, 16, 2), (2, 2, 1)
    fields = make_fields(domain, halo, np.float64, (2, 1, 0), 64, count=3)
    devs = [DeviceField(f, halo) for f in fields]
    try:
        strides = [s // 8 for s in fields[0].strides]
        geom = geometry(domain, [strides[0], -strides[1], strides[2]], halo)
        # logical row 0 is the last interior row of the allocation
        ptrs = [ctypes.c_void_p(d.pointer() + (domain[1] - 1) * strides[1] * 8) for d in devs]
        if entry == "sb_hdiff":
            args = (gpu.SB_F64, *ptrs)
            opts = gpu.launch_options(variant=2)
        else:
            args = (gpu.SB_F64, 1, 1, 0, ptrs[0], ptrs[1])
            opts = gpu.launch_options(variant=3)
        with pytest.raises(ExecutionError, match="positive row pitch"):
            gpu.call(entry, *args, geom, opts, None)
    finally:
        for d in devs:
            d.free()


HDIFF_DOMAINS = [(70, 37, 9), (64, 16, 1), (1, 1, 1), (129, 33, 3), (5, 70, 2)]


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("variant,tile,planes", [(1, 8, 1), (1, 16, 8), (1, 32, 2), (1, 16, 16),
                                                  (3, 0, 0)])
@pytest.mark.parametrize("domain", HDIFF_DOMAINS)
def test_hdiff(gpu, prec, variant, tile, planes, domain):
    halo = (2, 3, 0)
    dtype = _dtype(prec)
    inp, coeff = make_fields(domain, halo, dtype, count=2, seed=sum(domain))
    out = _with_sentinel(make_fields(domain, halo, dtype)[0])
    opts = gpu.launch_options(variant=variant, block=(64, tile, planes))
    got, = _run(gpu, "sb_hdiff", [inp, coeff, out], [2], domain, halo, _prec(gpu, prec),
                launch=opts)
    expected = _expected(lambda a, b: oracle.hdiff(a, b, domain, halo), [inp, coeff], prec)
    assert np.array_equal(got[inner(domain, halo)], expected)
    _check_halo(got, domain, halo)


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("tile", [8, 16])
@pytest.mark.parametrize("domain", [(64, 16, 1), (128, 48, 5), (192, 16, 3), (256, 32, 2)])
@pytest.mark.parametrize("halo,alignment", [((2, 3, 0), 0), ((3, 3, 3), 64), ((3, 2, 1), 4)])
def test_hdiff_lds_dma_tile(gpu, prec, tile, domain, halo, alignment):
    # variant 2: tiles staged by global-to-LDS loads (16-byte units for FP64
    # fields, 4-byte ones for float32), whole tiles only; halo 3 with no
    # padding leaves float32 tile corners 4-byte aligned only
    if prec == "f64" and alignment == 4:
        pytest.skip("FP64 fields need a 16-byte aligned origin")
    dtype = _dtype(prec)
    inp, coeff = make_fields(domain, halo, dtype, (2, 1, 0), alignment, count=2,
                             seed=sum(domain) + tile)
    out = _with_sentinel(make_fields(domain, halo, dtype, (2, 1, 0), alignment)[0])
    opts = gpu.launch_options(variant=2, block=(0, tile, 0))
    got, = _run(gpu, "sb_hdiff", [inp, coeff, out], [2], domain, halo, _prec(gpu, prec),
                launch=opts)
    expected = _expected(lambda a, b: oracle.hdiff(a, b, domain, halo), [inp, coeff], prec)
    assert np.array_equal(got[inner(domain, halo)], expected)
    _check_halo(got, domain, halo)


@pytest.mark.parametrize("domain,prec,tile,match", [((70, 16, 2), "f64", 0, "multiple of 64"),
                                                    ((64, 12, 2), "f64", 0, "multiple of 8"),
                                                    ((70, 16, 2), "f32_f64", 0, "multiple of 64"),
                                                    ((64, 24, 2), "f32", 16, "multiple of 16"),
                                                    ((64, 32, 2), "f32_f64", 32, "not in {8, 16}")])
def test_hdiff_lds_dma_tile_rejects(gpu, domain, prec, tile, match):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    halo = (2, 2, 0)
    dtype = _dtype(prec)
    inp, coeff, out = make_fields(domain, halo, dtype, count=3)
    opts = gpu.launch_options(variant=2, block=(0, tile, 0))
    with pytest.raises(ExecutionError, match=match):
        _run(gpu, "sb_hdiff", [inp, coeff, out], [2], domain, halo, _prec(gpu, prec), launch=opts)


@pytest.mark.parametrize("layout,alignment", [((0, 1, 2), 0), ((1, 0, 2), 8)])
def test_hdiff_other_layouts(gpu, layout, alignment):
    domain, halo = (33, 21, 7), (3, 3, 3)
    inp, coeff = make_fields(domain, halo, np.float64, layout, alignment, count=2)
    out = _with_sentinel(make_fields(domain, halo, np.float64, layout, alignment)[0])
    got, = _run(gpu, "sb_hdiff", [inp, coeff, out], [2], domain, halo, gpu.SB_F64)
    assert np.array_equal(got[inner(domain, halo)], oracle.hdiff(inp, coeff, domain, halo))
    _check_halo(got, domain, halo)


VADV_DOMAINS = [(70, 37, 9), (64, 1, 2), (1, 1, 2), (65, 3, 33)]


def _vadv_expected(fields, domain, halo, prec, components):
    ncomp = len(components)
    wcon = fields[4 * ncomp]
    out = []
    for q, comp in enumerate(components):
        f = fields[4 * q: 4 * q + 4] + [wcon]
        shift = oracle.COMPONENT_SHIFTS[comp]
        out.append(_expected(lambda *a: oracle.vadv(*a, domain, halo, shift), f, prec))
    return out


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("all_components", [False, True])
@pytest.mark.parametrize("domain", VADV_DOMAINS)
def test_vadv(gpu, prec, variant, all_components, domain):
    halo = (1, 1, 0)
    dtype = _dtype(prec)
    components = ("u", "v", "w") if all_components else ("u",)
    n = 4 * len(components) + 1
    fields = make_fields(domain, halo, dtype, count=n, seed=7 + sum(domain))
    scratch = make_fields(domain, halo, dtype, count=3)
    written = [4 * q + 3 for q in range(len(components))]
    got = _run_vadv(gpu, fields + scratch, written, domain, halo, _prec(gpu, prec),
                    all_components, variant)
    expected = _vadv_expected(fields, domain, halo, prec, components)
    for g, e in zip(got, expected):
        assert np.array_equal(g[inner(domain, halo)], e)


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("domain", [(128, 700, 19), (70, 530, 9), (64, 1300, 2)])
@pytest.mark.parametrize("depth", [4, 8])
def test_vadv_pipelined_many_groups(gpu, prec, domain, depth):
    # more column groups than resident workgroups: every workgroup streams
    # several groups through the fused forward/backward pipeline
    halo = (1, 0, 0)
    dtype = _dtype(prec)
    fields = make_fields(domain, halo, dtype, count=8, seed=sum(domain))
    got, = _run_vadv(gpu, fields, [3], domain, halo, _prec(gpu, prec), False, 3,
                     extra=dict(unroll=depth, bz=1))
    expected = _expected(lambda *f: oracle.vadv(*f, domain, halo), fields[:5], prec)
    assert np.array_equal(got[inner(domain, halo)], expected)


# nz around the warp-specialised solver's block (4 levels) and prefetch ring
# (16 levels), up to the FP64 LDS limit of 160 levels; more column groups
# than resident workgroups
WS_DOMAINS = [(128, 700, 19), (70, 530, 9), (64, 1300, 2), (65, 90, 17), (130, 40, 20),
              (96, 3, 41), (200, 60, 160), (33, 17, 5)]


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("domain", WS_DOMAINS)
@pytest.mark.parametrize("nt", [1, -1])
@pytest.mark.parametrize("levels", [0, 8])  # levels per barrier step (0: 4)
def test_vadv_warp_specialised(gpu, prec, domain, nt, levels):
    halo = (1, 0, 0)
    dtype = _dtype(prec)
    fields = make_fields(domain, halo, dtype, count=8, seed=sum(domain))
    got, = _run_vadv(gpu, fields, [3], domain, halo, _prec(gpu, prec), False, 4,
                     extra=dict(bz=nt, unroll=levels))
    expected = _expected(lambda *f: oracle.vadv(*f, domain, halo), fields[:5], prec)
    assert np.array_equal(got[inner(domain, halo)], expected)


# nz around the register part of the one-wave solver (40 levels with FP64
# arithmetic, 80 with float32) and its LDS blocks (8 levels), the register
# part alone, and the LDS limit (nz = 201 with FP64 arithmetic)
HYBRID_DOMAINS = [(70, 37, 2), (65, 9, 3), (64, 20, 39), (64, 20, 40), (64, 20, 41),
                  (96, 30, 42), (130, 40, 48), (33, 17, 49), (200, 60, 80), (64, 70, 81),
                  (70, 11, 97), (64, 64, 160), (65, 3, 201)]


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("domain", HYBRID_DOMAINS)
@pytest.mark.parametrize("nt", [0, 1])
def test_vadv_one_wave_register_columns(gpu, prec, domain, nt):
    halo = (1, 0, 0)
    dtype = _dtype(prec)
    fields = make_fields(domain, halo, dtype, count=8, seed=3 * sum(domain))
    got, = _run_vadv(gpu, fields, [3], domain, halo, _prec(gpu, prec), False, 5,
                     extra=dict(bz=nt))
    expected = _expected(lambda *f: oracle.vadv(*f, domain, halo), fields[:5], prec)
    assert np.array_equal(got[inner(domain, halo)], expected)


# more column groups than the resident waves hold at once: each wave walks
# several rounds (1280 groups at nz 40 against 1024 waves; 1680 groups at
# nz 96, where two waves fit a CU)
@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("domain", [(320, 256, 40), (448, 240, 96)])
def test_vadv_one_wave_register_columns_several_rounds(gpu, prec, domain):
    halo = (1, 0, 0)
    dtype = _dtype(prec)
    fields = make_fields(domain, halo, dtype, count=8, seed=sum(domain))
    got, = _run_vadv(gpu, fields, [3], domain, halo, _prec(gpu, prec), False, 5)
    expected = _expected(lambda *f: oracle.vadv(*f, domain, halo), fields[:5], prec)
    assert np.array_equal(got[inner(domain, halo)], expected)


def test_vadv_one_wave_rejects_columns_beyond_lds(gpu):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    domain, halo = (8, 2, 202), (1, 0, 0)
    fields = make_fields(domain, halo, np.float64, count=8)
    with pytest.raises(ExecutionError, match="LDS"):
        _run_vadv(gpu, fields, [3], domain, halo, gpu.SB_F64, False, 5)


def test_vadv_warp_specialised_rejects_other_step_sizes(gpu):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    domain, halo = (64, 2, 8), (1, 0, 0)
    fields = make_fields(domain, halo, np.float64, count=8, seed=1)
    with pytest.raises(ExecutionError, match="levels per step"):
        _run_vadv(gpu, fields, [3], domain, halo, gpu.SB_F64, False, 4, extra=dict(unroll=16))


# Tolerance of the fast reciprocal (float32 storage, FP64 arithmetic): the
# result differs from the FP64 oracle on the float32 inputs by at most
# 1e-6 relative to max(|oracle|, 1e-3), well inside the 1e-5 bound of the
# float32 form.
FAST_RTOL = 1e-6


@pytest.mark.parametrize("domain", [(200, 60, 160), (70, 37, 9), (65, 90, 17), (256, 64, 80)])
def test_vadv_fast_reciprocal(gpu, domain):
    halo = (1, 0, 0)
    fields = make_fields(domain, halo, np.float32, count=8, seed=5 + sum(domain))
    got, = _run_vadv(gpu, fields, [3], domain, halo, gpu.SB_F32_F64, False, 4,
                     extra=dict(bz=4 | 1))
    expected = _expected(lambda *f: oracle.vadv(*f, domain, halo), fields[:5], "f32_f64")
    got = got[inner(domain, halo)].astype(np.float64)
    ref = expected.astype(np.float64)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)
    assert err.max() <= FAST_RTOL, err.max()


def test_vadv_fast_reciprocal_needs_float32_storage(gpu):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    domain, halo = (64, 2, 8), (1, 0, 0)
    fields = make_fields(domain, halo, np.float64, count=8, seed=1)
    with pytest.raises(ExecutionError, match="fast reciprocal"):
        _run_vadv(gpu, fields, [3], domain, halo, gpu.SB_F64, False, 4, extra=dict(bz=4))


@pytest.mark.parametrize("flags", [1, 2, 3])
@pytest.mark.parametrize("variant,depth", [(1, 4), (1, 16), (2, 8)])
def test_vadv_streaming_options(gpu, flags, variant, depth):
    domain, halo = (130, 5, 41), (1, 1, 0)
    fields = make_fields(domain, halo, np.float64, count=8, seed=11)
    got, = _run_vadv(gpu, fields, [3], domain, halo, gpu.SB_F64, False, variant,
                     extra=dict(bz=flags, unroll=depth))
    assert np.array_equal(got[inner(domain, halo)], oracle.vadv(*fields[:5], domain, halo))


def _run_vadv(native, fields, written, domain, halo, code, all_components, variant, extra=None):
    from stencil_benchmarks_amd.device import DeviceField, geometry

    devs = [DeviceField(f, halo) for f in fields]
    for d, f in zip(devs, fields):
        d.upload(f)
    geom = geometry(domain, [s // fields[0].itemsize for s in fields[0].strides], halo)
    array = (ctypes.c_void_p * len(devs))(*[d.pointer() for d in devs])
    opts = native.launch_options(variant=variant)
    for key, value in (extra or {}).items():
        setattr(opts, key, value)
    native.call("sb_vadv", code, int(all_components), array, geom, opts, None)
    native.call("sb_synchronize")
    results = []
    for w in written:
        out = np.ndarray(fields[w].shape, fields[w].dtype,
                         buffer=np.zeros(devs[w].span, np.uint8), strides=fields[w].strides)
        devs[w].download(out)
        results.append(out)
    return results


@pytest.mark.parametrize("nz", [161, 201, 202, 300])
def test_vadv_tall_columns_default(gpu, nz):
    # nz > 160 exceeds the LDS of the FP64 LDS-only solvers; variant 0 takes
    # the one-wave register + LDS solver up to nz = 201 and the global
    # scratch beyond, and must still match
    domain, halo = (67, 2, nz), (1, 0, 0)
    fields = make_fields(domain, halo, np.float64, count=8, seed=3)
    got, = _run_vadv(gpu, fields, [3], domain, halo, gpu.SB_F64, False, 0)
    expected = oracle.vadv(*fields[:5], domain, halo)
    assert np.array_equal(got[inner(domain, halo)], expected)


# (unroll, non-temporal bits, elements per vector (0: 16 bytes), workgroups
# per CU of a persistent grid (0: one pass), grid axis)
STREAM_SHAPES = [(1, 0, 0, 0, 0), (4, 3, 0, 0, 0), (8, 1, 0, 0, 0), (2, 0, 1, 0, 0),
                 (1, 3, 2, 2, 0), (4, 0, 1, 1, 0), (1, 0, 0, 0, 1), (2, 3, 1, 0, 2),
                 (1, 0, 0, 3, 2)]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("op", ["copy", "scale", "add", "triad"])
@pytest.mark.parametrize("unroll,nt,vector,per_cu,axis", STREAM_SHAPES)
def test_stream_kernels(gpu, dtype, op, unroll, nt, vector, per_cu, axis):
    from stencil_benchmarks_amd.device import DeviceBuffer

    # large enough for several passes of a persistent grid along y or z
    n = 100003 if per_cu == 0 else 3_000_017
    code = gpu.SB_F64 if dtype == "float64" else gpu.SB_F32
    rng = np.random.default_rng(5)
    host = [rng.random(n).astype(dtype) for _ in range(3)]
    bufs = [DeviceBuffer(n * host[0].itemsize) for _ in range(3)]
    for b, h in zip(bufs, host):
        gpu.call("sb_memcpy_h2d", ctypes.c_void_p(b.ptr), ctypes.c_void_p(h.ctypes.data), h.nbytes)
    opts = gpu.launch_options(variant=nt | axis << 2, block=(256, vector, per_cu), unroll=unroll)
    gpu.call("sb_stream_kernel", code, gpu.STREAM_OPS[op], *[ctypes.c_void_p(b.ptr) for b in bufs],
             n, 3.0, opts, None)
    gpu.call("sb_synchronize")
    got = []
    for b, h in zip(bufs, host):
        out = np.empty_like(h)
        gpu.call("sb_memcpy_d2h", ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(b.ptr), h.nbytes)
        got.append(out)
    a, b, c = host
    s = np.dtype(dtype).type(3.0)
    expected = {"copy": (a, b, a), "scale": (a, s * c, c), "add": (a, b, a + b),
                "triad": (b + s * c, b, c)}[op]
    for g, e in zip(got, expected):
        assert np.array_equal(g, e)


def test_stream_rejects_bad_shapes(gpu):
    from stencil_benchmarks_amd.benchmark import ExecutionError
    from stencil_benchmarks_amd.device import DeviceBuffer

    bufs = [DeviceBuffer(4096) for _ in range(3)]
    ptrs = [ctypes.c_void_p(b.ptr) for b in bufs]
    for opts in (gpu.launch_options(block=(256, 4, 0)),      # 4 doubles per vector
                 gpu.launch_options(block=(256, 0, 65)),     # persistent grid too large
                 gpu.launch_options(variant=3 << 2)):        # no fourth axis
        with pytest.raises(ExecutionError):
            gpu.call("sb_stream_kernel", gpu.SB_F64, 0, *ptrs, 512, 3.0, opts, None)


def test_fill_random_fields_matches_host_hash(gpu):
    from stencil_benchmarks_amd.device import DeviceBuffer
    from stencil_benchmarks_amd.tools.hashfill import uniform

    n = 3001
    for dtype, code in (("float64", gpu.SB_F64), ("float32", gpu.SB_F32)):
        bufs = [DeviceBuffer(n * np.dtype(dtype).itemsize) for _ in range(5)]
        seeds = [7, 7, 1 << 40, 3, 0]  # a repeated seed gives identical arrays
        ptrs = (ctypes.c_void_p * 5)(*[b.ptr for b in bufs])
        gpu.call("sb_fill_random_fields", code, ptrs, 5, n, (ctypes.c_uint64 * 5)(*seeds))
        gpu.call("sb_synchronize")
        for buf, seed in zip(bufs, seeds):
            out = np.empty(n, dtype)
            gpu.call("sb_memcpy_d2h", ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(buf.ptr),
                     out.nbytes)
            assert np.array_equal(out, uniform(seed, np.arange(n), dtype)), seed


def test_subdomain_fill_shares_the_boundary(gpu):
    from stencil_benchmarks_amd import batch

    sub = batch.Subdomain((37, 20, 9), (3, 3, 3), "float32", seed=5)
    u0, u1, wcon = sub.download("u0"), sub.download("u1"), sub.download("wcon")
    assert np.array_equal(u0, u1) and not np.array_equal(u0, wcon)
    sub.free()


def test_stream_random_fill_matches_host_hash(gpu):
    from stencil_benchmarks_amd.device import DeviceBuffer
    from stencil_benchmarks_amd.tools.hashfill import uniform

    n = 4099
    for dtype, code in (("float64", gpu.SB_F64), ("float32", gpu.SB_F32)):
        buf = DeviceBuffer(n * np.dtype(dtype).itemsize)
        gpu.call("sb_stream_fill_random", code, ctypes.c_void_p(buf.ptr), n, 11)
        out = np.empty(n, dtype)
        gpu.call("sb_memcpy_d2h", ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(buf.ptr), out.nbytes)
        assert np.array_equal(out, uniform(11, np.arange(n), dtype))
        assert out.min() >= 0 and out.max() < 1
