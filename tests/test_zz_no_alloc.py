"""sb_vadv_noalloc (the entry point that allocates nothing) against sb_vadv.

Both run the global-scratch column solver; the no-allocation entry keeps its
(c', d') scratch in a fixed array of the library and strides a persistent
grid over the columns.  Results must equal sb_vadv's bit for bit (and the
oracle's on the small domain), on two calls in a row, the second with a
smaller domain, so the array's leftover contents from the first call cannot
leak into the second.
"""

import ctypes

import numpy as np
import pytest

from conftest import inner, make_fields
from oracle import oracle

pytestmark = pytest.mark.gpu


def _vadv(native, entry, fields, domain, halo, code, all_components, **launch):
    from stencil_benchmarks_amd.device import DeviceField, geometry

    devs = [DeviceField(f, halo) for f in fields]
    for d, f in zip(devs, fields):
        d.upload(f)
    geom = geometry(domain, [s // fields[0].itemsize for s in fields[0].strides], halo)
    array = (ctypes.c_void_p * len(devs))(*[d.pointer() for d in devs])
    opts = native.launch_options(**launch)
    native.call(entry, code, int(all_components), array, geom, opts, None)
    native.call("sb_synchronize")
    written = [3, 7, 11] if all_components else [3]
    out = []
    for w in written:
        result = np.ndarray(fields[w].shape, fields[w].dtype,
                            buffer=np.zeros(devs[w].span, np.uint8), strides=fields[w].strides)
        devs[w].download(result)
        out.append(result[inner(domain, halo)])
    for d in devs:
        d.free()
    return out


# (domain, halo, dtype, precision name, all components); the first case has
# more 64-column tiles (4 x 300) than workgroups fit the scratch (409 at
# nz = 160 with FP64 arithmetic), so workgroups take several tiles each
CASES = [((200, 300, 160), (3, 3, 3), np.float64, "SB_F64", False),
         ((70, 37, 9), (3, 3, 3), np.float64, "SB_F64", False)]


def test_two_calls_match_sb_vadv_then_a_smaller_one(gpu):
    for domain, halo, dtype, prec, allc in CASES:
        print(f"noalloc case domain={domain} prec={prec}", flush=True)
        fields = make_fields(domain, halo, dtype, count=8, seed=21)
        code = getattr(gpu, prec)
        reference = _vadv(gpu, "sb_vadv", fields, domain, halo, code, allc, variant=2)
        got = _vadv(gpu, "sb_vadv_noalloc", fields, domain, halo, code, allc)
        for g, r in zip(got, reference):
            assert np.array_equal(g, r)
    # the smaller domain also against the oracle
    expected = oracle.vadv(*fields[:5], domain, halo)
    assert np.array_equal(got[0], expected)


@pytest.mark.parametrize("prec,dtype", [("SB_F32_F64", np.float32), ("SB_F32", np.float32)])
def test_float32_storage_and_all_components(gpu, prec, dtype):
    for domain in ((150, 20, 61), (33, 9, 7)):
        print(f"noalloc case domain={domain} prec={prec}", flush=True)
        fields = make_fields(domain, (1, 1, 1), dtype, count=16, seed=5)
        code = getattr(gpu, prec)
        reference = _vadv(gpu, "sb_vadv", fields, domain, (1, 1, 1), code, True, variant=2)
        got = _vadv(gpu, "sb_vadv_noalloc", fields, domain, (1, 1, 1), code, True)
        for g, r in zip(got, reference):
            assert np.array_equal(g, r)


def test_limits_are_errors(gpu):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    domain, halo = (8, 2, 4), (1, 1, 1)
    fields = make_fields(domain, halo, np.float64, count=8)
    with pytest.raises(ExecutionError, match="flush"):
        _vadv(gpu, "sb_vadv_noalloc", fields, domain, halo, gpu.SB_F64, False, flush=True)
    with pytest.raises(ExecutionError, match="variant"):
        _vadv(gpu, "sb_vadv_noalloc", fields, domain, halo, gpu.SB_F64, False, variant=4)
    # 65537 FP64 levels need more than 64 MiB of scratch per 64 columns
    tall = (1, 1, 65537)
    fields = make_fields(tall, (1, 0, 0), np.float64, count=8)
    with pytest.raises(ExecutionError, match="scratch"):
        _vadv(gpu, "sb_vadv_noalloc", fields, tall, (1, 0, 0), gpu.SB_F64, False)
