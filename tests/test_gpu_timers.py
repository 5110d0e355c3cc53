"""Timers of a timed launch: events around the launch ('gpu') and events the
kernels record themselves ('hip-ext', hipExtLaunchKernelGGL), as the
reference's `timers` choices (stencils/cuda_hip/mixin.py:57).  Both give a
positive time and leave the results bit-identical to the untimed launch."""

import ctypes

import numpy as np
import pytest

from conftest import inner, make_fields

pytestmark = pytest.mark.gpu


def _hdiff(native, fields, domain, halo, options, seconds):
    from stencil_benchmarks_amd.device import DeviceField, geometry

    devs = [DeviceField(f, halo) for f in fields]
    for d, f in zip(devs, fields):
        d.upload(f)
    geom = geometry(domain, [s // 8 for s in fields[0].strides], halo)
    native.call("sb_hdiff", native.SB_F64, *[ctypes.c_void_p(d.pointer()) for d in devs], geom,
                options, seconds)
    native.call("sb_synchronize")
    out = np.empty_like(fields[2])
    devs[2].download(out)
    for d in devs:
        d.free()
    return out


@pytest.mark.parametrize("timer", [0, 1])
def test_timers_time_hdiff_and_keep_results(gpu, timer):
    domain, halo = (70, 33, 9), (3, 3, 3)
    fields = make_fields(domain, halo, np.float64, count=3, seed=4)
    plain = _hdiff(gpu, fields, domain, halo, gpu.launch_options(), None)
    seconds = ctypes.c_double(0.0)
    timed = _hdiff(gpu, fields, domain, halo,
                   gpu.launch_options(dry_runs=1, flush=True, timer=timer), ctypes.byref(seconds))
    assert 0 < seconds.value < 1.0
    assert np.array_equal(plain[inner(domain, halo)], timed[inner(domain, halo)])


def test_unknown_timer_is_rejected(gpu):
    from stencil_benchmarks_amd.benchmark import ExecutionError

    domain, halo = (8, 8, 4), (3, 3, 3)
    fields = make_fields(domain, halo, np.float64, count=3)
    with pytest.raises(ExecutionError, match="unknown timer"):
        _hdiff(gpu, fields, domain, halo, gpu.launch_options(timer=7),
               ctypes.byref(ctypes.c_double()))


@pytest.mark.parametrize("timers", ["gpu", "hip-ext", "wall"])
def test_plugin_timers(gpu, timers):
    from stencil_benchmarks_amd.benchmarks_collection.stencils.mi355x import (
        basic, horizontal_diffusion, vertical_advection)

    for cls in (horizontal_diffusion.Fused, vertical_advection.WarpSpecialised, basic.Laplacian):
        bench = cls(domain=(64, 40, 24), timers=timers)
        assert bench.run()["time"] > 0


def test_stream_hip_ext_timer(gpu):
    from stencil_benchmarks_amd.device import DeviceBuffer

    # a length that is not a multiple of the vector width: a second (tail)
    # kernel runs, and the stop event is the one recorded after it
    n = (1 << 20) + 3
    bufs = [DeviceBuffer(n * 8) for _ in range(3)]
    seconds = ctypes.c_double(0.0)
    gpu.call("sb_stream_kernel", gpu.SB_F64, gpu.STREAM_OPS["triad"],
             *[ctypes.c_void_p(b.ptr) for b in bufs], n, 3.0,
             gpu.launch_options(timer=gpu.SB_TIMER_HIP_EXT), ctypes.byref(seconds))
    assert seconds.value > 0
    for b in bufs:
        b.free()
