"""Serving loop: results equal alone and in the mix; oracle spot check."""

import numpy as np
import pytest

from conftest import inner
from oracle import oracle

pytestmark = pytest.mark.gpu


def _requests():
    from stencil_benchmarks_amd import serve

    reqs = serve.poisson_requests(rate=0.0, count=7, seed=3, nmin=8, nmax=96,
                                  nz_choices=(2, 9, 40), smin=1, smax=4)
    return reqs


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_serve_results_alone_equal_mix(gpu, dtype):
    from stencil_benchmarks_amd import serve

    mixed = serve.Server(dtype=dtype, keep_results=True)
    stats = mixed.run(_requests(), realtime=False)
    assert stats.requests == 7 and stats.max_active > 1
    for req in _requests():
        alone = serve.Server(dtype=dtype, keep_results=True)
        alone.run([req], realtime=False)
        u_mix, ts_mix = mixed.results[req.rid]
        u_one, ts_one = alone.results[req.rid]
        assert np.array_equal(u_mix, u_one) and np.array_equal(ts_mix, ts_one), req


def test_serve_under_memory_pressure_keeps_results(gpu):
    # an HBM budget of the largest request plus 1 MiB: admission waits (head of line)
    # for retirements, with a tick in flight, and reuses freed arena blocks
    from stencil_benchmarks_amd import serve

    largest = max(serve.footprint_bytes(r.domain, (3, 3, 3), "float64") for r in _requests())
    tight = serve.Server(dtype="float64", keep_results=True, hbm_budget=largest + (1 << 20))
    stats = tight.run(_requests(), realtime=False)
    assert stats.requests == 7
    roomy = serve.Server(dtype="float64", keep_results=True)
    roomy.run(_requests(), realtime=False)
    assert stats.max_active < 7 and stats.ticks > max(r.steps for r in _requests())
    for rid, (u, ts) in roomy.results.items():
        assert np.array_equal(u, tight.results[rid][0]) and np.array_equal(ts, tight.results[rid][1])


def test_serve_realtime_arrivals_keep_results(gpu):
    # arrivals spread over time (idle waits, joins at later ticks, a tick in
    # flight while the next is enqueued): same results as the closed loop
    from stencil_benchmarks_amd import serve

    def stream():
        return serve.poisson_requests(rate=3000.0, count=12, seed=5, nmin=8, nmax=96,
                                      nz_choices=(2, 9, 40), smin=1, smax=6)

    timed = serve.Server(dtype="float32", keep_results=True)
    stats = timed.run(stream(), realtime=True)
    assert stats.requests == 12 and stats.ticks >= 6
    assert all(lat > 0 for lat in stats.latencies_s)
    closed = serve.Server(dtype="float32", keep_results=True)
    closed.run(stream(), realtime=False)
    for rid, (u, ts) in closed.results.items():
        assert np.array_equal(u, timed.results[rid][0]) and np.array_equal(ts, timed.results[rid][1])


def test_events_order_host_waits(gpu):
    import ctypes

    from stencil_benchmarks_amd.device import DeviceBuffer

    n = 1 << 20
    buf = DeviceBuffer(8 * n)
    marker = gpu.Event()
    try:
        gpu.call("sb_stream_fill", gpu.SB_F64, ctypes.c_void_p(buf.ptr), n, 2.5)
        marker.record()
        marker.wait()
        worst = ctypes.c_double(1.0)
        gpu.call("sb_stream_check", gpu.SB_F64, ctypes.c_void_p(buf.ptr), n, 2.5,
                 ctypes.byref(worst))
        assert worst.value == 0.0
    finally:
        marker.close()
        buf.free()


def test_serve_matches_oracle(gpu):
    from stencil_benchmarks_amd import batch, serve
    from test_gpu_batch import oracle_step

    req = _requests()[0]
    server = serve.Server(dtype="float64", keep_results=True)
    server.run([req], realtime=False)
    # rebuild the request's initial fields on the device and read them back
    sub = batch.Subdomain(req.domain, (3, 3, 3), "float64", seed=req.seed)
    fields = {"u": sub.download("u0")}
    for name in ("coeff", "upos", "utens", "utensstage", "wcon"):
        fields[name] = sub.download(name)
    for _ in range(req.steps):
        oracle_step(fields, req.domain, (3, 3, 3), np.float64)
    sl = inner(req.domain, (3, 3, 3))
    u, ts = server.results[req.rid]
    assert np.array_equal(u, fields["u"][sl]) and np.array_equal(ts, fields["utensstage"][sl])


@pytest.mark.gpu
def test_poisson_stream_shape():
    from stencil_benchmarks_amd import serve

    reqs = serve.poisson_requests(rate=100.0, count=2000, seed=0)
    sizes = np.array([r.domain[0] for r in reqs])
    assert sizes.min() >= 16 and sizes.max() <= 1024
    assert {r.domain[2] for r in reqs} == {40, 80, 137}
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 0.01) < 0.002
    assert 1 <= min(r.steps for r in reqs) and max(r.steps for r in reqs) <= 20
