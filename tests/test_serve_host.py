"""Host-side bookkeeping of the serving loop (CPU only): request streams,
footprints and the rejection of requests that have nothing to do."""

import numpy as np
import pytest

from stencil_benchmarks_amd import serve


def test_requests_without_steps_are_rejected_before_any_device_work():
    good = serve.Request(0, 0.0, (16, 16, 40), 3, 1)
    for steps in (0, -2):
        bad = serve.Request(1, 0.0, (16, 16, 40), steps, 2)
        with pytest.raises(ValueError, match="steps must be at least 1"):
            serve.Server(hbm_budget=1 << 30).run([good, bad], realtime=False)


def test_poisson_stream_follows_its_distribution():
    reqs = serve.poisson_requests(rate=500.0, count=2000, seed=7)
    arrivals = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arrivals) >= 0)
    # mean inter-arrival time 1/rate (2000 samples: within 10 %)
    assert abs(np.diff(arrivals).mean() * 500.0 - 1.0) < 0.1
    sizes = np.array([r.domain[0] for r in reqs])
    assert all(r.domain[0] == r.domain[1] for r in reqs)
    assert sizes.min() >= 16 and sizes.max() <= 1024
    # log-uniform: the median of exp(U[ln 16, ln 1024]) is 128
    assert 110 <= np.median(sizes) <= 150
    assert {r.domain[2] for r in reqs} == {40, 80, 137}
    steps = np.array([r.steps for r in reqs])
    assert steps.min() == 1 and steps.max() == 20
    assert len({r.seed for r in reqs}) == len(reqs)


def test_footprint_covers_seven_padded_fields():
    one = serve.footprint_bytes((100, 100, 40), (3, 3, 3), "float32")
    # 106 x 106 x 46 points, rows padded to 64 bytes (112 floats)
    assert one >= 7 * 112 * 106 * 46 * 4
    assert serve.footprint_bytes((100, 100, 40), (3, 3, 3), "float64") > 1.9 * one
