"""The autotuner on the device: it picks options from its spaces, caches
them, and the tuned options give bit-identical results to the defaults."""

import numpy as np
import pytest

from stencil_benchmarks_amd import autotune, batch

pytestmark = pytest.mark.gpu

SPECS = [batch.SubdomainSpec((70, 37, 50)), batch.SubdomainSpec((130, 20, 97)),
         batch.SubdomainSpec((33, 66, 40))]


@pytest.mark.parametrize("storage", ["float64", "float32"])
def test_tune_batch_caches_and_keeps_results(gpu, tmp_path, storage):
    cache = autotune.TuneCache(tmp_path / "tune.json")
    tuned = autotune.tune_batch(SPECS, storage, "float64", cache=cache)
    assert not tuned.from_cache and tuned.seconds > 0
    assert tuned.hdiff in autotune.BATCH_HDIFF_SPACE and tuned.vadv in autotune.BATCH_VADV_SPACE
    again = autotune.tune_batch(SPECS, storage, "float64", cache=autotune.TuneCache(cache.path))
    assert again.from_cache and (again.hdiff, again.vadv) == (tuned.hdiff, tuned.vadv)

    results = []
    for options in ((None, None), (tuned.hdiff_options, tuned.vadv_options)):
        subs = [batch.Subdomain(s.domain, s.halo, storage, seed=i) for i, s in enumerate(SPECS)]
        for _ in range(2):
            batch.step(subs, compute="float64", hdiff_options=options[0], vadv_options=options[1])
        results.append([(s.state(), s.download("utensstage")) for s in subs])
        for s in subs:
            s.free()
    for (u0, t0), (u1, t1) in zip(*results):
        assert np.array_equal(u0, u1) and np.array_equal(t0, t1)


def test_tune_hdiff_picks_a_tile(gpu, tmp_path):
    cache = autotune.TuneCache(tmp_path / "tune.json")
    tuned = autotune.tune_hdiff((96, 80, 24), "float64", cache=cache, repeats=1)
    assert tuned.hdiff in autotune.PLAIN_HDIFF_SPACE and len(tuned.trials) == 60


@pytest.mark.parametrize("storage", ["float64", "float32"])
def test_every_batch_candidate_gives_the_default_results(gpu, storage):
    # the tuner may pick any candidate: each must compute the same bits
    def run(hdiff, vadv):
        subs = [batch.Subdomain(s.domain, s.halo, storage, seed=i) for i, s in enumerate(SPECS)]
        for _ in range(2):
            batch.step(subs, compute="float64", hdiff_options=autotune.launch_from(hdiff),
                       vadv_options=autotune.launch_from(vadv))
        out = [(s.state(), s.download("utensstage")) for s in subs]
        for s in subs:
            s.free()
        return out

    reference = run({}, {})
    candidates = ([(h, {}) for h in autotune.BATCH_HDIFF_SPACE]
                  + [({}, v) for v in autotune.BATCH_VADV_SPACE])
    for hdiff, vadv in candidates:
        for (u0, t0), (u1, t1) in zip(reference, run(hdiff, vadv)):
            assert np.array_equal(u0, u1) and np.array_equal(t0, t1), (hdiff, vadv)


def _vadv_with(native, fields, domain, halo, options):
    import ctypes

    from stencil_benchmarks_amd.device import DeviceField, geometry

    devs = [DeviceField(f, halo) for f in fields]
    for d, f in zip(devs, fields):
        d.upload(f)
    geom = geometry(domain, [st // fields[0].itemsize for st in fields[0].strides], halo)
    pointers = (ctypes.c_void_p * 8)(*([d.pointer() for d in devs] + [0, 0, 0]))
    native.call("sb_vadv", native.SB_F64, 0, pointers, geom, autotune.launch_from(options), None)
    native.call("sb_synchronize")
    out = np.ndarray(fields[3].shape, fields[3].dtype, buffer=np.zeros(devs[3].span, np.uint8),
                     strides=fields[3].strides)
    devs[3].download(out)
    for d in devs:
        d.free()
    return out


@pytest.mark.parametrize("domain,solver", [((96, 80, 24), None), ((64, 4, 200), (2, 5)),
                                           ((64, 4, 230), (2,))])
def test_tune_vadv_picks_a_solver_that_keeps_results(gpu, tmp_path, domain, solver):
    from conftest import inner, make_fields
    from oracle import oracle

    cache = autotune.TuneCache(tmp_path / "tune.json")
    tuned = autotune.tune_vadv(domain, "float64", cache=cache, repeats=1)
    assert tuned.vadv in autotune.PLAIN_VADV_SPACE
    assert len(tuned.trials) == len(autotune.PLAIN_VADV_SPACE)
    if solver is not None:
        # columns beyond the LDS of the other solvers: the one-wave solver
        # (registers + LDS, nz <= 201) and the global scratch take them
        assert tuned.vadv["variant"] in solver
    again = autotune.tune_vadv(domain, "float64", cache=autotune.TuneCache(cache.path))
    assert again.from_cache and again.vadv == tuned.vadv
    halo = (3, 3, 3)
    fields = make_fields(domain, halo, np.float64, count=5, seed=7)
    got = _vadv_with(gpu, fields, domain, halo, tuned.vadv)
    assert np.array_equal(got[inner(domain, halo)], oracle.vadv(*fields, domain, halo))
