"""Data-set cycling through the GPU plugins (the reference's `data_sets`
option, stencils/base.py:47-51 and :152): runs take the sets in turn, each
run is verified, and the sets evolve independently (the device mirror of one
set never sees another set's run)."""

import copy

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def test_hdiff_data_sets_cycle_and_stay_apart(gpu):
    from stencil_benchmarks_amd.benchmarks_collection.stencils.mi355x import horizontal_diffusion

    bench = horizontal_diffusion.Fused(domain=(96, 40, 24), data_sets=2, verify=True)
    assert [bench.data_set_index] == [0]
    for run in range(4):
        assert bench.run()["time"] > 0
        assert bench.data_set_index == (run + 1) % 2
    inner = bench.inner_slice()
    outs = []
    for data in bench._data:
        expected = oracle.hdiff(data.inp, data.coeff, bench.domain, bench.halo)
        assert np.array_equal(data.out[inner], expected)
        outs.append(data.out[inner].copy())
    assert not np.array_equal(outs[0], outs[1])


def test_vadv_data_sets_advance_separately(gpu):
    from stencil_benchmarks_amd.benchmarks_collection.stencils.mi355x import vertical_advection

    bench = vertical_advection.WarpSpecialised(domain=(80, 30, 40), data_sets=2, verify=True)
    initial = [copy.deepcopy(data) for data in bench._data]
    for _ in range(3):  # set 0, set 1, set 0
        assert bench.run()["time"] > 0
    domain, halo = bench.domain, bench.halo
    inner = bench.inner_slice()
    for data, start, times in zip(bench._data, initial, (2, 1)):
        state = bench.empty_field()  # the fields' own strides, as the oracle needs
        state[...] = start.utensstage
        for _ in range(times):
            stepped = oracle.vadv(start.ustage, start.upos, start.utens, state, start.wcon,
                                  domain, halo)
            state[inner] = stepped
        assert np.array_equal(data.utensstage[inner], state[inner])
