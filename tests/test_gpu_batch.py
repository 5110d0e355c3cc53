"""Batched hdiff + vadv steps vs the oracle, through the C ABI.

FP64 batches equal the oracle step loop bit for bit; float32 storage with
float64 arithmetic equals the float64 oracle applied to the stored float32
state and rounded once per step.  A subdomain's result does not depend on
the other subdomains in the batch (alone == in the mix), and replays of a
captured hipGraph equal eager execution.
"""

import numpy as np
import pytest

from conftest import inner
from oracle import oracle

pytestmark = pytest.mark.gpu

DOMAINS = [(70, 37, 9), (64, 16, 2), (1, 1, 2), (130, 5, 41), (33, 66, 17)]


def _host_fields(sub, rng):
    fields = {}
    for name in ("u", "coeff", "upos", "utens", "utensstage", "wcon"):
        a = sub.empty_host()
        a[...] = rng.random(a.shape).astype(sub.dtype)
        fields[name] = a
    return fields


def _widen(a):
    out = np.zeros(a.shape, np.float64)
    out[...] = a
    return out


def oracle_step(fields, domain, halo, dtype):
    """One step of the batch wiring on host fields (mutates and returns them)."""
    wide = dtype != np.float64
    get = (lambda n: _widen(fields[n])) if wide else (lambda n: fields[n])
    u = get("u")
    coeff = get("coeff")
    u_next = u.copy() if wide else oracle._clone(u, copy_values=True)
    u_next[inner(domain, halo)] = oracle.hdiff(u, coeff, domain, halo)
    if wide:
        # the device stores u_next in float32 before vadv reads it
        u_next = _widen(u_next.astype(dtype))
    ts = oracle.vadv(u_next, get("upos"), get("utens"), get("utensstage"), get("wcon"),
                     domain, halo, (1, 0))
    fields["u"] = u_next.astype(dtype) if wide else u_next
    fields["utensstage"][inner(domain, halo)] = ts.astype(dtype) if wide else ts
    return fields


def _make_batch(dtype, seed=1, domains=DOMAINS):
    from stencil_benchmarks_amd import batch

    rng = np.random.default_rng(seed)
    subs, hosts = [], []
    for domain in domains:
        sub = batch.Subdomain(domain, (3, 3, 3), dtype, alignment=64, fill=False)
        hf = _host_fields(sub, rng)
        sub.upload(hf)
        subs.append(sub)
        hosts.append(hf)
    return subs, hosts


def _check(subs, hosts, dtype):
    for sub, hf in zip(subs, hosts):
        sl = inner(sub.domain, sub.halo)
        assert np.array_equal(sub.state()[sl], hf["u"][sl]), sub.domain
        assert np.array_equal(sub.download("utensstage")[sl], hf["utensstage"][sl]), sub.domain


# columns up to the LDS limit of the FP64 solvers (nz = 160)
TALL_DOMAINS = [(100, 20, 160), (64, 8, 133), (40, 40, 50), (3, 70, 97)]


@pytest.mark.parametrize("prec", ["f64", "f32_f64"])
# (variant, unroll): single-wave pipeline; warp-specialised solver at 4 and
# 8 levels per barrier step
@pytest.mark.parametrize("variant,unroll", [(3, 0), (4, 0), (4, 8)], ids=["pipe", "ws4", "ws8"])
@pytest.mark.parametrize("domains", [DOMAINS, TALL_DOMAINS], ids=["short", "tall"])
def test_batch_steps_match_oracle(gpu, prec, variant, unroll, domains):
    from stencil_benchmarks_amd import batch, native

    dtype = np.float64 if prec == "f64" else np.float32
    subs, hosts = _make_batch(dtype, domains=domains)
    options = native.launch_options(variant=variant, unroll=unroll)
    for _ in range(3):
        batch.step(subs, compute=np.float64, vadv_options=options)
        for sub, hf in zip(subs, hosts):
            oracle_step(hf, sub.domain, sub.halo, dtype)
    _check(subs, hosts, dtype)


def test_batch_plan_graph_replays_match_eager(gpu):
    from stencil_benchmarks_amd import batch

    subs, hosts = _make_batch(np.float64, seed=4)
    twins, _ = _make_batch(np.float64, seed=4)
    plan = batch.Plan(subs, steps=2)
    for replay in range(3):
        plan.run(use_graph=True)
        # eager twin between replays
        batch.step(twins)
        batch.step(twins)
        for sub, hf in zip(subs, hosts):
            oracle_step(hf, sub.domain, sub.halo, np.float64)
            oracle_step(hf, sub.domain, sub.halo, np.float64)
        _check(subs, hosts, np.float64)
        _check(twins, hosts, np.float64)
    plan.close()


def test_batch_alone_equals_mix(gpu):
    from stencil_benchmarks_amd import batch

    subs, _ = _make_batch(np.float64, seed=9)
    alone, _ = _make_batch(np.float64, seed=9)
    for _ in range(2):
        batch.step(subs)
        for sub in alone:
            batch.step([sub])
    for mixed, single in zip(subs, alone):
        sl = inner(mixed.domain, mixed.halo)
        assert np.array_equal(mixed.state()[sl], single.state()[sl])
        assert np.array_equal(mixed.download("utensstage")[sl], single.download("utensstage")[sl])


def test_batch_rejects_other_step_sizes(gpu):
    from stencil_benchmarks_amd import batch, native
    from stencil_benchmarks_amd.benchmark import ExecutionError

    sub = batch.Subdomain((64, 2, 20), (3, 3, 3), np.float64, seed=1)
    with pytest.raises(ExecutionError, match="levels per step"):
        batch.step([sub], vadv_options=native.launch_options(variant=4, unroll=6))
    sub.free()


def test_batch_rejects_tall_columns(gpu):
    from stencil_benchmarks_amd import batch
    from stencil_benchmarks_amd.benchmark import ExecutionError

    sub = batch.Subdomain((64, 2, 200), (3, 3, 3), np.float64, seed=1)
    with pytest.raises(ExecutionError, match="LDS"):
        batch.step([sub])
