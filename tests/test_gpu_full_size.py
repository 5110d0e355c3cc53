"""Parity at the benchmark workloads' full sizes, through the C ABI.

The other GPU tests sweep small and ragged shapes; these run the shapes the
measurements are quoted on: horizontal diffusion and the x/y Laplacian at
512x512x160 (halo 3, 64-byte alignment, layout (2, 1, 0)), vertical
advection at 512x512x160 and 256x256x80, STREAM at 2**29 elements, the
96-subdomain batch, a batch of one 512x512x160 subdomain, and the served
request mix.  The stencils are compared
with the C oracle bit for bit (FP64, and float32 storage with FP64
arithmetic against the FP64 oracle on the float32 inputs, rounded once);
STREAM with the reference's closed-form recurrence; the batch and the
served requests (too large for the oracle in seconds) with the plain
kernels, which the oracle pins, applied subdomain by subdomain.
"""

import ctypes

import numpy as np
import pytest

from conftest import inner, make_fields
from oracle import oracle

pytestmark = pytest.mark.gpu

HALO = (3, 3, 3)


def _device(fields, domain, halo):
    from stencil_benchmarks_amd.device import DeviceField, geometry

    devs = [DeviceField(f, halo) for f in fields]
    for d, f in zip(devs, fields):
        d.upload(f)
    geom = geometry(domain, [s // fields[0].itemsize for s in fields[0].strides], halo)
    return devs, geom


def _download(dev, like):
    out = np.ndarray(like.shape, like.dtype, buffer=np.zeros(dev.span, np.uint8),
                     strides=like.strides)
    dev.download(out)
    return out


def _widened(fields):
    """float64 copies of float32 fields (same shape, C layout)."""
    return [f.astype(np.float64) for f in fields]


@pytest.mark.parametrize("storage,compute", [("float64", "float64"), ("float32", "float64"),
                                             ("float32", "float32")])
def test_hdiff_512x512x160(gpu, storage, compute):
    domain = (512, 512, 160)
    dtype = np.dtype(storage)
    inp, coeff, out = make_fields(domain, HALO, dtype, alignment=64, count=3, seed=21)
    out[...] = -1.0
    devs, geom = _device([inp, coeff, out], domain, HALO)
    try:
        code = gpu.precision_code(dtype, compute)
        ptrs = [ctypes.c_void_p(d.pointer()) for d in devs]
        gpu.call("sb_hdiff", code, *ptrs, geom, gpu.launch_options(), None)
        gpu.call("sb_synchronize")
        got = _download(devs[2], out)
    finally:
        for d in devs:
            d.free()
    if dtype == np.float64 or compute == "float32":  # the oracle in the storage type
        expected = oracle.hdiff(inp, coeff, domain, HALO)
    else:
        wide_inp, wide_coeff = _widened([inp, coeff])
        expected = oracle.hdiff(wide_inp, wide_coeff, domain, HALO).astype(np.float32)
    assert np.array_equal(got[inner(domain, HALO)], expected)
    halo_mask = np.ones(got.shape, bool)
    halo_mask[inner(domain, HALO)] = False
    assert np.all(got[halo_mask] == -1.0)


def test_laplacian_512x512x160(gpu):
    domain = (512, 512, 160)
    inp, out = make_fields(domain, HALO, np.float64, alignment=64, count=2, seed=22)
    devs, geom = _device([inp, out], domain, HALO)
    try:
        ptrs = [ctypes.c_void_p(d.pointer()) for d in devs]
        gpu.call("sb_laplacian", gpu.SB_F64, 1, 1, 0, *ptrs, geom, gpu.launch_options(), None)
        gpu.call("sb_synchronize")
        got = _download(devs[1], out)
    finally:
        for d in devs:
            d.free()
    expected = oracle.laplacian(inp, (1, 1, 0), domain, HALO)
    assert np.array_equal(got[inner(domain, HALO)], expected)


@pytest.mark.parametrize("domain,storage", [((512, 512, 160), "float64"),
                                            ((256, 256, 80), "float64"),
                                            ((256, 256, 80), "float32")])
def test_vadv_full_size(gpu, domain, storage):
    dtype = np.dtype(storage)
    fields = make_fields(domain, HALO, dtype, alignment=64, count=5, seed=23)
    devs, geom = _device(fields, domain, HALO)
    try:
        code = gpu.precision_code(dtype, np.float64)
        pointers = (ctypes.c_void_p * 8)(*([d.pointer() for d in devs] + [0, 0, 0]))
        gpu.call("sb_vadv", code, 0, pointers, geom, gpu.launch_options(), None)
        gpu.call("sb_synchronize")
        got = _download(devs[3], fields[3])
    finally:
        for d in devs:
            d.free()
    if dtype == np.float64:
        expected = oracle.vadv(*fields, domain, HALO)
    else:
        expected = oracle.vadv(*_widened(fields), domain, HALO).astype(np.float32)
    assert np.array_equal(got[inner(domain, HALO)], expected)


def test_vadv_all_components_256x256x80(gpu):
    # u, v and w with the wcon neighbour shifts (1, 0), (0, 1) and (0, 0)
    domain = (256, 256, 80)
    fields = make_fields(domain, HALO, np.float64, alignment=64, count=13, seed=24)
    devs, geom = _device(fields, domain, HALO)
    try:
        pointers = (ctypes.c_void_p * 16)(*([d.pointer() for d in devs] + [0, 0, 0]))
        gpu.call("sb_vadv", gpu.SB_F64, 1, pointers, geom, gpu.launch_options(), None)
        gpu.call("sb_synchronize")
        got = [_download(devs[3 + 4 * c], fields[3 + 4 * c]) for c in range(3)]
    finally:
        for d in devs:
            d.free()
    wcon = fields[12]
    for c, shift in enumerate([(1, 0), (0, 1), (0, 0)]):
        stage, pos, tens, ts = fields[4 * c:4 * c + 4]
        expected = oracle.vadv(stage, pos, tens, ts, wcon, domain, HALO, shift)
        assert np.array_equal(got[c][inner(domain, HALO)], expected), shift


@pytest.mark.parametrize("init", ["constant", "random"])
def test_stream_2_pow_29(gpu, init):
    from stencil_benchmarks_amd.benchmarks_collection.stream import mi355x

    bench = mi355x.Native(array_size=2 ** 29, ntimes=3, dtype="float64", init=init, verify=True)
    try:
        rows = bench.run()  # raises ExecutionError when the check fails
    finally:
        for buffer in bench.arrays:
            buffer.free()
    assert {row["name"] for row in rows} == {"copy", "scale", "add", "triad"}
    assert all(row["bandwidth"] > 0 for row in rows)


def _plain_step(native, sub, code):
    f = sub.fields
    cur, nxt = f[f"u{sub.parity}"], f[f"u{sub.parity ^ 1}"]
    native.call("sb_hdiff", code, ctypes.c_void_p(cur.pointer()),
                ctypes.c_void_p(f["coeff"].pointer()), ctypes.c_void_p(nxt.pointer()), sub.geom,
                native.launch_options(), None)
    ptrs = [nxt, f["upos"], f["utens"], f["utensstage"], f["wcon"]]
    pointers = (ctypes.c_void_p * 8)(*([p.pointer() for p in ptrs] + [0, 0, 0]))
    native.call("sb_vadv", code, 0, pointers, sub.geom, native.launch_options(), None)
    sub.advance(1)


@pytest.mark.parametrize("storage", ["float64", "float32"])
def test_batch_of_96_subdomains_equals_plain_kernels(gpu, storage):
    from stencil_benchmarks_amd import batch

    specs = batch.random_specs(96)
    code = gpu.precision_code(np.dtype(storage), np.float64)
    batched = [batch.Subdomain(s.domain, s.halo, storage, seed=s.seed) for s in specs]
    try:
        for _ in range(2):
            batch.step(batched, compute="float64")
        for spec, sub in zip(specs, batched):
            plain = batch.Subdomain(spec.domain, spec.halo, storage, seed=spec.seed)
            try:
                for _ in range(2):
                    _plain_step(gpu, plain, code)
                gpu.call("sb_synchronize")
                sl = inner(sub.domain, sub.halo)
                assert np.array_equal(sub.state()[sl], plain.state()[sl]), spec.domain
                assert np.array_equal(sub.download("utensstage")[sl],
                                      plain.download("utensstage")[sl]), spec.domain
            finally:
                plain.free()
    finally:
        for sub in batched:
            sub.free()


@pytest.mark.parametrize("planes_per_item", [1, 2])
def test_batch_of_one_512x512x160_subdomain_equals_plain_kernels(gpu, planes_per_item):
    # one large subdomain: 40,960 or 20,480 hdiff items over the per-XCD
    # queues, so every workgroup draws many tickets and steals from the
    # other queues at the end
    from stencil_benchmarks_amd import batch

    domain, halo = (512, 512, 160), (3, 3, 3)
    batched = batch.Subdomain(domain, halo, "float64", seed=7)
    plain = batch.Subdomain(domain, halo, "float64", seed=7)
    try:
        options = gpu.launch_options(block=(0, 16, planes_per_item))
        for _ in range(2):
            batch.step([batched], compute="float64", hdiff_options=options)
            _plain_step(gpu, plain, gpu.SB_F64)
        gpu.call("sb_synchronize")
        sl = inner(domain, halo)
        assert np.array_equal(batched.state()[sl], plain.state()[sl])
        assert np.array_equal(batched.download("utensstage")[sl], plain.download("utensstage")[sl])
    finally:
        batched.free()
        plain.free()


def test_served_request_mix_equals_plain_kernels(gpu):
    # the serving workload's request mix (nx = ny from 16 to 1024, nz 40 / 80 /
    # 137, 1-20 steps; float32 storage, FP64 arithmetic), all admitted at once
    from stencil_benchmarks_amd import batch, serve

    requests = serve.poisson_requests(rate=0.0, count=40, seed=1)
    server = serve.Server(dtype="float32", compute="float64", keep_results=True)
    stats = server.run(requests, realtime=False)
    assert stats.requests == 40 and stats.max_active > 1
    code = gpu.precision_code(np.float32, np.float64)
    for req in serve.poisson_requests(rate=0.0, count=40, seed=1):
        plain = batch.Subdomain(req.domain, (3, 3, 3), "float32", seed=req.seed)
        try:
            for _ in range(req.steps):
                _plain_step(gpu, plain, code)
            gpu.call("sb_synchronize")
            sl = inner(req.domain, (3, 3, 3))
            u, ts = server.results[req.rid]
            assert np.array_equal(u, plain.state()[sl]), req
            assert np.array_equal(ts, plain.download("utensstage")[sl]), req
        finally:
            plain.free()
