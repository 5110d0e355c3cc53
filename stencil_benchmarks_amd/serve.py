"""Continuous-batching request server for hdiff + vadv solves.

No counterpart in the reference (which times single launches).  A request
is a subdomain solve: nx = ny, nz, a number of timesteps and a seed for
its U[0,1) fields.  The server

* admits requests in arrival order once they have arrived and their
  fields fit in the free-HBM budget (head-of-line: a request that does not
  fit blocks later ones, so large requests are not starved);
* keeps every admitted request's fields resident in HBM until it finishes,
  in a device arena allocated once per run, before its clock starts
  (admission and retirement are host bookkeeping: no allocation or release
  of device memory, whose release waits for the device, per request);
* advances all active requests together, one timestep per tick, in one
  batched launch pair (``sb_batch_step``): requests that arrive while
  others are on the GPU join at the next tick and share the persistent
  kernels' work queues with them (continuous batching);
* keeps one tick in flight: tick t+1 is built and enqueued before the host
  waits for tick t (a marker event behind each tick), so the per-tick host
  work overlaps the device's; a request is complete, and retired, once the
  tick that holds its last step has passed (arena blocks it frees are only
  reused by work enqueued later on the same stream);
* records per request the latency from arrival to completion.

Timestep semantics are those of :mod:`stencil_benchmarks_amd.batch`; a
request's results do not depend on which other requests it was batched
with (the kernels have no cross-subdomain arithmetic), which the tests
check bit for bit.
"""

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import batch, native
from .device import DeviceArena


@dataclass
class Request:
    rid: int
    arrival: float
    domain: tuple
    steps: int
    seed: int
    admitted: float = -1.0
    finished: float = -1.0
    remaining: int = 0
    sub: object = field(default=None, repr=False)

    @property
    def latency(self):
        return self.finished - self.arrival


def poisson_requests(rate, count, seed=0, nmin=16, nmax=1024, nz_choices=(40, 80, 137),
                     smin=1, smax=20):
    """`count` requests with exponential inter-arrival times (mean 1/rate s).

    nx = ny = round(exp(U[ln nmin, ln nmax])), nz uniform over `nz_choices`,
    steps ~ U{smin..smax}, per-request field seed.
    """
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        n = int(round(float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))))
        nz = int(rng.choice(nz_choices))
        steps = int(rng.integers(smin, smax + 1))
        out.append(Request(rid, t, (n, n, nz), steps, seed * 1000003 + rid))
    return out


def footprint_bytes(domain, halo, dtype, alignment=64):
    """Device bytes of a request's seven fields (padded like alloc_array,
    plus the 256-byte alignment of each field's interior origin)."""
    from .tools import array

    shape = tuple(d + 2 * h for d, h in zip(domain, halo))
    strides, span = array.compute_strides(shape, np.dtype(dtype).itemsize, (2, 1, 0), alignment)
    last = sum((n - 1) * s for n, s in zip(shape, strides)) + np.dtype(dtype).itemsize
    return 7 * (last + 3 * 256)


@dataclass
class ServeStats:
    requests: int
    wall_s: float
    ticks: int
    latencies_s: list
    bytes_moved: float
    point_steps: float
    max_active: int

    def summary(self):
        lat = np.array(self.latencies_s) if self.latencies_s else np.zeros(1)
        return {
            "requests": self.requests,
            "wall_s": self.wall_s,
            "ticks": self.ticks,
            "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_latency_ms": float(lat.mean() * 1e3),
            "throughput_requests_per_s": self.requests / self.wall_s if self.wall_s else 0.0,
            "aggregate_gbs": self.bytes_moved / self.wall_s / 1e9 if self.wall_s else 0.0,
            "point_steps_per_s": self.point_steps / self.wall_s if self.wall_s else 0.0,
            "max_active": self.max_active,
        }


class Server:
    def __init__(self, dtype="float32", compute="float64", halo=(3, 3, 3), hbm_budget=None,
                 hdiff_options=None, vadv_options=None, keep_results=False):
        self.dtype = np.dtype(dtype)
        self.compute = np.dtype(compute)
        self.halo = tuple(halo)
        self.budget = None if hbm_budget is None else int(hbm_budget)  # None: 60% of free HBM
        self.arena = None
        self.hdiff_options = hdiff_options
        self.vadv_options = vadv_options
        self.keep_results = keep_results
        self.results = {}

    def _admit(self, req, now):
        req.sub = batch.Subdomain(req.domain, self.halo, self.dtype, seed=req.seed,
                                  arena=self.arena)
        req.admitted = now
        req.remaining = req.steps

    def _retire(self, req, now):
        req.finished = now
        if self.keep_results:
            sl = tuple(slice(h, h + d) for d, h in zip(req.domain, self.halo))
            self.results[req.rid] = (req.sub.state()[sl].copy(),
                                     req.sub.download("utensstage")[sl].copy())
        req.sub.free()
        req.sub = None

    def run(self, requests, realtime=True):
        """Serve `requests` (arrival times relative to start); returns ServeStats.

        With realtime=False every request is available at t=0 (closed loop,
        measures capacity).
        """
        requests = list(requests)
        for req in requests:
            if req.steps < 1:
                raise ValueError(f"request {req.rid}: steps must be at least 1, got {req.steps}")
        budget = self.budget
        if budget is None:
            free, _ = native.mem_info()
            budget = int(free * 0.6)
        # no larger than every request resident at once needs
        need = sum(footprint_bytes(r.domain, self.halo, self.dtype) for r in requests)
        self.arena = DeviceArena(max(1, min(budget, need)))
        try:
            return self._serve(requests, realtime)
        finally:
            self.arena.close()
            self.arena = None

    def _serve(self, requests, realtime):
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        active, done = [], []
        ticks = 0
        bytes_moved = 0.0
        point_steps = 0.0
        max_active = 0
        # one tick in flight while the next is prepared and enqueued: the
        # host's per-tick work overlaps the device's; `inflight` holds the
        # marker behind the last enqueued tick and the requests whose last
        # step it holds (their results are complete once it has passed)
        markers = [native.Event(), native.Event()]
        inflight = None
        start = time.perf_counter()
        try:
            while pending or active or inflight:
                now = time.perf_counter() - start
                while pending and (not realtime or pending[0].arrival <= now):
                    req = pending[0]
                    try:
                        self._admit(req, now)
                    except MemoryError:
                        if active or inflight:
                            break  # wait for memory (head of line)
                        raise
                    pending.popleft()
                    if not realtime:
                        req.arrival = 0.0
                    active.append(req)
                tick = None
                if active:
                    max_active = max(max_active, len(active))
                    batch.step([r.sub for r in active], compute=self.compute,
                               hdiff_options=self.hdiff_options, vadv_options=self.vadv_options)
                    marker = markers[ticks & 1]
                    marker.record()
                    ticks += 1
                    for r in active:
                        bytes_moved += r.sub.compulsory_bytes_per_step()
                        point_steps += r.sub.points
                        r.remaining -= 1
                    tick = (marker, [r for r in active if r.remaining == 0])
                    active = [r for r in active if r.remaining > 0]
                if inflight is not None:
                    inflight[0].wait()
                    now = time.perf_counter() - start
                    for r in inflight[1]:
                        self._retire(r, now)
                        done.append(r)
                inflight = tick
                if inflight is None and not active and pending:
                    # idle until the next arrival
                    wait = pending[0].arrival - (time.perf_counter() - start)
                    if wait > 0:
                        time.sleep(min(wait, 0.001))
        finally:
            for marker in markers:
                marker.close()
        wall = time.perf_counter() - start
        return ServeStats(len(done), wall, ticks, [r.latency for r in done], bytes_moved,
                          point_steps, max_active)


def run_plain(requests, dtype="float64", compute=None, halo=(3, 3, 3)):
    """Serve requests one at a time with the plain (unbatched) kernels.

    Each step is one ``sb_hdiff`` and one ``sb_vadv`` launch on the
    request's own fields, which come from a device arena allocated before
    the clock starts (as the server's), in storage type `dtype` with
    arithmetic type `compute` (default: the storage type); returns (wall
    seconds, bytes, point-steps).
    """
    dtype = np.dtype(dtype)
    prec = native.precision_code(dtype, compute)
    total_bytes = 0.0
    point_steps = 0.0
    arena = DeviceArena(max(footprint_bytes(r.domain, halo, dtype) for r in requests))
    try:
        native.call("sb_synchronize")
        start = time.perf_counter()
        for req in requests:
            sub = batch.Subdomain(req.domain, halo, dtype, seed=req.seed, arena=arena)
            f = sub.fields
            for _ in range(req.steps):
                cur, nxt = f[f"u{sub.parity}"], f[f"u{sub.parity ^ 1}"]
                native.call("sb_hdiff", prec, ctypes.c_void_p(cur.pointer()),
                            ctypes.c_void_p(f["coeff"].pointer()), ctypes.c_void_p(nxt.pointer()),
                            sub.geom, native.launch_options(), None)
                ptrs = [nxt, f["upos"], f["utens"], f["utensstage"], f["wcon"]]
                arr = (ctypes.c_void_p * 8)(*([p.pointer() for p in ptrs] + [0, 0, 0]))
                native.call("sb_vadv", prec, 0, arr, sub.geom, native.launch_options(), None)
                sub.advance(1)
                total_bytes += sub.compulsory_bytes_per_step()
                point_steps += sub.points
            native.call("sb_synchronize")
            sub.free()
        elapsed = time.perf_counter() - start
    finally:
        arena.close()
    return elapsed, total_bytes, point_steps
