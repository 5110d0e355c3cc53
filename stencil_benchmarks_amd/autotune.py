"""Launch-shape autotuner for the batched and plain kernels.

The reference sweeps block sizes by hand (its collection scripts run fixed
`block_size` grids, stencil_benchmarks/scripts/sbench_*_collection.py, and
tools/multirun.py); here the sweep is a tuner that times candidate launch
options on the device and keeps the fastest per device and workload:

* batched steps (`tune_batch`): hdiff tile height, k-planes per item and
  cache hints, then the vadv solver (warp-specialised with 8/16/32 column
  groups per item, 4 or 8 levels per barrier step and 4 or 5 waves, or the
  single-wave pipeline at prefetch depth 4/8),
  chosen by coordinate descent (hdiff with the default vadv, then vadv with
  the chosen hdiff), each candidate timed over whole steps on scratch
  subdomains of the workload's geometry (user fields are never advanced);
* plain horizontal diffusion (`tune_hdiff`): tile height x k-planes x cache
  hints over one domain;
* plain vertical advection (`tune_vadv`): the warp-specialised solver at 4 or
  8 levels per barrier step or the single-wave pipeline at prefetch depth
  4/8/16, each with plain or non-temporal streamed loads, over one domain.

Every candidate computes the same arithmetic (the tests check that results
are bit-identical across launch options), so tuning changes speed only.
Choices persist as JSON (``$SBMI355X_TUNE_CACHE``, default
``~/.cache/stencil_benchmarks_amd/autotune.json``) keyed by device, storage
and arithmetic types and the workload's shapes; an unwritable cache only
costs re-tuning.
"""

import hashlib
import json
import os
import tempfile
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import native

# hdiff options of the batched step (launch->by tile rows, ->bz k-planes per
# item, ->unroll 4 = plain coeff loads instead of non-temporal ones): the
# combinations the batched kernel is built for
BATCH_HDIFF_SPACE = (
    {"by": 16, "bz": 8}, {"by": 16, "bz": 8, "unroll": 4}, {"by": 16, "bz": 4},
    {"by": 16, "bz": 2}, {"by": 16, "bz": 2, "unroll": 4}, {"by": 32, "bz": 2},
    {"by": 16, "bz": 1}, {"by": 8, "bz": 1},
)
# vadv options of the batched step (variant 4: ->by column groups per item,
# ->unroll levels per barrier step, ->bx waves per workgroup; variant 3:
# ->unroll prefetch depth)
BATCH_VADV_SPACE = (
    {"variant": 4, "by": 16}, {"variant": 4, "by": 8}, {"variant": 4, "by": 32},
    {"variant": 4, "by": 16, "unroll": 8}, {"variant": 4, "by": 8, "unroll": 8},
    {"variant": 4, "by": 32, "unroll": 8},
    {"variant": 4, "by": 8, "bx": 5}, {"variant": 4, "by": 16, "bx": 5},
    {"variant": 4, "by": 16, "unroll": 8, "bx": 5},
    {"variant": 3, "unroll": 8}, {"variant": 3, "unroll": 4},
)
# plain hdiff: register-staged tile rows x k-planes x cache hints (->unroll
# 0: non-temporal coeff loads; else 4 | bits: bit 0 non-temporal out stores,
# bit 1 non-temporal coeff loads); tune_hdiff adds the LDS-DMA tile's two
# heights (variant 2) for FP64 domains of whole tiles
PLAIN_HDIFF_SPACE = tuple({"variant": 1, "by": by, "bz": bz, "unroll": u}
                          for by in (8, 16, 32) for bz in (1, 2, 4, 8, 16) for u in (0, 4, 5, 7))
# plain vadv: variant 4 (->unroll levels per barrier step) or 3 (->unroll
# prefetch depth), ->bz 1 non-temporal streamed loads, -1 plain loads; and
# variant 2 (global scratch), the one solver for columns beyond the LDS
PLAIN_VADV_SPACE = tuple({"variant": v, "unroll": u, "bz": bz}
                         for v, depths in ((4, (4, 8)), (3, (4, 8, 16)))
                         for u in depths for bz in (1, -1)) + (
    {"variant": 5, "unroll": 0, "bz": -1}, {"variant": 5, "unroll": 0, "bz": 1},
    {"variant": 2, "unroll": 8, "bz": 1})


def default_cache_path():
    env = os.environ.get("SBMI355X_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "stencil_benchmarks_amd" / "autotune.json"


class TuneCache:
    """JSON file of tuned choices: {key: {"hdiff": {...}, "vadv": {...}, ...}}."""

    def __init__(self, path=None):
        self.path = Path(path) if path is not None else default_cache_path()
        try:
            self.entries = json.loads(self.path.read_text())
            if not isinstance(self.entries, dict):
                self.entries = {}
        except (OSError, ValueError):
            self.entries = {}

    def get(self, key):
        return self.entries.get(key)

    def put(self, key, value):
        self.entries[key] = value
        try:  # write a temporary file, then rename: readers never see half a file
            self.path.parent.mkdir(parents=True, exist_ok=True)
            fd, tmp = tempfile.mkstemp(dir=str(self.path.parent), suffix=".tmp")
            with os.fdopen(fd, "w") as handle:
                json.dump(self.entries, handle, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass


def launch_from(options):
    """sb_launch from an options dict (missing fields: library defaults)."""
    launch = native.launch_options()
    for name, value in (options or {}).items():
        setattr(launch, name, int(value))
    return launch


def device_key():
    info = native.device_info()
    return f"{info['name']}|{info['arch']}|{info['compute_units']}"


def workload_key(kind, domains, storage, compute, halo=(3, 3, 3)):
    shapes = json.dumps(sorted(tuple(int(d) for d in dom) for dom in domains))
    digest = hashlib.sha1(shapes.encode()).hexdigest()[:16]
    return (f"{kind}|{np.dtype(storage).name}|{np.dtype(compute).name}|halo={tuple(halo)}|"
            f"n={len(domains)}|{digest}")


def descend(space_a, space_b, measure, default_b):
    """Coordinate descent over two option spaces: the best `a` with
    `default_b`, then the best `b` with that `a`.  `measure(a, b)` returns
    seconds (lower is better).  Returns (a, b, seconds, trials)."""
    trials = []
    best_a, best_t = None, float("inf")
    for a in space_a:
        t = measure(a, default_b)
        trials.append((a, default_b, t))
        if t < best_t:
            best_a, best_t = a, t
    best_b = default_b
    for b in space_b:
        if b == default_b:
            continue
        t = measure(best_a, b)
        trials.append((best_a, b, t))
        if t < best_t:
            best_b, best_t = b, t
    return best_a, best_b, best_t, trials


@dataclass
class TuneResult:
    key: str
    hdiff: dict
    vadv: dict
    seconds: float
    from_cache: bool
    trials: list = field(default_factory=list)

    @property
    def hdiff_options(self):
        return launch_from(self.hdiff)

    @property
    def vadv_options(self):
        return launch_from(self.vadv)


def tune_batch(specs, storage, compute=None, cache=None, repeats=3, force=False, warmup=4):
    """Tuned launch options for batched steps over subdomains of `specs`
    (objects with .domain, .halo) in `storage` precision with `compute`
    arithmetic; times whole steps on scratch subdomains of the same shapes."""
    from . import batch

    storage = np.dtype(storage)
    compute = np.dtype(compute or storage)
    halo = tuple(specs[0].halo) if specs else (3, 3, 3)
    key = device_key() + "|" + workload_key("batch", [s.domain for s in specs], storage, compute,
                                            halo)
    cache = cache if cache is not None else TuneCache()
    hit = None if force else cache.get(key)
    if hit:
        return TuneResult(key, hit["hdiff"], hit["vadv"], hit["seconds"], True)
    scratch = [batch.Subdomain(s.domain, s.halo, storage, seed=index)
               for index, s in enumerate(specs)]
    try:
        # untimed steps first: the first launches on fresh fields and a cold
        # device run slower and would count against the first candidate
        for _ in range(warmup):
            batch.step(scratch, compute=compute)
        native.call("sb_synchronize")

        def measure(h, v):
            hl, vl = launch_from(h), launch_from(v)
            batch.step(scratch, compute=compute, hdiff_options=hl, vadv_options=vl)  # warm-up
            return min(batch.step(scratch, compute=compute, hdiff_options=hl, vadv_options=vl,
                                  timed=True) for _ in range(repeats))

        h, v, seconds, trials = descend(BATCH_HDIFF_SPACE, BATCH_VADV_SPACE, measure,
                                        BATCH_VADV_SPACE[0])
    finally:
        for sub in scratch:
            sub.free()
    cache.put(key, {"hdiff": h, "vadv": v, "seconds": seconds, "trials": len(trials)})
    return TuneResult(key, h, v, seconds, False, trials)


def tune_hdiff(domain, storage, compute=None, halo=(3, 3, 3), cache=None, repeats=3,
               force=False, warmup=20):
    """Tuned sb_hdiff launch options for one domain (cold-cache launches)."""
    import ctypes

    from .device import DeviceField
    from .tools import array

    storage = np.dtype(storage)
    compute = np.dtype(compute or storage)
    key = device_key() + "|" + workload_key("hdiff", [domain], storage, compute, halo)
    cache = cache if cache is not None else TuneCache()
    hit = None if force else cache.get(key)
    if hit:
        return TuneResult(key, hit["hdiff"], {}, hit["seconds"], True)
    shape = tuple(d + 2 * h for d, h in zip(domain, halo))
    template = array.alloc_array(shape, storage, (2, 1, 0), 64, halo)
    strides = [s // storage.itemsize for s in template.strides]
    fields = DeviceField.group(shape, storage, template.strides, 3, halo)
    code = native.precision_code(storage, compute)
    seeds = (ctypes.c_uint64 * 3)(1, 2, 3)
    bases = (ctypes.c_void_p * 3)(*[f.base for f in fields])
    native.call("sb_fill_random_fields", native.SB_F64 if storage == np.float64 else native.SB_F32,
                bases, 3, fields[0].span // storage.itemsize, seeds)
    from .device import geometry

    geom = geometry(domain, strides, halo)
    ptrs = [ctypes.c_void_p(f.pointer()) for f in fields]
    try:
        def measure(options, _unused):
            launch = launch_from(options)
            launch.flush = 1
            seconds = ctypes.c_double()
            runs = []
            for _ in range(repeats):
                native.call("sb_hdiff", code, *ptrs, geom, launch, ctypes.byref(seconds))
                runs.append(seconds.value)
            return min(runs)

        # untimed, with the library's default launch (it takes any domain): a
        # cold device would count against the first candidate
        for _ in range(warmup):
            native.call("sb_hdiff", code, *ptrs, geom, native.launch_options(), None)

        space = PLAIN_HDIFF_SPACE
        if domain[0] % 64 == 0:
            # the LDS-DMA tile (variant 2) takes whole 64 x 8 / 64 x 16 tiles
            # (the tuning fields have 64-byte aligned rows, as FP64 needs)
            space = space + tuple({"variant": 2, "by": by} for by in (8, 16) if domain[1] % by == 0)
        h, _, seconds, trials = descend(space, (), measure, {})
    finally:
        fields[0].free()
    cache.put(key, {"hdiff": h, "vadv": {}, "seconds": seconds, "trials": len(trials)})
    return TuneResult(key, h, {}, seconds, False, trials)


def tune_vadv(domain, storage, compute=None, halo=(3, 3, 3), cache=None, repeats=3,
              force=False, warmup=20):
    """Tuned sb_vadv launch options (u component) for one domain (cold-cache
    launches).  Solvers that the domain does not fit (LDS, 4 GiB buffer
    offsets) are skipped."""
    import ctypes

    from .benchmark import ExecutionError
    from .device import DeviceField, geometry
    from .tools import array

    storage = np.dtype(storage)
    compute = np.dtype(compute or storage)
    key = device_key() + "|" + workload_key("vadv", [domain], storage, compute, halo)
    cache = cache if cache is not None else TuneCache()
    hit = None if force else cache.get(key)
    if hit:
        return TuneResult(key, {}, hit["vadv"], hit["seconds"], True)
    shape = tuple(d + 2 * h for d, h in zip(domain, halo))
    template = array.alloc_array(shape, storage, (2, 1, 0), 64, halo)
    strides = [s // storage.itemsize for s in template.strides]
    # ustage, upos, utens, utensstage, wcon (the reference's ccol, dcol and
    # datacol are never touched)
    fields = DeviceField.group(shape, storage, template.strides, 5, halo)
    seeds = (ctypes.c_uint64 * 5)(1, 2, 3, 4, 5)
    bases = (ctypes.c_void_p * 5)(*[f.base for f in fields])
    native.call("sb_fill_random_fields", native.SB_F64 if storage == np.float64 else native.SB_F32,
                bases, 5, fields[0].span // storage.itemsize, seeds)
    code = native.precision_code(storage, compute)
    geom = geometry(domain, strides, halo)
    pointers = (ctypes.c_void_p * 8)(*([f.pointer() for f in fields] + [0, 0, 0]))
    try:
        def measure(_unused, options):
            launch = launch_from(options)
            launch.flush = 1
            seconds = ctypes.c_double()
            runs = []
            try:
                for _ in range(repeats):
                    native.call("sb_vadv", code, 0, pointers, geom, launch, ctypes.byref(seconds))
                    runs.append(seconds.value)
            except ExecutionError:  # this solver does not take the domain
                return float("inf")
            return min(runs)

        # untimed, with the library's default launch (it takes any domain): a
        # cold device would count against the first candidate
        for _ in range(warmup):
            native.call("sb_vadv", code, 0, pointers, geom, native.launch_options(), None)
        _, v, seconds, trials = descend(({},), PLAIN_VADV_SPACE, measure, PLAIN_VADV_SPACE[0])
    finally:
        fields[0].free()
    if seconds == float("inf"):
        raise ExecutionError(f"no vadv solver takes the domain {tuple(domain)}")
    cache.put(key, {"hdiff": {}, "vadv": v, "seconds": seconds, "trials": len(trials)})
    return TuneResult(key, {}, v, seconds, False, trials)
