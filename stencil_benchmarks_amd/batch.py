"""Batched hdiff + vadv timesteps over many device-resident subdomains.

The reference runs one domain per kernel launch and has no batching; this
module drives the batched entry points of libsbmi355x.so
(``sb_batch_step``, ``sb_batch_plan_*``).  One timestep of a subdomain is

    u_next     = hdiff(u, coeff)                    (interior)
    utensstage = vadv(ustage = u_next, upos, utens, utensstage, wcon)

followed by swapping u and u_next.  The halos of u and u_next hold the same
boundary values and are never written.  All seven fields of a subdomain
share one geometry (padded like the reference's ``alloc_array``) and stay
resident in HBM across steps.
"""

import ctypes
from dataclasses import dataclass

import numpy as np

from . import native
from .device import DeviceField, geometry
from .tools import array

FIELDS = ("u0", "u1", "coeff", "upos", "utens", "utensstage", "wcon")


class SubdomainStruct(ctypes.Structure):
    _fields_ = [
        ("u", ctypes.c_void_p * 2),
        ("coeff", ctypes.c_void_p),
        ("upos", ctypes.c_void_p),
        ("utens", ctypes.c_void_p),
        ("utensstage", ctypes.c_void_p),
        ("wcon", ctypes.c_void_p),
        ("geom", native.Geom),
        ("parity", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


@dataclass
class SubdomainSpec:
    domain: tuple
    halo: tuple = (3, 3, 3)
    steps: int = 1
    seed: int = 0


class Subdomain:
    """Seven device fields of one subdomain plus its step parity."""

    def __init__(self, domain, halo=(3, 3, 3), dtype="float64", alignment=64, seed=0,
                 layout=(2, 1, 0), host_fields=None, fill=True, arena=None):
        self.domain = tuple(int(d) for d in domain)
        self.halo = tuple(int(h) for h in halo)
        self.dtype = np.dtype(dtype)
        # the geometry of alloc_array (layout, innermost padding), no host array
        self.shape = tuple(d + 2 * h for d, h in zip(self.domain, self.halo))
        self.byte_strides, _ = array.compute_strides(self.shape, self.dtype.itemsize,
                                                     tuple(layout), alignment)
        self.strides = tuple(s // self.dtype.itemsize for s in self.byte_strides)
        self.geom = geometry(self.domain, self.strides, self.halo)
        # all seven fields in one allocation (a block of `arena` if given)
        self.fields = dict(zip(FIELDS, DeviceField.group(
            self.shape, self.dtype, self.byte_strides, len(FIELDS), self.halo, arena)))
        nx, ny, nz = self.domain
        self.points = nx * ny * nz
        s = self.dtype.itemsize
        # HBM bytes one step must move: hdiff (inp + halo, coeff, out) + vadv
        self._step_bytes = (((nx + 4) * (ny + 4) * nz + 2 * nx * ny * nz) * s
                            + (5 * nx * ny * nz + (nx + 1) * ny * nz) * s)
        self._packed = {}  # parity -> bytes of the SubdomainStruct
        self.parity = 0
        self.steps_done = 0
        if host_fields is not None:
            self.upload(host_fields)
        elif fill:
            self.fill_random(seed)

    # -- initialisation ------------------------------------------------------
    def fill_random(self, seed):
        """U[0,1) on the device (counter hash of (seed * 64 + field index,
        element)), u1 = u0 (the shared boundary): one asynchronous launch."""
        code = native.SB_F64 if self.dtype == np.float64 else native.SB_F32
        n = self.fields["u0"].span // self.dtype.itemsize
        index = {name: (0 if name == "u1" else position) for position, name in enumerate(FIELDS)}
        bases = (ctypes.c_void_p * len(FIELDS))(*[self.fields[name].base for name in FIELDS])
        seeds = (ctypes.c_uint64 * len(FIELDS))(*[seed * 64 + index[name] for name in FIELDS])
        native.call("sb_fill_random_fields", code, bases, len(FIELDS), n, seeds)

    def upload(self, host_fields):
        """Upload host fields (dict with u, coeff, upos, utens, utensstage, wcon)."""
        for name in FIELDS:
            source = host_fields["u" if name in ("u0", "u1") else name]
            self.fields[name].upload(self._conform(source))

    def _conform(self, host):
        if host.strides == self.byte_strides and host.dtype == self.dtype:
            return host
        out = self.empty_host()
        out[...] = host
        return out

    def empty_host(self):
        buf = np.zeros(self.fields["u0"].span, dtype=np.uint8)
        return np.ndarray(self.shape, self.dtype, buffer=buf, strides=self.byte_strides)

    def download(self, name):
        host = self.empty_host()
        self.fields[name].download(host)
        return host

    def state(self):
        """Current u (after the steps run so far)."""
        return self.download(f"u{self.parity}")

    def struct(self):
        f = self.fields
        s = SubdomainStruct()
        s.u[0] = f["u0"].pointer()
        s.u[1] = f["u1"].pointer()
        s.coeff = f["coeff"].pointer()
        s.upos = f["upos"].pointer()
        s.utens = f["utens"].pointer()
        s.utensstage = f["utensstage"].pointer()
        s.wcon = f["wcon"].pointer()
        s.geom = self.geom
        s.parity = self.parity
        return s

    def packed(self):
        """The SubdomainStruct as bytes (cached per parity: the fields and
        geometry never change)."""
        blob = self._packed.get(self.parity)
        if blob is None:
            blob = self._packed[self.parity] = bytes(self.struct())
        return blob

    def advance(self, steps=1):
        self.parity = (self.parity + steps) & 1
        self.steps_done += steps

    def compulsory_bytes_per_step(self):
        """HBM bytes one step must move: hdiff (inp + halo, coeff, out) + vadv."""
        return self._step_bytes

    def free(self):
        for field in self.fields.values():
            field.free()


def _array(subs):
    n = max(1, len(subs))
    blob = b"".join(sub.packed() for sub in subs)
    return (SubdomainStruct * n).from_buffer_copy(blob.ljust(n * ctypes.sizeof(SubdomainStruct),
                                                            b"\0"))


def precision_of(subs, compute):
    dtype = subs[0].dtype if subs else np.dtype("float64")
    return native.precision_code(dtype, compute or dtype)


def step(subs, compute=None, hdiff_options=None, vadv_options=None, timed=False):
    """Advance every subdomain by one step (one batched launch pair)."""
    if not subs:
        return 0.0
    arr = _array(subs)
    seconds = ctypes.c_double(0.0)
    native.call(
        "sb_batch_step", precision_of(subs, compute), arr, len(subs),
        hdiff_options or native.launch_options(), vadv_options or native.launch_options(),
        ctypes.byref(seconds) if timed else None,
    )
    for sub in subs:
        sub.advance(1)
    return seconds.value


class Plan:
    """A fixed batch advanced `steps` steps per run, optionally as a hipGraph."""

    def __init__(self, subs, steps, compute=None, hdiff_options=None, vadv_options=None):
        self.subs = list(subs)
        self.steps = int(steps)
        handle = ctypes.c_void_p()
        arr = _array(self.subs)
        native.call(
            "sb_batch_plan_create", precision_of(self.subs, compute), arr, len(self.subs),
            self.steps, hdiff_options or native.launch_options(),
            vadv_options or native.launch_options(), ctypes.byref(handle),
        )
        self.handle = handle

    def run(self, use_graph=True, timed=True):
        seconds = ctypes.c_double(0.0)
        native.call("sb_batch_plan_run", self.handle, int(bool(use_graph)),
                    ctypes.byref(seconds) if timed else None)
        for sub in self.subs:
            sub.advance(self.steps)
        return seconds.value

    def close(self):
        if self.handle:
            native.call("sb_batch_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_specs(count=96, nmin=32, nmax=512, nzmin=40, nzmax=160, seed=0, halo=(3, 3, 3)):
    """The batched workload: nx, ny ~ U{nmin..nmax}, nz ~ U{nzmin..nzmax}."""
    rng = np.random.default_rng(seed)
    specs = []
    for index in range(count):
        nx, ny = (int(v) for v in rng.integers(nmin, nmax + 1, size=2))
        nz = int(rng.integers(nzmin, nzmax + 1))
        specs.append(SubdomainSpec((nx, ny, nz), halo, seed=seed * 100003 + index))
    return specs
