"""Stencil semantics: fields, halo, layout, verification and bandwidth.

Restates stencils/base.py:45-501 of the reference:

* :class:`Stencil` parameters domain, data_sets, halo, dtype, layout,
  alignment, huge_pages, offset_allocations and verify (``base.py:45-64``),
  plus ``seed``: the reference fills fields from unseeded generators
  (tools/parallel.cpp:51-68); here every field of every data set draws from
  ``numpy.random.default_rng((seed, data_set, field))`` so runs repeat;
* ``setup`` validates and allocates every argument for every data set
  (``base.py:66-85``); arguments listed in ``outputs`` start at zero,
  every other one is U[0,1) over its halo-inclusive extent;
* ``run`` cycles the data sets, snapshots the inputs when verifying, calls
  ``run_stencil``, verifies and adds ``bandwidth = data_size / time / 1e9``
  (``base.py:151-166``);
* the ``verify_stencil`` oracles keep the reference's operation order, so
  an FP64 kernel that follows the same order matches them bit for bit.

Subclasses may set ``oracle_dtype`` to verify in a wider type than the
storage type: the oracle then runs on the stored (already rounded) inputs
widened to that type and its result is rounded once to the storage type.
"""

import abc
import collections
import copy

import numpy as np

from ...benchmark import Benchmark, ExecutionError, Parameter, ParameterError
from ...tools import array, fill, validation

DTR_STAGE = 3 / 20
BET_M = 0.5 * (1 - 0)
BET_P = 0.5 * (1 + 0)


def _names_numpy_type(name):
    try:
        np.dtype(name)
    except TypeError:
        return False
    return True


class Stencil(Benchmark):
    domain = Parameter("domain size", (128, 128, 80))
    data_sets = Parameter(
        "number of data sets, if bigger than one, data sets are cycled before "
        "each execution to start with cold cache",
        1,
    )
    halo = Parameter("halo size", (3, 3, 3))
    dtype = Parameter("data type in NumPy format, e.g. float32 or float64", "float64")
    layout = Parameter("data layout, 2 means innermost dimension, 0 outermost", (2, 1, 0))
    alignment = Parameter("data alignment in bytes", 0)
    huge_pages = Parameter("use huge pages", "none", choices=["none", "transparent", "explicit"])
    offset_allocations = Parameter(
        "offset allocated data by some bytes to minimize cache conflicts", False
    )
    verify = Parameter("enable verification", True)
    seed = Parameter("seed of the uniform [0, 1) field values", 0)

    oracle_dtype = None

    # (test, complaint) pairs over the parameter values; setup raises
    # ParameterError with the complaint of the first failing test
    _CONFIGURATION_RULES = (
        (lambda s: _names_numpy_type(s.dtype), "dtype {s.dtype!r} is not a NumPy data type"),
        (lambda s: min(s.halo) >= 0, "halo widths must not be negative: {s.halo}"),
        (lambda s: min(s.domain) >= 0, "domain extents must not be negative: {s.domain}"),
        (lambda s: sorted(s.layout) == [0, 1, 2],
         "layout {s.layout} is not a permutation of (0, 1, 2)"),
        (lambda s: s.data_sets >= 1, "at least one data set is needed, got {s.data_sets}"),
        (lambda s: s.alignment >= 0, "alignment must not be negative: {s.alignment} bytes"),
        (lambda s: s.alignment % s.dtype_size == 0,
         "alignment of {s.alignment} bytes is not a multiple of the {s.dtype_size}-byte "
         "element size"),
    )

    def setup(self):
        super().setup()
        for test, complaint in self._CONFIGURATION_RULES:
            if not test(self):
                raise ParameterError(complaint.format(s=self))
        self._data = [self._new_data_set(number) for number in range(self.data_sets)]
        self._runs_started = 0

    def _new_data_set(self, number):
        """Every argument of data set `number`: outputs zero, the other
        fields U[0,1) from the generator seeded (seed, number, argument)."""
        record = collections.namedtuple("StencilData", self.args)
        fields = []
        for position, name in enumerate(self.args):
            field = self.empty_field()
            if name in self.outputs:
                field.fill(0)
            else:
                fill.random_fill(field, fill.field_rng(self.seed, number, position))
            fields.append(field)
        return record(*fields)

    @property
    def data_set_index(self):
        """The data set of the current (or next) run; runs cycle through the
        sets."""
        return self._runs_started % self.data_sets

    # ---- allocation / geometry -------------------------------------------
    def alloc_field(self, domain_with_halo, layout, index_to_align):
        if self.huge_pages == "none":
            allocate = array.alloc_smallpages
        else:
            transparent = self.huge_pages == "transparent"

            def allocate(nbytes):
                return array.alloc_hugepages(nbytes, transparent)

        return array.alloc_array(
            domain_with_halo,
            self.dtype,
            layout,
            self.alignment,
            index_to_align=index_to_align,
            alloc=allocate,
            apply_offset=self.offset_allocations,
        )

    def empty_field(self):
        return self.alloc_field(self.domain_with_halo, self.layout, self.halo)

    def random_field(self, rng=None):
        field = self.empty_field()
        return fill.random_fill(field, rng or np.random.default_rng())

    @property
    def outputs(self):
        return ()

    @property
    def dtype_size(self):
        return np.dtype(self.dtype).itemsize

    @property
    def domain_with_halo(self):
        return tuple(int(n) for n in np.add(self.domain, np.multiply(2, self.halo)))

    @property
    def strides(self):
        first_field = self._data[0][0]
        return tuple(s // self.dtype_size for s in first_field.strides)

    @property
    def data_size(self):
        return len(self.args) * int(np.prod(self.domain)) * self.dtype_size

    def inner_slice(self, shift=None, expand=None):
        """Slices of the interior, moved by `shift` and widened by `expand`
        (an int for every axis, or one value per axis) on both sides."""
        rank = len(self.domain)
        shift = np.zeros(rank, dtype=int) if shift is None else np.asarray(shift, dtype=int)
        grow = np.broadcast_to(np.asarray(0 if expand is None else expand, dtype=int), (rank,))
        first = np.asarray(self.halo) + shift - grow
        last = first + np.asarray(self.domain) + 2 * grow
        return tuple(slice(int(a), int(b)) for a, b in zip(first, last))

    # ---- lifecycle ------------------------------------------------------------
    @abc.abstractmethod
    def run_stencil(self, data):
        pass

    @abc.abstractmethod
    def verify_stencil(self, data_before, data_after):
        pass

    @property
    @abc.abstractmethod
    def args(self):
        pass

    def run(self):
        """One execution on the next data set: the kernel, its verification
        against the oracle on a snapshot of the inputs, and the reference
        bandwidth figure (data_size / time, GB/s)."""
        data = self._data[self.data_set_index]
        snapshot = copy.deepcopy(data) if self.verify else None
        result = self.run_stencil(data)
        self._runs_started += 1
        if snapshot is not None:
            self.verify_stencil(snapshot, data)
        if result.get("time", 0) <= 0:
            raise ExecutionError(f"{type(self).__name__}: run reported no positive time")
        if "bandwidth" in result:
            raise ExecutionError(f"{type(self).__name__}: run_stencil must leave "
                                 "'bandwidth' to the base class")
        return dict(result, bandwidth=self.data_size / result["time"] / 1e9)

    # ---- oracle helpers ---------------------------------------------------
    def _oracle_input(self, field):
        """Field as the oracle sees it (widened when oracle_dtype is set)."""
        if self.oracle_dtype is None:
            return field
        return field.astype(self.oracle_dtype)

    def _stored(self, values, like):
        return np.asarray(values).astype(like.dtype, copy=False)


class BasicStencil(Stencil):
    @property
    def args(self):
        return "inp", "out"

    @property
    def outputs(self):
        return ("out",)


class EmptyStencil(BasicStencil):
    def verify_stencil(self, data_before, data_after):
        pass


class CopyStencil(BasicStencil):
    def verify_stencil(self, data_before, data_after):
        validation.check_equality("inp", data_before.inp, data_after.inp)
        inner = self.inner_slice()
        validation.check_equality("out", data_after.out[inner], data_before.inp[inner])


def _unit(axis):
    shift = np.zeros(3, dtype=int)
    shift[axis] = 1
    return shift


class OnesidedAverageStencil(BasicStencil):
    axis = Parameter("axis along which to average", 0, choices=[0, 1, 2])

    def setup(self):
        super().setup()
        if self.halo[self.axis] < 1:
            raise ParameterError(f"positive halo size required along axis {self.axis}")

    def expected(self, inp):
        inp = self._oracle_input(inp)
        shifted = inp[self.inner_slice(_unit(self.axis))]
        return (shifted + inp[self.inner_slice()]) / 2

    def verify_stencil(self, data_before, data_after):
        validation.check_equality("inp", data_before.inp, data_after.inp)
        expected = self._stored(self.expected(data_before.inp), data_after.out)
        validation.check_equality("out", data_after.out[self.inner_slice()], expected)


class SymmetricAverageStencil(BasicStencil):
    axis = Parameter("axis along which to average", 0, choices=[0, 1, 2])

    def setup(self):
        super().setup()
        if self.halo[self.axis] < 1:
            raise ParameterError(f"positive halo size required along axis {self.axis}")

    def expected(self, inp):
        inp = self._oracle_input(inp)
        up = _unit(self.axis)
        return (inp[self.inner_slice(up)] + inp[self.inner_slice(-up)]) / 2

    def verify_stencil(self, data_before, data_after):
        validation.check_equality("inp", data_before.inp, data_after.inp)
        expected = self._stored(self.expected(data_before.inp), data_after.out)
        validation.check_equality("out", data_after.out[self.inner_slice()], expected)


class LaplacianStencil(BasicStencil):
    along_x = Parameter("include x-axis in Laplacian", True)
    along_y = Parameter("include y-axis in Laplacian", True)
    along_z = Parameter("include z-axis in Laplacian", False)

    @property
    def along_axes(self):
        return (self.along_x, self.along_y, self.along_z)

    def setup(self):
        super().setup()
        short = [axis for axis, (h, on) in enumerate(zip(self.halo, self.along_axes)) if on and h < 1]
        if short:
            raise ParameterError(f"the Laplacian needs a halo of 1 along axes {short}, "
                                 f"got halo {self.halo}")

    def expected(self, inp):
        inp = self._oracle_input(inp)
        centre = inp[self.inner_slice()]
        result = np.zeros(self.domain, dtype=inp.dtype)
        for axis, enabled in enumerate(self.along_axes):
            if enabled:
                up = _unit(axis)
                result += (2 * centre - inp[self.inner_slice(up)]) - inp[self.inner_slice(-up)]
        return result

    def verify_stencil(self, data_before, data_after):
        validation.check_equality("inp", data_before.inp, data_after.inp)
        expected = self._stored(self.expected(data_before.inp), data_after.out)
        validation.check_equality("out", data_after.out[self.inner_slice()], expected)


def horizontal_diffusion(inp, coeff):
    """hdiff over the whole array (base.py:284-307); border rows stay 0."""
    lap = np.zeros_like(inp)
    c = inp[1:-1, 1:-1, :]
    lap[1:-1, 1:-1, :] = 4 * c - (
        inp[2:, 1:-1, :] + inp[:-2, 1:-1, :] + inp[1:-1, 2:, :] + inp[1:-1, :-2, :]
    )
    flx = np.zeros_like(inp)
    flux = lap[1:, :, :] - lap[:-1, :, :]
    flx[:-1, :, :] = np.where(flux * (inp[1:, :, :] - inp[:-1, :, :]) > 0, 0, flux)
    fly = np.zeros_like(inp)
    flux = lap[:, 1:, :] - lap[:, :-1, :]
    fly[:, :-1, :] = np.where(flux * (inp[:, 1:, :] - inp[:, :-1, :]) > 0, 0, flux)
    out = np.zeros_like(inp)
    out[1:-1, 1:-1, :] = c - coeff[1:-1, 1:-1, :] * (
        flx[1:-1, 1:-1, :] - flx[:-2, 1:-1, :] + fly[1:-1, 1:-1, :] - fly[1:-1, :-2, :]
    )
    return out


class HorizontalDiffusionStencil(Stencil):
    def setup(self):
        super().setup()
        if min(self.halo[:2]) < 2:
            raise ParameterError(f"horizontal diffusion reads two points beyond the domain in i "
                                 f"and j: halo {self.halo} is too narrow")

    @property
    def args(self):
        return "inp", "coeff", "out"

    @property
    def outputs(self):
        return ("out",)

    @property
    def data_size(self):
        n = int(np.prod(self.domain))
        return (2 * n + int(np.prod(np.array(self.domain) + 4))) * self.dtype_size

    def expected(self, inp, coeff):
        return horizontal_diffusion(self._oracle_input(inp), self._oracle_input(coeff))[
            self.inner_slice()
        ]

    def verify_stencil(self, data_before, data_after):
        validation.check_equality("inp", data_before.inp, data_after.inp)
        validation.check_equality("coeff", data_before.coeff, data_after.coeff)
        expected = self._stored(self.expected(data_before.inp, data_before.coeff), data_after.out)
        validation.check_equality("out", data_after.out[self.inner_slice()], expected)


def vertical_advection(fields, domain, halo, shift):
    """Thomas solve of one component (base.py:415-473); returns tensstage.

    `fields` maps stage, pos, tens, tensstage and wcon to full (halo
    inclusive) arrays; `shift` is the (i, j) offset of the wcon neighbour.
    The returned array has the shape of the interior.
    """
    nx, ny, nz = domain
    hi, hj, hk = halo

    def plane(name, k, di=0, dj=0):
        return fields[name][hi + di : hi + di + nx, hj + dj : hj + dj + ny, hk + k]

    si, sj = shift
    ccol = [None] * nz
    dcol = [None] * nz
    for k in range(nz):
        pos, tens, ts = plane("pos", k), plane("tens", k), plane("tensstage", k)
        stage_k = plane("stage", k)
        if k == 0:
            gcv = 0.25 * (plane("wcon", k + 1, si, sj) + plane("wcon", k + 1))
            cs = gcv * BET_M
            c = gcv * BET_P
            b = DTR_STAGE - c
            correction = -cs * (plane("stage", k + 1) - stage_k)
            d = DTR_STAGE * pos + tens + ts + correction
            ccol[k] = c / b
            dcol[k] = d / b
        elif k < nz - 1:
            gav = -0.25 * (plane("wcon", k, si, sj) + plane("wcon", k))
            gcv = 0.25 * (plane("wcon", k + 1, si, sj) + plane("wcon", k + 1))
            as_, cs = gav * BET_M, gcv * BET_M
            a = gav * BET_P
            c = gcv * BET_P
            b = DTR_STAGE - a - c
            correction = -as_ * (plane("stage", k - 1) - stage_k) - cs * (
                plane("stage", k + 1) - stage_k
            )
            d = DTR_STAGE * pos + tens + ts + correction
            r = 1.0 / (b - ccol[k - 1] * a)
            ccol[k] = c * r
            dcol[k] = (d - dcol[k - 1] * a) * r
        else:
            gav = -0.25 * (plane("wcon", k, si, sj) + plane("wcon", k))
            as_ = gav * BET_M
            a = gav * BET_P
            b = DTR_STAGE - a
            correction = -as_ * (plane("stage", k - 1) - stage_k)
            d = DTR_STAGE * pos + tens + ts + correction
            dcol[k] = (d - dcol[k - 1] * a) / (b - ccol[k - 1] * a)
    result = np.empty((nx, ny, nz), dtype=fields["stage"].dtype)
    x = None
    for k in range(nz - 1, -1, -1):
        x = dcol[k] if k == nz - 1 else dcol[k] - ccol[k] * x
        result[:, :, k] = DTR_STAGE * (x - plane("pos", k))
    return result


class VerticalAdvectionStencil(Stencil):
    all_components = Parameter(
        "advect all velocity components (like in the COSMO dycore) "
        "instead of the u component (like in the GridTools benchmark)",
        False,
    )

    COMPONENTS = (("u", (1, 0)), ("v", (0, 1)), ("w", (0, 0)))

    def setup(self):
        if self.domain[2] < 2:
            raise ParameterError("vertical advection needs at least 2 vertical levels")
        super().setup()
        if self.halo[0] < 1 or (self.all_components and any(h < 1 for h in self.halo)):
            raise ParameterError(f"vertical advection reads wcon one point beyond the domain: "
                                 f"halo {self.halo} is too narrow")

    @property
    def components(self):
        return self.COMPONENTS if self.all_components else self.COMPONENTS[:1]

    @property
    def args(self):
        names = []
        for comp, _ in self.components:
            names += [f"{comp}stage", f"{comp}pos", f"{comp}tens", f"{comp}tensstage"]
        return tuple(names) + ("wcon", "ccol", "dcol", "datacol")

    @property
    def data_size(self):
        reads, writes = (15, 5) if self.all_components else (7, 3)
        return (reads + writes) * int(np.prod(self.domain)) * self.dtype_size

    def expected(self, data_before):
        """Expected tensstage of every component, interior shape."""
        out = {}
        for comp, shift in self.components:
            fields = {
                "stage": self._oracle_input(getattr(data_before, f"{comp}stage")),
                "pos": self._oracle_input(getattr(data_before, f"{comp}pos")),
                "tens": self._oracle_input(getattr(data_before, f"{comp}tens")),
                "tensstage": self._oracle_input(getattr(data_before, f"{comp}tensstage")),
                "wcon": self._oracle_input(data_before.wcon),
            }
            out[comp] = vertical_advection(fields, self.domain, self.halo, shift)
        return out

    def verify_stencil(self, data_before, data_after):
        for comp, _ in self.components:
            for name in ("stage", "pos", "tens"):
                field = f"{comp}{name}"
                validation.check_equality(
                    field, getattr(data_before, field), getattr(data_after, field)
                )
        validation.check_equality("wcon", data_before.wcon, data_after.wcon)
        expected = self.expected(data_before)
        for comp, _ in self.components:
            result = getattr(data_after, f"{comp}tensstage")
            validation.check_equality(
                f"{comp}tensstage",
                result[self.inner_slice()],
                self._stored(expected[comp], result),
            )
