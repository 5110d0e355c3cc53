"""Command-line helpers: options that take value ranges, their expansion into
runs, the per-run parameter columns and a progress line.

The option grammar and the column names are the reference's interface
(stencil_benchmarks/tools/cli.py:231-493), so command lines and CSV files
carry over unchanged:

* a plain value, or a bracketed range: ``[1,2,5]`` (a list), ``[1-16]``
  (``start-stop``, step +1) or ``[1-16:*2]`` (``start-stop:<op><step>`` with
  op one of ``+ - * /``; the values run from start while they stay inside
  ``[start, stop]``);
* ``[name=...]`` names a range.  Options that use the same name take their
  values together (zipped); different ranges multiply (cartesian product);
  ``[name=]`` uses a range whose values another option defines;
* result columns: the parameter name with ``_`` replaced by ``-``, tuple
  values split into ``<name>-0``, ``<name>-1``, ..., plus ``sbench-version``
  and ``benchmark-name`` (module path and class name).

>>> kind = range_type(int)
>>> kind.convert("[n=2-32:*2]", None, None)
ArgRange(name='n', values=(2, 4, 8, 16, 32))
>>> runs = unpack_ranges(size=(kind.convert("[s=8,16]", None, None), 4),
...                      depth=kind.convert("[s=]", None, None))
>>> [(r["size"], r["depth"]) for r in runs]
[((8, 4), 8), ((16, 4), 16)]
"""

import itertools
import re
import sys
from typing import Any, Dict, Iterable, Iterator, List, NamedTuple, Optional

import click

from .. import __version__


class ArgRange(NamedTuple):
    """A named set of values for an option; ``values`` None means the values
    are defined by another option using the same name."""

    name: str
    values: Optional[tuple]

    def unique_values(self, *others) -> tuple:
        """The one definition of this range among all its uses.

        >>> ArgRange("r", (1, 2)).unique_values(ArgRange("r", None))
        (1, 2)
        >>> ArgRange("r", None).unique_values()
        Traceback (most recent call last):
            ...
        ValueError: range 'r' is used but its values are never given
        """
        defined = {use.values for use in (self, *others) if use.values}
        if len(defined) == 1:
            return defined.pop()
        if not defined:
            raise ValueError(f"range {self.name!r} is used but its values are never given")
        raise ValueError(f"range {self.name!r} is given different values: "
                         + " vs ".join(map(str, sorted(defined, key=str))))


# ``start-stop`` with optional signs, then an optional ``:<op><step>``
_PROGRESSION = re.compile(r"([+-]?\d+)-([+-]?\d+)(?::([-+*/])([+-]?\d+))?")
_MAX_RANGE_VALUES = 1 << 16


def _progression(start, stop, op, step):
    """Values start, start op step, ... while inside [start, stop]."""
    advance = {"+": lambda v: v + step, "-": lambda v: v - step,
               "*": lambda v: v * step, "/": lambda v: v / step}[op]
    values = []
    value = start
    while start <= value <= stop:
        if len(values) == _MAX_RANGE_VALUES or (values and value == values[-1]):
            raise click.BadParameter(f"range from {start} to {stop} does not progress "
                                     f"with {op}{step}")
        values.append(value)
        value = advance(value)
    return tuple(values)


class _RangeAware(click.ParamType):
    """A click type that reads a plain value of `element`, or a range."""

    name = "value or range"
    _unnamed = itertools.count()

    def __init__(self, element: click.ParamType):
        self.element = element

    def _scalar(self, text, param, ctx):
        return self.element.convert(text, param, ctx)

    def convert(self, value, param, ctx):
        if not (isinstance(value, str) and value.startswith("[") and value.endswith("]")
                and len(value) >= 2):
            return self._scalar(value, param, ctx)
        body = value[1:-1]
        name, equals, spec = body.partition("=")
        if not equals:
            name, spec = f"~{next(self._unnamed)}", body
        if not spec:
            return ArgRange(name, None)
        if "," in spec:
            return ArgRange(name, tuple(self._scalar(v, param, ctx) for v in spec.split(",")))
        match = _PROGRESSION.fullmatch(spec)
        if match is None:
            self.fail(f"cannot read {value!r} as a range: expected [a,b,...], [start-stop] "
                      "or [start-stop:<op><step>]", param, ctx)
        start, stop = (self._scalar(match.group(g), param, ctx) for g in (1, 2))
        op = match.group(3) or "+"
        step = self._scalar(match.group(4), param, ctx) if match.group(4) else 1
        return ArgRange(name, _progression(start, stop, op, step))

    def get_metavar(self, param, *args, **kwargs):
        try:
            inner = self.element.get_metavar(param, *args, **kwargs)
        except TypeError:  # click < 8.2 takes the parameter only
            inner = self.element.get_metavar(param)
        inner = inner or self.element.name.upper()
        return f"{inner}|[RANGE]"


_BUILTIN = {int: click.INT, float: click.FLOAT, str: click.STRING}
_TYPES: Dict[Any, _RangeAware] = {}


def range_type(dtype) -> click.ParamType:
    """Click type of `dtype` (int, float, str or a click type such as
    ``click.Choice``) that also accepts the range syntax.

    >>> @click.command()
    ... @click.option("--n", type=range_type(int))
    ... def show(n):
    ...     return n
    >>> show(["--n", "3"], standalone_mode=False)
    3
    >>> show(["--n", "[1,3]"], standalone_mode=False)  # doctest: +ELLIPSIS
    ArgRange(name='~...', values=(1, 3))
    """
    if isinstance(dtype, click.ParamType):
        return _RangeAware(dtype)
    if dtype not in _TYPES:
        _TYPES[dtype] = _RangeAware(_BUILTIN[dtype])
    return _TYPES[dtype]


def _each_range(value) -> Iterator[ArgRange]:
    if isinstance(value, ArgRange):
        yield value
    elif isinstance(value, tuple):
        for item in value:
            yield from _each_range(item)


def _resolve(value, choice: Dict[str, Any]):
    if isinstance(value, ArgRange):
        return choice[value.name]
    if isinstance(value, tuple):
        return tuple(_resolve(item, choice) for item in value)
    return value


def unpack_ranges(**kwargs: Any) -> List[Dict[str, Any]]:
    """One kwargs dict per combination of range values (the range met first
    varies slowest); kwargs without ranges give a single dict.

    >>> unpack_ranges(a=3, b=5)
    [{'a': 3, 'b': 5}]
    >>> unpack_ranges(a=ArgRange("p", (1, 2)), b=ArgRange("q", (3, 4)))
    [{'a': 1, 'b': 3}, {'a': 1, 'b': 4}, {'a': 2, 'b': 3}, {'a': 2, 'b': 4}]
    """
    uses: Dict[str, List[ArgRange]] = {}
    for value in kwargs.values():
        for found in _each_range(value):
            uses.setdefault(found.name, []).append(found)
    axes = {name: found[0].unique_values(*found[1:]) for name, found in uses.items()}
    runs = []
    for combination in itertools.product(*axes.values()):
        choice = dict(zip(axes, combination))
        runs.append({key: _resolve(value, choice) for key, value in kwargs.items()})
    return runs


def range_args(**kwargs: Any) -> Iterator[str]:
    """Names of the options in `kwargs` that hold a range.

    >>> list(range_args(a=1, b=ArgRange("r", (1, 2)), c=(1, ArgRange("r", None))))
    ['b', 'c']
    """
    return (key for key, value in kwargs.items() if next(_each_range(value), None))


def pretty_parameters(bmark, include_version: bool = True,
                      include_name: bool = True) -> Dict[str, Any]:
    """The result columns describing one benchmark instance.

    >>> from stencil_benchmarks_amd.benchmark import Benchmark, Parameter
    >>> class Probe(Benchmark):
    ...     grid_size = Parameter("grid", (4, 2))
    ...     kind = Parameter("kind", "x")
    ...     def run(self):
    ...         return {}
    >>> pretty_parameters(Probe(), include_version=False, include_name=False)
    {'grid-size-0': 4, 'grid-size-1': 2, 'kind': 'x'}
    """
    columns: Dict[str, Any] = {}
    for name, value in bmark.parameters.items():
        column = name.replace("_", "-")
        if isinstance(value, tuple):
            columns.update((f"{column}-{i}", item) for i, item in enumerate(value))
        else:
            columns[column] = value
    if include_version:
        columns["sbench-version"] = __version__
    if include_name:
        cls = type(bmark)
        columns["benchmark-name"] = f"{cls.__module__}.{cls.__name__}"
    return columns


class ProgressBar:
    """Progress of nested loops on one terminal line of standard error.

    ``report(iterable)`` yields the items and advances; nested calls refine
    the fraction of the enclosing step.

    >>> bar = ProgressBar()
    >>> bar.fraction
    0.0
    """

    WIDTH = 50

    def __init__(self, stream=None):
        self._stream = stream
        self._loops: List[List[int]] = []  # [done, total] per nesting level

    @property
    def fraction(self) -> float:
        """Completed fraction of the outermost loop, in [0, 1]."""
        total, scale = 0.0, 1.0
        for done, count in self._loops:
            if not count:
                break
            total += scale * done / count
            scale /= count
        return min(total, 1.0)

    @property
    def progress(self) -> float:
        """The completed fraction in percent."""
        return 100.0 * self.fraction

    def report(self, iterable: Iterable[Any]) -> Iterator[Any]:
        items = list(iterable)
        loop = [0, len(items)]
        self._loops.append(loop)
        self._draw()
        try:
            for item in items:
                yield item
                loop[0] += 1
                self._draw()
        finally:
            self._loops.remove(loop)

    def _out(self):
        return self._stream or sys.stderr

    def _draw(self):
        filled = int(self.fraction * self.WIDTH)
        counts = " ".join(f"{done}/{count}" for done, count in self._loops)
        self._out().write(f"\r[{'#' * filled}{'.' * (self.WIDTH - filled)}] "
                          f"{self.progress:5.1f}% {counts}  ")
        self._out().flush()

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, traceback):
        # a finished bar is erased; an interrupted one is left on its line
        self._out().write("\r" + " " * (self.WIDTH + 40) + "\r" if exc_type is None else "\n")
        self._out().flush()
        return False
