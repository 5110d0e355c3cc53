"""Click helpers: range-valued options, range unpacking, result columns.

Range grammar of stencil_benchmarks/tools/cli.py:231-416:

* ``[a-b]`` / ``[a-b:op s]`` with op one of ``+ - * /`` (default ``+1``);
* ``[x,y,z]`` explicit values;
* ``[name=...]`` names a range; ranges sharing a name are zipped, while
  distinct ranges form a cartesian product; ``[name=]`` reuses the values
  defined elsewhere under that name.

Result columns follow ``pretty_parameters`` (``cli.py:448-493``): ``_`` in
names becomes ``-``, tuple values expand to ``name-0``, ``name-1``, ...,
plus ``sbench-version`` and ``benchmark-name``.

>>> unpack_ranges(a=3, b=ArgRange(name='r0', values=(1, 2, 3)))
[{'a': 3, 'b': 1}, {'a': 3, 'b': 2}, {'a': 3, 'b': 3}]
>>> unpack_ranges(a=ArgRange(name='r0', values=(1, 2)),
...               b=ArgRange(name='r0', values=None))
[{'a': 1, 'b': 1}, {'a': 2, 'b': 2}]
"""

import itertools
import operator
import re
import sys
import types
from typing import Any, Dict, Iterator, List, NamedTuple, Optional

import click

from .. import __version__


def colorize(string, rgb=None, hsv=None):
    """Wrap `string` in a 24-bit terminal colour escape.

    >>> colorize('foo', rgb=(1, 0, 0)) # doctest: +ELLIPSIS
    '...foo...'
    """
    if hsv:
        hue, sat, val = hsv

        def channel(n):
            k = (n + hue * 6) % 6
            return val - val * sat * max(min(k, 4 - k, 1), 0)

        rgb = channel(5), channel(3), channel(1)
    if not rgb:
        raise ValueError("missing color specification")
    red, green, blue = (int(255 * c) for c in rgb)
    return f"\x1b[38;2;{red};{green};{blue}m{string}\x1b[0m"


class ProgressBar:
    """Nested terminal progress bar.

    >>> ProgressBar().progress
    0.0
    """

    def __init__(self):
        self._levels = []

    def report(self, iterable):
        items = list(iterable)
        depth = len(self._levels)
        self._levels.append(types.SimpleNamespace(current=0, max=len(items)))
        self._print()
        try:
            for item in items:
                yield item
                self._levels[depth].current += 1
                self._print()
        finally:
            self._levels.pop()

    @property
    def progress(self) -> float:
        percent, width = 0.0, 100.0
        for level in self._levels:
            if level.max:
                percent += level.current / level.max * width
                width /= level.max
        return percent

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        if exc_type is None:
            click.echo("\r" + " " * 110 + "\r", nl=False)
        else:
            click.echo()
        return False

    def _print(self):
        percent = round(self.progress)
        bar = colorize("█" * percent + "-" * (100 - percent), hsv=(percent / 300, 1, 1))
        click.echo("\r|" + bar + f"| {percent:3}%", nl=False)
        sys.stdout.flush()


class ArgRange(NamedTuple):
    """A named range of option values (``values`` None: defined elsewhere)."""

    name: str
    values: Optional[tuple]

    def unique_values(self, *others) -> tuple:
        """Single definition of this range's values among its occurrences.

        >>> ArgRange('foo', (1, 2)).unique_values(ArgRange('foo', None))
        (1, 2)
        """
        definitions = {r.values for r in (self,) + others if r.values}
        if not definitions:
            raise ValueError(f'missing value definition for argument range "{self.name}"')
        if len(definitions) > 1:
            raise ValueError(f'multiple definitions for argument range "{self.name}"')
        return definitions.pop()


_RANGE_RE = re.compile(
    r"^(?P<start>[+-]?\d+)-(?P<stop>[+-]?\d+)(:(?P<op>[*/+-])(?P<step>[+-]?\d+))?$"
)
_OPS = {"+": operator.add, "-": operator.sub, "*": operator.mul, "/": operator.truediv}


class _MaybeRange(click.ParamType):
    """Click type accepting a plain value or a bracketed range."""

    name = "maybe range"

    def __init__(self, base_type):
        self.base_type = base_type
        self._anonymous = 0

    def _interval(self, text, param, ctx):
        match = _RANGE_RE.match(text)
        if not match:
            self.fail(f'could not parse range "{text}"')
        start = self.base_type.convert(match.group("start"), param, ctx)
        stop = self.base_type.convert(match.group("stop"), param, ctx)
        step_op = _OPS[match.group("op")] if match.group("op") else operator.add
        step = (
            self.base_type.convert(match.group("step"), param, ctx)
            if match.group("step")
            else 1
        )
        values, current = [], start
        while start <= current <= stop:
            values.append(current)
            current = step_op(current, step)
        return tuple(values)

    def convert(self, value, param, ctx):
        if isinstance(value, str) and len(value) >= 2 and value[0] == "[" and value[-1] == "]":
            body = value[1:-1]
            if "=" in body:
                name, body = body.split("=", 1)
            else:
                name = f"{id(self)}_{self._anonymous}"
                self._anonymous += 1
            if not body:
                values = None
            elif "," in body:
                values = tuple(
                    self.base_type.convert(v, param, ctx) for v in body.split(",")
                )
            else:
                values = self._interval(body, param, ctx)
            return ArgRange(name, values)
        return self.base_type.convert(value, param, ctx)

    def get_metavar(self, param, *args, **kwargs):
        try:
            base = self.base_type.get_metavar(param, *args, **kwargs)
        except TypeError:
            base = self.base_type.get_metavar(param)
        base = base or self.base_type.name
        return f"{base} [or {base} range]"


_RANGE_TYPES = {
    int: _MaybeRange(click.INT),
    float: _MaybeRange(click.FLOAT),
    str: _MaybeRange(click.STRING),
}


def range_type(dtype):
    """Click parameter type for `dtype` that also accepts ranges.

    >>> @click.command()
    ... @click.argument('arg', type=range_type(int))
    ... @click.option('--option', type=range_type(int))
    ... def foo(arg, option):
    ...     return arg, option
    >>> foo(args=['1', '--option', '3'], standalone_mode=False)
    (1, 3)
    >>> foo(args=['[x=1-3]', '--option', '[3,4]'], standalone_mode=False)
    (ArgRange(name='x', values=(1, 2, 3)), ArgRange(name='...', values=(3, 4)))
    """
    if isinstance(dtype, click.ParamType):
        return _MaybeRange(dtype)
    return _RANGE_TYPES[dtype]


def _ranges_in(values) -> Iterator[ArgRange]:
    for value in values:
        if isinstance(value, ArgRange):
            yield value
        elif isinstance(value, tuple):
            yield from _ranges_in(value)


def _substitute(value, chosen):
    if isinstance(value, ArgRange):
        return chosen[value.name]
    if isinstance(value, tuple):
        return tuple(_substitute(v, chosen) for v in value)
    return value


def unpack_ranges(**kwargs: Any) -> List[Dict[str, Any]]:
    """Expand all ArgRanges in `kwargs` into a list of plain kwargs dicts."""
    by_name = {}
    for rng in _ranges_in(kwargs.values()):
        by_name.setdefault(rng.name, []).append(rng)
    names = sorted(by_name)
    domains = [ArgRange.unique_values(*by_name[n]) for n in names]
    return [
        {key: _substitute(val, dict(zip(names, combo))) for key, val in kwargs.items()}
        for combo in itertools.product(*domains)
    ]


def range_args(**kwargs: Any) -> Iterator[str]:
    """Names of the arguments in `kwargs` that hold ranges.

    >>> list(range_args(a=1, b=ArgRange(name='r0', values=(1, 2))))
    ['b']
    """
    for key, value in kwargs.items():
        if isinstance(value, ArgRange) or (
            isinstance(value, tuple) and any(isinstance(v, ArgRange) for v in value)
        ):
            yield key


def pretty_parameters(bmark, include_version=True, include_name=True) -> Dict[str, Any]:
    """Flattened, dash-named parameter dict of a benchmark instance.

    >>> from stencil_benchmarks_amd.benchmark import Benchmark, Parameter
    >>> class BMark(Benchmark):
    ...     foo_bar = Parameter('foo param', dtype=int, nargs=1)
    ...     baz = Parameter('bar param', default=(1, 2))
    ...     def run(self):
    ...         pass
    >>> pretty_parameters(BMark(foo_bar=42), include_version=False,
    ...                   include_name=False)
    {'foo-bar': 42, 'baz-0': 1, 'baz-1': 2}
    """
    columns = {}
    for name, value in bmark.parameters.items():
        key = name.replace("_", "-")
        if isinstance(value, tuple):
            for position, item in enumerate(value):
                columns[f"{key}-{position}"] = item
        else:
            columns[key] = value
    if include_version:
        columns["sbench-version"] = __version__
    if include_name:
        columns["benchmark-name"] = type(bmark).__module__ + "." + type(bmark).__name__
    return columns
