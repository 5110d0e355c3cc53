"""Benchmark core: typed parameters, class registry and benchmark lifecycle.

Behaviour follows stencil_benchmarks/benchmark.py:39-220 of the reference:

* a :class:`Parameter` carries a description, a default, a value type, an
  argument count and optional choices; a missing default makes it required
  (``benchmark.py:50-120``);
* :class:`BenchmarkMeta` gathers the Parameters declared on a class and its
  bases into ``cls.parameters`` and registers every concrete class in
  :data:`REGISTRY` (``benchmark.py:140-164``);
* :class:`Benchmark` validates its keyword arguments, stores the values in
  ``self.parameters``, exposes them as attributes (re-validated on
  assignment) and calls :meth:`Benchmark.setup` (``benchmark.py:167-220``).
"""

import abc
import inspect
from typing import Any, Dict, Iterable, List, Optional, Type, Union

REGISTRY = set()


class ParameterError(ValueError):
    """Invalid benchmark configuration (CLI: skipped with ``-s``)."""


class ExecutionError(RuntimeError):
    """Failure while running a benchmark (CLI: skipped with ``-q``)."""


def _infer_type_and_count(default):
    if isinstance(default, (tuple, list)):
        if len(default) == 0:
            raise ValueError("can not use empty tuple as default")
        kinds = {type(v) for v in default}
        if len(kinds) != 1:
            raise ValueError("different types in default tuple")
        return kinds.pop(), len(default)
    return type(default), 1


class Parameter:
    """A typed benchmark parameter.

    >>> p = Parameter("domain size", (128, 128, 80))
    >>> p.dtype, p.nargs
    (<class 'int'>, 3)
    >>> p.validate(None)
    (128, 128, 80)
    >>> Parameter("required", dtype=str, nargs=1).validate(None)
    Traceback (most recent call last):
        ...
    stencil_benchmarks_amd.benchmark.ParameterError: value is required
    """

    def __init__(
        self,
        description: str,
        default: Any = None,
        dtype: Optional[Type] = None,
        nargs: Optional[int] = None,
        choices: Optional[Iterable] = None,
    ):
        if default is None:
            if dtype is None or nargs is None:
                raise ValueError("dtype and nargs must be given if default is None")
        else:
            inferred_type, inferred_nargs = _infer_type_and_count(default)
            if dtype is None:
                dtype = inferred_type
            elif dtype is not inferred_type:
                raise ValueError("iconsistent default and dtype values")
            if nargs is None:
                nargs = inferred_nargs
            elif nargs != inferred_nargs:
                raise ValueError("inconsistent default and nargs values")
        self.description = description
        self.default = default
        self.dtype = dtype
        self.nargs = nargs
        self.choices = None if choices is None else tuple(choices)

    def validate(self, value):
        """Return `value` (or the default when None) after type checks."""
        if value is None:
            if self.default is None:
                raise ParameterError("value is required")
            value = self.default
        if self.nargs == 1:
            if not isinstance(value, self.dtype):
                raise ParameterError(
                    f'wrong type of argument "{value}", '
                    f' expected one of type "{self.dtype.__name__}"'
                )
        else:
            if not isinstance(value, (tuple, list)):
                raise ParameterError(
                    f"{self.nargs} arguments of type "
                    f'"{self.dtype.__name__}" required, found "{value}"'
                )
            if len(value) != self.nargs:
                raise ParameterError(f'wrong number of arguments in argument "{value}"')
            if not all(isinstance(v, self.dtype) for v in value):
                raise ParameterError(f'wrong type in argument "{value}"')
        if self.choices is not None and value not in self.choices:
            listed = ", ".join(f'"{c}"' for c in self.choices)
            raise ParameterError(
                f'unsupported argument value "{value}", choices are {listed}'
            )
        return value

    def __repr__(self):
        return (
            f"{type(self).__name__}(description={self.description!r}, "
            f"dtype={self.dtype!r}, nargs={self.nargs!r}, default={self.default!r})"
        )

    def __eq__(self, other):
        return (
            isinstance(other, Parameter)
            and self.description == other.description
            and self.dtype is other.dtype
            and self.nargs == other.nargs
            and self.default == other.default
        )

    __hash__ = None


class BenchmarkMeta(abc.ABCMeta):
    """Collects class-level Parameters and registers concrete benchmarks."""

    def __new__(mcs, name, bases, namespace):
        if "parameters" in namespace:
            raise AttributeError(
                "Benchmark classes must not define an attribute `parameters`"
            )
        collected = {}
        for base in bases:
            collected.update(getattr(base, "parameters", {}))
        attributes = {}
        for key, value in namespace.items():
            if isinstance(value, Parameter):
                collected[key] = value
            else:
                attributes[key] = value
        attributes["parameters"] = collected
        cls = super().__new__(mcs, name, bases, attributes)
        if not inspect.isabstract(cls):
            REGISTRY.add(cls)
        return cls


class Benchmark(metaclass=BenchmarkMeta):
    """Base class of all benchmarks; see the module docstring."""

    def __init__(self, **kwargs):
        unknown = set(kwargs) - set(self.parameters)
        if unknown:
            raise ValueError(
                "unsupported arguments " + ", ".join(f'"{a}"' for a in unknown)
            )
        values = {}
        for name, param in self.parameters.items():
            try:
                values[name] = param.validate(kwargs.get(name))
            except ParameterError as error:
                raise ParameterError(
                    f'invalid value for argument "{name}": {error.args[0]}'
                ) from None
        # instance attribute shadows the class-level Parameter dict
        object.__setattr__(self, "parameters", values)
        self.setup()

    def __setattr__(self, name, value):
        if name in self.parameters:
            self.parameters[name] = type(self).parameters[name].validate(value)
        else:
            object.__setattr__(self, name, value)

    def __getattr__(self, name):
        # only reached when normal lookup fails, i.e. for parameter values
        params = self.__dict__.get("parameters")
        if params is not None and name in params:
            return params[name]
        raise AttributeError(name)

    def setup(self) -> None:
        """Prepare the benchmark; raise ParameterError for bad configurations."""

    @abc.abstractmethod
    def run(self) -> Union[Dict[str, Any], List[Dict[str, Any]]]:
        """Run once and return a result dict (or a list of them)."""

    def __call__(self):
        return self.run()

    def __repr__(self):
        args = ", ".join(f"{k}={v!r}" for k, v in self.parameters.items())
        return f"{type(self).__name__}({args})"
