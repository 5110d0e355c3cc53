"""Benchmark core: declared parameters, the class registry, one benchmark's
lifecycle.

The public contract is that of the reference's ``stencil_benchmarks.benchmark``
module (stencil_benchmarks/benchmark.py:39-220), which every benchmark plugin
and the ``sbench`` command line build on:

* ``Parameter(description, default=None, dtype=None, nargs=None,
  choices=None)``: a typed option; without a default it is required, and
  then ``dtype`` and ``nargs`` must be given; a tuple default declares a
  fixed-length tuple of one element type;
* ``Cls.parameters`` is the dict name -> Parameter of a benchmark class (its
  own declarations after those of its bases, a later base overriding an
  earlier one); ``obj.parameters`` is the dict name -> value of an instance;
* every concrete (non-abstract) benchmark class lands in :data:`REGISTRY`
  when it is defined; a class may not define an attribute ``parameters``;
* ``Cls(**values)`` rejects unknown names (``ValueError``), checks each value
  (``ParameterError``), then calls :meth:`Benchmark.setup`; parameter values
  read and assign as attributes, assignments being checked again;
* ``ParameterError`` (a ``ValueError``) marks an invalid configuration,
  ``ExecutionError`` (a ``RuntimeError``) a failed run.

The design here: a :class:`Parameter` is a data descriptor that stays on the
class (the values live in one dict per instance), and the metaclass only
records, after the class exists, which declarations are in effect.
"""

import abc
import inspect
from typing import Any, Dict, Iterable, List, Optional, Tuple, Type, Union

REGISTRY = set()


class ParameterError(ValueError):
    """The parameter values do not describe a runnable configuration."""


class ExecutionError(RuntimeError):
    """A benchmark run failed."""


def _element_signature(default) -> Tuple[type, int]:
    """(element type, count) that a non-None default implies."""
    if not isinstance(default, (tuple, list)):
        return type(default), 1
    kinds = {type(item) for item in default}
    if not kinds:
        raise ValueError("a tuple default needs at least one element")
    if len(kinds) != 1:
        raise ValueError(f"tuple default {default!r} mixes element types")
    return kinds.pop(), len(default)


class Parameter:
    """One declared benchmark parameter.

    >>> size = Parameter("domain size", (128, 128, 80))
    >>> size.dtype.__name__, size.nargs
    ('int', 3)
    >>> size.validate(None), size.validate([4, 4, 4])
    ((128, 128, 80), [4, 4, 4])
    >>> Parameter("needed", dtype=str, nargs=1).validate(None)
    Traceback (most recent call last):
        ...
    stencil_benchmarks_amd.benchmark.ParameterError: a value is required (no default)
    >>> Parameter("kind", "a", choices="ab").validate("c")
    Traceback (most recent call last):
        ...
    stencil_benchmarks_amd.benchmark.ParameterError: 'c' is not one of 'a', 'b'
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
                raise ValueError(f"parameter {description!r}: a required parameter "
                                 "(no default) must state dtype and nargs")
        else:
            implied_type, implied_count = _element_signature(default)
            if dtype is not None and dtype is not implied_type:
                raise ValueError(f"parameter {description!r}: default {default!r} is not "
                                 f"of dtype {dtype.__name__}")
            if nargs is not None and nargs != implied_count:
                raise ValueError(f"parameter {description!r}: default {default!r} does not "
                                 f"hold nargs={nargs} values")
            dtype, nargs = implied_type, implied_count
        self.description = description
        self.default = default
        self.dtype = dtype
        self.nargs = nargs
        self.choices = tuple(choices) if choices is not None else None
        self.name = None  # attribute name, set when the owning class is created

    # ---- checking -----------------------------------------------------------
    def _complaint(self, value) -> Optional[str]:
        """Why `value` is not acceptable, or None."""
        kind = self.dtype.__name__
        if self.nargs == 1:
            elements = (value,)
        elif not isinstance(value, (tuple, list)):
            return f"expected {self.nargs} values of type {kind}, got {value!r}"
        elif len(value) != self.nargs:
            return f"expected {self.nargs} values, got {len(value)} in {value!r}"
        else:
            elements = value
        for element in elements:
            if not isinstance(element, self.dtype):
                return f"{element!r} is not of type {kind}"
        if self.choices is not None and value not in self.choices:
            return f"{value!r} is not one of " + ", ".join(repr(c) for c in self.choices)
        return None

    def validate(self, value):
        """`value`, or the default for None, once it passes the checks."""
        if value is None:
            if self.default is None:
                raise ParameterError("a value is required (no default)")
            value = self.default
        complaint = self._complaint(value)
        if complaint:
            raise ParameterError(complaint)
        return value

    # ---- descriptor: the value of this parameter on an instance ---------------
    def __set_name__(self, owner, name):
        self.name = name

    def __get__(self, instance, owner=None):
        if instance is None:
            return self
        try:
            return instance.__dict__["_parameter_values"][self.name]
        except KeyError:
            raise AttributeError(self.name) from None

    def __set__(self, instance, value):
        # the declaration in effect may come from another base than this one
        declared = type(instance).parameters[self.name]
        instance.__dict__["_parameter_values"][self.name] = declared.validate(value)

    # ---- identity -------------------------------------------------------------
    def _key(self):
        return (self.description, self.dtype, self.nargs, self.default)

    def __eq__(self, other):
        if not isinstance(other, Parameter):
            return NotImplemented
        return self._key() == other._key()

    __hash__ = None

    def __repr__(self):
        return (f"{type(self).__name__}(description={self.description!r}, "
                f"dtype={self.dtype!r}, nargs={self.nargs!r}, default={self.default!r})")


class _ParameterTable:
    """``Cls.parameters``: the declarations; ``obj.parameters``: the values."""

    def __get__(self, instance, owner=None):
        if instance is None:
            return owner.__dict__["_declared_parameters"]
        return instance.__dict__["_parameter_values"]


class BenchmarkMeta(abc.ABCMeta):
    """Metaclass of benchmarks: resolves the declarations in effect for each
    new class and registers the concrete ones."""

    def __init__(cls, name, bases, namespace, **kwargs):
        super().__init__(name, bases, namespace, **kwargs)
        is_root = not any(isinstance(base, BenchmarkMeta) for base in bases)
        if not is_root and "parameters" in namespace:
            raise AttributeError(f"benchmark class {name} may not define `parameters`; "
                                 "declare Parameter attributes instead")
        declared: Dict[str, Parameter] = {}
        for base in bases:
            declared.update(base.__dict__.get("_declared_parameters", {}) if isinstance(
                base, BenchmarkMeta) else {})
        declared.update((key, value) for key, value in namespace.items()
                        if isinstance(value, Parameter))
        cls._declared_parameters = declared
        if not inspect.isabstract(cls):
            REGISTRY.add(cls)


class Benchmark(metaclass=BenchmarkMeta):
    """Base class of every benchmark.  Subclasses declare Parameters as class
    attributes, override :meth:`setup` and implement :meth:`run`."""

    parameters = _ParameterTable()

    def __init__(self, **kwargs):
        declared = type(self).parameters
        unknown = sorted(set(kwargs) - set(declared))
        if unknown:
            raise ValueError(f"{type(self).__name__} has no parameter " + ", ".join(unknown))
        values = {}
        for name, parameter in declared.items():
            try:
                values[name] = parameter.validate(kwargs.get(name))
            except ParameterError as error:
                raise ParameterError(f"parameter {name}: {error}") from None
        self.__dict__["_parameter_values"] = values
        self.setup()

    def __getattr__(self, name):
        # a declaration shadowed by a plain attribute of a more derived base
        # is still readable by name
        values = self.__dict__.get("_parameter_values")
        if values is not None and name in values:
            return values[name]
        raise AttributeError(f"{type(self).__name__} object has no attribute {name!r}")

    def setup(self) -> None:
        """Prepare to run; raise ParameterError for unsupported settings."""

    @abc.abstractmethod
    def run(self) -> Union[Dict[str, Any], List[Dict[str, Any]]]:
        """One run: a result dict, or a list of them."""

    def __call__(self):
        return self.run()

    def __repr__(self):
        settings = ", ".join(f"{name}={value!r}" for name, value in self.parameters.items())
        return f"{type(self).__name__}({settings})"
