"""Benchmark core API (mirrors stencil_benchmarks/test/test_benchmark.py)."""

import pytest

from stencil_benchmarks_amd import benchmark


class SomeBenchmark(benchmark.Benchmark):
    parameter1 = benchmark.Parameter("some parameter", False)
    parameter2 = benchmark.Parameter("some other parameter", 1)

    def setup(self):
        self.somevalue = self.parameter2 // 2

    def run(self):
        return self.somevalue


class DerivedBenchmark(SomeBenchmark):
    parameter3 = benchmark.Parameter("third parameter", "0")

    def setup(self):
        super().setup()
        self.paramsum = sum(int(p) for p in self.parameters.values())

    def run(self):
        return self.paramsum


def test_registration():
    assert SomeBenchmark in benchmark.REGISTRY
    assert DerivedBenchmark in benchmark.REGISTRY


def test_params():
    assert SomeBenchmark.parameters == {
        "parameter1": benchmark.Parameter("some parameter", False),
        "parameter2": benchmark.Parameter("some other parameter", 1),
    }


def test_derived_params():
    assert DerivedBenchmark.parameters == {
        "parameter1": benchmark.Parameter("some parameter", False, bool),
        "parameter2": benchmark.Parameter("some other parameter", 1, int),
        "parameter3": benchmark.Parameter("third parameter", "0", str),
    }


def test_init():
    bmark = SomeBenchmark(parameter1=True, parameter2=42)
    assert bmark.parameter1 is True and bmark.parameter2 == 42 and bmark.somevalue == 21
    assert bmark.parameters == {"parameter1": True, "parameter2": 42}
    with pytest.raises(benchmark.ParameterError):
        SomeBenchmark(parameter1=3, parameter2="42")


def test_derived_init():
    bmark = DerivedBenchmark(parameter1=True, parameter2=42, parameter3="5")
    assert bmark.parameters == {"parameter1": True, "parameter2": 42, "parameter3": "5"}
    assert bmark.somevalue == 21 and bmark.paramsum == 48


def test_run():
    assert SomeBenchmark(parameter1=False, parameter2=42).run() == 21


def test_reassign():
    bmark = DerivedBenchmark(parameter1=True, parameter2=42, parameter3="5")
    bmark.parameter1 = False
    assert bmark.parameter1 is False and bmark.parameters["parameter1"] is False
    with pytest.raises(benchmark.ParameterError):
        bmark.parameter1 = 3


def test_unknown_argument_and_required():
    class Required(benchmark.Benchmark):
        needed = benchmark.Parameter("needed", dtype=str, nargs=1)

        def run(self):
            return None

    with pytest.raises(ValueError):
        SomeBenchmark(unknown=1)
    with pytest.raises(benchmark.ParameterError, match="required"):
        Required()
    assert Required(needed="x").needed == "x"


def test_parameter_validation():
    p = benchmark.Parameter("choice", "a", choices=["a", "b"])
    assert p.validate("b") == "b"
    with pytest.raises(benchmark.ParameterError):
        p.validate("c")
    t = benchmark.Parameter("tuple", (1, 2, 3))
    with pytest.raises(benchmark.ParameterError):
        t.validate((1, 2))
    with pytest.raises(benchmark.ParameterError):
        t.validate((1, 2, 3.0))
    with pytest.raises(ValueError):
        benchmark.Parameter("mixed", (1, 2.0))
    with pytest.raises(ValueError):
        benchmark.Parameter("no default")


def test_parameters_attribute_forbidden():
    with pytest.raises(AttributeError):

        class Bad(benchmark.Benchmark):
            parameters = {}

            def run(self):
                return None
