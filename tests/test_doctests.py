"""The docstring examples of the package's tools run and hold - CPU only.

They restate the reference's own doctests for these interfaces (range
arguments and pretty parameters, stencil_benchmarks/tools/cli.py:336-350,
393-405, 432-437, 466-477; array geometry, tools/array.py:97-121, 198-206;
parameters, benchmark.py) so the known answers are checked here as there.
"""

import doctest
import importlib

import pytest

MODULES = ["stencil_benchmarks_amd.tools.cli", "stencil_benchmarks_amd.tools.array",
           "stencil_benchmarks_amd.benchmark"]


@pytest.mark.parametrize("name", MODULES)
def test_docstring_examples(name):
    result = doctest.testmod(importlib.import_module(name), optionflags=doctest.ELLIPSIS)
    assert result.attempted > 0 and result.failed == 0, result
