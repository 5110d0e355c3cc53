"""The reference-side binding shown in INTEGRATION.md runs as written.

The code block is executed with the reference's module names
(``stencil_benchmarks.benchmark``, ``...stencils.base``) mapped onto this
package, which keeps their interface, and the library path taken from
SBMI355X_LIBRARY; the stub's HorizontalDiffusion then runs with the
reference's own verification (base.verify_stencil against numpy).
"""

import re
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def _stub_source():
    text = (ROOT / "INTEGRATION.md").read_text()
    section = text.split("## Binding the C ABI from the reference", 1)[1]
    return re.search(r"```python\n(.*?)```", section, re.S).group(1)


@pytest.fixture
def reference_names(monkeypatch):
    import stencil_benchmarks_amd
    from stencil_benchmarks_amd import benchmark, benchmarks_collection
    from stencil_benchmarks_amd.benchmarks_collection import stencils
    from stencil_benchmarks_amd.benchmarks_collection.stencils import base

    for name, module in {"stencil_benchmarks": stencil_benchmarks_amd,
                         "stencil_benchmarks.benchmark": benchmark,
                         "stencil_benchmarks.benchmarks_collection": benchmarks_collection,
                         "stencil_benchmarks.benchmarks_collection.stencils": stencils,
                         "stencil_benchmarks.benchmarks_collection.stencils.base": base}.items():
        monkeypatch.setitem(sys.modules, name, module)
    monkeypatch.setenv("SBMI355X_LIBRARY",
                       str(ROOT / "stencil_benchmarks_amd" / "libsbmi355x.so"))
    yield benchmark


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_integration_stub_runs_and_verifies(gpu, reference_names, dtype):
    namespace = {"__name__": "stencil_benchmarks.benchmarks_collection.stencils.mi355x"}
    exec(compile(_stub_source(), "INTEGRATION.md", "exec"), namespace)
    cls = namespace["HorizontalDiffusion"]
    try:
        bench = cls(domain=(70, 45, 12), dtype=dtype, data_sets=2)
        for _ in range(3):  # both data sets, verified each time
            assert bench.run()["time"] > 0
    finally:
        reference_names.REGISTRY.discard(cls)
