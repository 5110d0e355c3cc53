import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")
    # the CPU oracle (test infrastructure) is built here when missing, so the
    # CPU suite does not depend on a prior build step; where it already
    # exists (built by build()), it is used as is
    oracle_lib = ROOT / "oracle" / "liboracle.so"
    if not oracle_lib.exists() and shutil.which("gcc"):
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "liboracle.so"], check=False,
                       capture_output=True)


@pytest.fixture(scope="session")
def gpu():
    """Initialise HIP device 0 through the C ABI; fails loudly without it."""
    from stencil_benchmarks_amd import native

    native.init(0)
    return native


def make_fields(domain, halo, dtype, layout=(2, 1, 0), alignment=0, count=1, seed=0):
    from stencil_benchmarks_amd.tools import array

    rng = np.random.default_rng(seed)
    shape = tuple(d + 2 * h for d, h in zip(domain, halo))
    fields = []
    for _ in range(count):
        a = array.alloc_array(shape, dtype, layout, alignment, halo)
        a[...] = rng.random(shape).astype(dtype)
        fields.append(a)
    return fields


def inner(domain, halo):
    return tuple(slice(h, h + d) for d, h in zip(domain, halo))
