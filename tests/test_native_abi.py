"""The C ABI library without a GPU - CPU only.

libsbmi355x.so loads, exports every entry point include/sbmi355x.h declares,
the ctypes binding covers them all, and the product fails loudly (no CPU
fallback) when the library is missing.  No entry point that touches a
device is called.
"""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "sbmi355x.h"
LIBRARY = ROOT / "stencil_benchmarks_amd" / "libsbmi355x.so"


def declared_functions():
    text = HEADER.read_text()
    return re.findall(r"^\s*(?:int|const char \*)\s*(sb_[a-z0-9_]+)\s*\(", text, re.M)


@pytest.fixture(scope="module")
def library():
    if not LIBRARY.exists():  # build once (hipcc cross-compiles gfx950 without a GPU)
        subprocess.run(["make", "-C", str(ROOT / "stencil_benchmarks_amd" / "csrc")], check=True,
                       capture_output=True)
    return ctypes.CDLL(str(LIBRARY))


def test_header_declares_the_entry_points():
    names = declared_functions()
    assert len(names) == len(set(names))
    for required in ("sb_hdiff", "sb_vadv", "sb_laplacian", "sb_copy", "sb_stream_kernel",
                     "sb_batch_step", "sb_batch_plan_create", "sb_last_error"):
        assert required in names


def test_library_exports_every_declared_function(library):
    missing = [name for name in declared_functions() if not hasattr(library, name)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_every_declared_function():
    from stencil_benchmarks_amd import native

    bound = set(native.SIGNATURES) | {"sb_last_error"}
    assert set(declared_functions()) == bound


def test_abi_version_and_error_string_without_a_device(library):
    from stencil_benchmarks_amd import native

    lib = native.load(LIBRARY)
    version = int(re.search(r"#define SB_ABI_VERSION (\d+)", HEADER.read_text()).group(1))
    assert lib.sb_abi_version() == version
    assert isinstance(lib.sb_last_error(), (bytes, type(None)))


def test_missing_library_fails_loudly(tmp_path):
    from stencil_benchmarks_amd import native

    with pytest.raises(native.NativeLibraryMissing):
        native.load(tmp_path / "libsbmi355x.so")


def test_event_entry_points_reject_null_without_a_device(library):
    # argument checks come before any device call
    from stencil_benchmarks_amd import native

    lib = native.load(LIBRARY)
    assert lib.sb_event_record(None) != 0 and b"NULL" in lib.sb_last_error()
    assert lib.sb_event_synchronize(None) != 0 and b"NULL" in lib.sb_last_error()
    assert lib.sb_event_create(None) != 0 and b"NULL" in lib.sb_last_error()
    assert lib.sb_event_destroy(None) == 0
