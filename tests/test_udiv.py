"""The multiply-shift division and tile decode that the one-plane-per-
workgroup kernels and the batch item decode use, checked on the host:
`udiv_check.cpp` compiles against the library's own header
(stencil_benchmarks_amd/csrc/sb_internal.h) and compares with the hardware
division and with xcd_remap followed by plain division."""

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not available")
def test_udiv_and_tile_decode(tmp_path):
    exe = tmp_path / "udiv_check"
    subprocess.run([HIPCC, "-std=c++17", "-O2", "-x", "hip", "--offload-arch=gfx950",
                    str(HERE / "udiv_check.cpp"), "-o", str(exe)], check=True,
                   capture_output=True, timeout=300)
    done = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert done.returncode == 0, done.stdout + done.stderr
    assert done.stdout.startswith("ok:"), done.stdout
