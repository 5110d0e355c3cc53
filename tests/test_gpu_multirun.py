"""Sweeps of the MI355X plugins and the collection script's CSV (GPU).

The sweep driver runs the real plugin classes (verification on, so every
row is also a parity check against the numpy oracle of stencils/base.py)
and the collection command writes a CSV with the reference's columns.
"""

import pandas as pd
import pytest
from click.testing import CliRunner

from stencil_benchmarks_amd.tools.multirun import Configuration, run_scaling_benchmark

pytestmark = pytest.mark.gpu


def test_sweep_over_plugins_verifies_and_reports(gpu):
    from stencil_benchmarks_amd.benchmarks_collection.stencils.mi355x import (
        horizontal_diffusion as hdiff,
        vertical_advection as vadv,
    )

    configs = [Configuration(hdiff.Fused, verify=True, halo=(3, 3, 3)),
               Configuration(vadv.WarpSpecialised, verify=True, halo=(3, 3, 3)),
               Configuration(vadv.Pipelined, verify=True, halo=(3, 3, 3))]
    table = run_scaling_benchmark(configs, 2, domain_range=[(32, 32, 40), (70, 33, 97)])
    assert len(table) == 3 * 2 * 2
    assert (table["bandwidth"] > 0).all() and (table["time"] > 0).all()
    assert set(table["name"]) == {"fused", "warp-specialised", "pipelined"}
    assert {"domain-0", "halo-2", "dtype", "benchmark-name", "sbench-version"} <= set(table.columns)


def test_stream_collection_writes_csv(gpu, tmp_path):
    from stencil_benchmarks_amd.scripts import sbench_mi355x_collection as collection

    path = tmp_path / "stream.csv"
    result = CliRunner().invoke(collection.main, ["stream-bandwidth", str(path), "-e", "1",
                                                  "-o", "ntimes=3"])
    assert result.exit_code == 0, result.output
    table = pd.read_csv(path, index_col=0)
    assert set(table["name"]) == {"copy", "scale", "add", "triad"}
    assert set(table["configuration"]) == {"stream", "stream-nt"}
    assert (table["bandwidth"] > 0).all() and "array-size" in table.columns
