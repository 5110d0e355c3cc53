"""Sweep driver, collection CSV columns and the CSV report - CPU only.

The sweep runs a stand-in benchmark (no device): one row per run with the
run's results, the configuration name and the parameters as
pretty_parameters columns (the reference's CSV layout,
stencil_benchmarks/tools/multirun.py:42-92); the report groups and pivots
such a CSV like the reference's sbench-analyze print.
"""

import pandas as pd
from click.testing import CliRunner

from stencil_benchmarks_amd.benchmark import Benchmark, Parameter
from stencil_benchmarks_amd.scripts import sbench_analyze
from stencil_benchmarks_amd.tools import multirun


class FakeStencil(Benchmark):
    domain = Parameter("domain", (8, 8, 8))
    block_size = Parameter("block", (64, 16, 4))
    dtype = Parameter("dtype", "float64")
    scale = Parameter("scale", 1.0)

    def run(self):
        n = 1
        for d in self.domain:
            n *= d
        return {"time": n * 1e-9, "bandwidth": self.scale * n}


def test_configuration_names_follow_the_reference():
    assert multirun.Configuration(FakeStencil).name == "fake-stencil"
    assert multirun._kebab("JScanShuffleSystolic") == "j-scan-shuffle-systolic"
    assert multirun._kebab("OnTheFly") == "on-the-fly"
    assert multirun.Configuration(FakeStencil, name="x").name == "x"


def test_scaling_sweep_rows_and_columns():
    configs = [multirun.Configuration(FakeStencil, name="a"),
               multirun.Configuration(FakeStencil, name="b", scale=2.0)]
    table = multirun.run_scaling_benchmark(configs, 3, domain_range=[(4, 4, 2), (8, 8, 2)])
    assert len(table) == 2 * 2 * 3
    for column in ("time", "bandwidth", "name", "domain-0", "domain-1", "domain-2",
                   "block-size-0", "dtype", "scale", "sbench-version", "benchmark-name"):
        assert column in table.columns
    assert table["benchmark-name"].iloc[0].endswith("test_multirun.FakeStencil")
    b = table[(table["name"] == "b") & (table["domain-0"] == 8)]
    assert (b["bandwidth"] == 2.0 * 128).all()


def test_domains_and_block_truncation_and_overrides():
    assert list(multirun.domains(5, 8, 40)) == [(32, 32, 40), (64, 64, 40), (128, 128, 40)]
    kw = multirun.truncate_block_size_to_domain(domain=(32, 8, 80), block_size=(64, 16, 4),
                                                storage_block_size=128)
    assert kw["block_size"] == (32, 8, 4) and kw["storage_block_size"] == 32
    override = multirun.default_kwargs(dtype="float64", verify=False)
    assert override(["dtype='float32'", "dry-runs=5"], halo=(1, 1, 1)) == {
        "dtype": "float32", "verify": False, "dry_runs": 5, "halo": (1, 1, 1)}


def test_report_groups_and_pivots_a_sweep_csv(tmp_path):
    configs = [multirun.Configuration(FakeStencil, name="a"),
               multirun.Configuration(FakeStencil, name="b", scale=2.0)]
    path = tmp_path / "sweep.csv"
    multirun.run_scaling_benchmark(configs, 2, domain_range=[(4, 4, 2), (8, 8, 2)]).to_csv(path)
    pivot = sbench_analyze.Arrangement(groups=("name", "domain-0"), unstack=True,
                                       select=("bandwidth",))
    table = pivot.apply(sbench_analyze.read_csv(path))
    assert table.loc["a", 4] == 32 and table.loc["b", 8] == 256
    assert pivot.common["dtype"] == "float64"
    result = CliRunner().invoke(sbench_analyze.main, ["print", str(path), "-g", "name", "-g",
                                                      "domain-0", "-u", "-s", "bandwidth"])
    assert result.exit_code == 0, result.output
    assert "256" in result.output
    filtered = sbench_analyze.Arrangement(query="name == 'a'").apply(
        sbench_analyze.read_csv(path))
    assert isinstance(filtered, pd.DataFrame) and len(filtered) == 4
    auto = sbench_analyze.Arrangement(auto_group=True, select=("bandwidth",))
    grouped = auto.apply(sbench_analyze.read_csv(path))
    assert grouped.loc[("a", 4, 4)] == 32 and grouped.index.names == ["name", "domain-0", "domain-1"]
    plot = tmp_path / "plot.png"
    result = CliRunner().invoke(sbench_analyze.main, ["plot", str(path), "-g", "name", "-g",
                                                      "domain-0", "-s", "bandwidth", "-r",
                                                      "peak=300", "-o", str(plot)])
    assert result.exit_code == 0, result.output
    assert plot.stat().st_size > 1000
