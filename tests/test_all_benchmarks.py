"""Every registered benchmark runs with its default parameters.

The counterpart of the reference's collection test
(stencil_benchmarks/test/benchmarks_collection/test_benchmarks_collection.py:
39-55): instantiate each benchmark in the registry with defaults and run it
(the stencils verify their output against the numpy oracles of
stencils/base.py by default).  On the CPU the benchmarks that need no GPU
run (numpy backend, host STREAM); with the GPU every benchmark runs.
"""

import pytest

from stencil_benchmarks_amd import benchmark
from stencil_benchmarks_amd import benchmarks_collection  # noqa: F401  (registers)


def _name(bmark):
    return bmark.__module__ + "." + bmark.__name__


def _needs_gpu(bmark):
    return ".mi355x" in bmark.__module__


CPU_BENCHMARKS = sorted((b for b in benchmark.REGISTRY if not _needs_gpu(b)), key=_name)
ALL_BENCHMARKS = sorted(benchmark.REGISTRY, key=_name)


def _run_defaults(bmark):
    result = bmark().run()
    rows = result if isinstance(result, list) else [result]
    assert rows
    for row in rows:
        if "time" in row:
            assert row["time"] >= 0


@pytest.mark.parametrize("bmark", CPU_BENCHMARKS, ids=_name)
def test_cpu_benchmark_runs_with_defaults(bmark):
    _run_defaults(bmark)


@pytest.mark.gpu
@pytest.mark.parametrize("bmark", ALL_BENCHMARKS, ids=_name)
def test_every_benchmark_runs_with_defaults(gpu, bmark):
    _run_defaults(bmark)


def test_registry_covers_the_mi355x_backends():
    names = {_name(b) for b in benchmark.REGISTRY}
    for expected in ("stencils.mi355x.horizontal_diffusion.Fused",
                     "stencils.mi355x.vertical_advection.WarpSpecialised",
                     "stencils.mi355x.basic.Laplacian", "stream.mi355x.Native",
                     "stream.mc_calpin.Native", "stencils.numpy.horizontal_diffusion.OnTheFly"):
        assert any(n.endswith(expected) for n in names), expected


@pytest.mark.gpu
def test_cli_runs_a_gpu_benchmark_and_writes_csv(gpu, tmp_path):
    import types

    import pandas as pd
    from click.testing import CliRunner

    from stencil_benchmarks_amd import cli

    path = tmp_path / "hdiff.csv"
    result = CliRunner().invoke(
        cli.command_group(),
        ["-e", "3", "-o", str(path), "stencils", "mi355x", "horizontal-diffusion", "fused",
         "--domain", "64", "64", "16",
         "--dtype", "[float32,float64]"],
        obj=types.SimpleNamespace())
    assert result.exit_code == 0, result.output
    table = pd.read_csv(path, index_col=0)
    assert len(table) == 6 and set(table["dtype"]) == {"float32", "float64"}
    assert (table["bandwidth"] > 0).all() and "benchmark-name" in table.columns


def test_warp_specialised_levels_per_step_are_checked():
    from stencil_benchmarks_amd.benchmarks_collection.stencils.mi355x import vertical_advection

    with pytest.raises(benchmark.ParameterError):
        vertical_advection.WarpSpecialised(levels_per_step=5)


# dry runs advance the read-write utensstage, so they run without verification
@pytest.mark.gpu
@pytest.mark.parametrize("timers,dry_runs,verify", [("gpu", 0, True), ("wall", 0, True),
                                                    ("wall", 2, False)])
def test_warp_specialised_plugin_runs_8_levels_per_step(gpu, timers, dry_runs, verify):
    from stencil_benchmarks_amd.benchmarks_collection.stencils.mi355x import vertical_advection

    bmark = vertical_advection.WarpSpecialised(levels_per_step=8, timers=timers,
                                              dry_runs=dry_runs, domain=(96, 40, 37),
                                              verify=verify)
    assert bmark.launch_options().unroll == 8
    assert bmark.run()["time"] > 0
