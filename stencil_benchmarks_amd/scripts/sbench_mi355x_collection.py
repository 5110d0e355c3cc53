"""Bandwidth sweeps of the MI355X stencils, one CSV per collection.

The MI355X counterpart of the reference's per-GPU collection scripts
(stencil_benchmarks/scripts/sbench_mi100_collection.py:56-157): each command
runs a fixed set of configurations over square domains 32^2 .. 1024^2 x 80
(`multirun.domains`) and writes every run as a row of a CSV whose columns
are the run's results plus the benchmark parameters as `pretty_parameters`
names, which the reference's `sbench-analyze` (and `sbench_analyze` here)
reads.  `-o name=value` overrides a parameter of every configuration (e.g.
`-o dtype="'float32'"`), `-e` sets the runs per configuration and domain.

    python -m stencil_benchmarks_amd.scripts.sbench_mi355x_collection \\
        vertical-advection-bandwidth vadv.csv -e 11
"""

import click
import pandas as pd

from ..benchmarks_collection.stencils.mi355x import basic
from ..benchmarks_collection.stencils.mi355x import horizontal_diffusion as hdiff
from ..benchmarks_collection.stencils.mi355x import vertical_advection as vadv
from ..benchmarks_collection.stream import mi355x as stream
from ..tools import cli
from ..tools.multirun import Configuration, default_kwargs, run_scaling_benchmark

# Defaults of every stencil configuration: no host verification (the
# kernels are verified by the test suite), warm-up launches, the
# reference's alignment of its GPU collections, FP64.
common_kwargs = default_kwargs(verify=False, dry_runs=10, alignment=512, dtype="float64")


@click.group()
def main():
    """MI355X bandwidth collections (CSV output)."""


def _options(function):
    function = click.option("--option", "-o", multiple=True,
                            help="parameter override name=value (Python literal)")(function)
    function = click.option("--executions", "-e", type=int, default=101,
                            help="runs per configuration and domain")(function)
    return click.argument("output", type=click.Path())(function)


@main.command()
@_options
def basic_bandwidth(output, executions, option):
    """Copy, averages along each axis and the horizontal Laplacian."""
    kwargs = common_kwargs(option, halo=(1, 1, 1))
    configurations = [
        Configuration(basic.Empty, name="empty", **kwargs),
        Configuration(basic.Copy, name="copy", **kwargs),
        Configuration(basic.OnesidedAverage, name="avg-i", axis=0, **kwargs),
        Configuration(basic.OnesidedAverage, name="avg-j", axis=1, **kwargs),
        Configuration(basic.OnesidedAverage, name="avg-k", axis=2, **kwargs),
        Configuration(basic.SymmetricAverage, name="sym-avg-i", axis=0, **kwargs),
        Configuration(basic.SymmetricAverage, name="sym-avg-j", axis=1, **kwargs),
        Configuration(basic.SymmetricAverage, name="sym-avg-k", axis=2, **kwargs),
        Configuration(basic.Laplacian, name="lap-ij", along_x=True, along_y=True,
                      along_z=False, **kwargs),
    ]
    run_scaling_benchmark(configurations, executions).to_csv(output)


@main.command()
@_options
def horizontal_diffusion_bandwidth(output, executions, option):
    """The LDS-tiled kernels at each tile height and the on-the-fly kernel."""
    kwargs = common_kwargs(option)
    configurations = [
        Configuration(hdiff.Fused, name="fused", **kwargs),
        Configuration(hdiff.LdsDma, name="lds-dma-64x16", block_size=(64, 16, 0), **kwargs),
        Configuration(hdiff.Fused, name="fused-64x16", block_size=(64, 16, 0), **kwargs),
        Configuration(hdiff.Fused, name="fused-64x8", block_size=(64, 8, 0), **kwargs),
        Configuration(hdiff.Fused, name="fused-64x32", block_size=(64, 32, 0), **kwargs),
        Configuration(hdiff.OnTheFly, name="on-the-fly", **kwargs),
    ]
    run_scaling_benchmark(configurations, executions).to_csv(output)


@main.command()
@_options
def vertical_advection_bandwidth(output, executions, option):
    """Every vertical-advection solver (the default picks per nz)."""
    kwargs = common_kwargs(option)
    configurations = [
        Configuration(vadv.Fused, name="fused", **kwargs),
        Configuration(vadv.WarpSpecialised, **kwargs),
        Configuration(vadv.Pipelined, **kwargs),
        Configuration(vadv.LdsColumn, **kwargs),
        Configuration(vadv.GlobalScratch, **kwargs),
    ]
    run_scaling_benchmark(configurations, executions).to_csv(output)


@main.command()
@_options
def stream_bandwidth(output, executions, option):
    """GPU STREAM at 2^20 to 2^29 elements per array (cache-resident to
    HBM-bound), with and without non-temporal loads and stores; one row per
    STREAM kernel and run."""
    kwargs = default_kwargs(dtype="float64")(option)
    rows = []
    with cli.ProgressBar() as progress:
        for exponent in progress.report(range(20, 30, 3)):
            for name, nt in progress.report((("stream", False), ("stream-nt", True))):
                bmark = stream.Native(array_size=2**exponent, streaming_loads=nt,
                                      streaming_stores=nt, **kwargs)
                for _ in progress.report(range(executions)):
                    for result in bmark.run():
                        rows.append({**result, "configuration": name,
                                     **cli.pretty_parameters(bmark)})
    pd.DataFrame(rows).to_csv(output)


if __name__ == "__main__":
    main()
