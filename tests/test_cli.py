"""The ``sbench`` / ``sbench-analyze`` command-line contract, by behaviour.

Pinned here: command paths, options generated from Parameters, range
expansion, exit statuses (1 parameter/execution error or nothing collected,
2 verification failure), the skip flags, report modes, CSV columns and the
console-script entry points declared in pyproject.toml (the reference's
setup.py:95-116 names).  CPU only: the benchmarks driven here are numpy ones
or fakes defined below.
"""

import importlib
import types
from pathlib import Path

import pandas as pd
import pytest
import tomli
from click.testing import CliRunner

from stencil_benchmarks_amd import benchmark, cli
from stencil_benchmarks_amd.tools import validation

ROOT = Path(__file__).resolve().parents[1]


class Scripted(benchmark.Benchmark):
    """Fake benchmark: `mode` chooses what run() does."""

    mode = benchmark.Parameter("behaviour of run()", "ok",
                               choices=["ok", "fail", "invalid-output", "interrupt-late"])
    size = benchmark.Parameter("a size", 4)
    shape = benchmark.Parameter("a shape", (2, 3))
    fast = benchmark.Parameter("a flag", False)

    def setup(self):
        super().setup()
        if self.size < 0:
            raise benchmark.ParameterError("size must not be negative")
        self.calls = 0

    def run(self):
        self.calls += 1
        if self.mode == "fail":
            raise benchmark.ExecutionError("device lost")
        if self.mode == "invalid-output":
            raise validation.ValidationError("out differs in 3 points")
        if self.mode == "interrupt-late" and self.calls == 2:
            raise KeyboardInterrupt
        return {"time": 0.001 * self.size * self.calls, "label": f"s{self.size}"}


def _invoke(args, bmark=Scripted):
    group = cli.build([(["fake", "scripted"], cli._command(bmark))])
    return CliRunner().invoke(group, args, obj=types.SimpleNamespace())


def test_command_paths_follow_module_and_kebab_class_name():
    from stencil_benchmarks_amd.benchmarks_collection.stencils.mi355x import (
        horizontal_diffusion, vertical_advection)
    from stencil_benchmarks_amd.benchmarks_collection.stencils.numpy import (
        horizontal_diffusion as numpy_hdiff)
    from stencil_benchmarks_amd.benchmarks_collection.stream import mi355x

    assert cli.command_path(horizontal_diffusion.Fused) == [
        "stencils", "mi355x", "horizontal_diffusion", "fused"]
    assert cli.command_path(vertical_advection.WarpSpecialised)[-1] == "warp-specialised"
    assert cli.command_path(numpy_hdiff.OnTheFly)[-1] == "on-the-fly"
    assert cli.command_path(mi355x.Native) == ["stream", "mi355x", "native"]
    assert [cli.kebab(n) for n in ("ABc", "KMiddle", "Blocked2D", "X")] == [
        "a-bc", "k-middle", "blocked-2d", "x"]


def test_every_registered_benchmark_has_a_command():
    group = cli.command_group()
    result = CliRunner().invoke(group, ["stencils", "mi355x", "horizontal-diffusion", "--help"],
                                obj=types.SimpleNamespace())
    assert result.exit_code == 0 and "fused" in result.output
    result = CliRunner().invoke(group, ["stream", "mi355x", "native", "--help"],
                                obj=types.SimpleNamespace())
    assert result.exit_code == 0 and "--array-size" in result.output


def test_options_come_from_parameters():
    result = _invoke(["fake", "scripted", "--help"])
    assert result.exit_code == 0
    for flag in ("--mode", "--size", "--shape", "--fast / --no-fast"):
        assert flag in result.output
    assert _invoke(["fake", "scripted", "--mode", "nope"]).exit_code == 2  # click rejects


def test_ranges_expand_into_runs_and_csv_rows(tmp_path):
    path = tmp_path / "runs.csv"
    result = _invoke(["-e", "2", "-o", str(path), "fake", "scripted", "--size", "[1-4:*2]",
                      "--shape", "[s=2,3]", "[s=]"])
    assert result.exit_code == 0, result.output
    table = pd.read_csv(path, index_col=0)
    # sizes 1, 2, 4 times shapes (2, 2), (3, 3), two executions each
    assert len(table) == 12
    assert set(zip(table["shape-0"], table["shape-1"])) == {(2, 2), (3, 3)}
    assert set(table["size"]) == {1, 2, 4}
    for column in ("time", "label", "mode", "fast", "sbench-version", "benchmark-name"):
        assert column in table.columns
    assert table["benchmark-name"].iloc[0].endswith("test_cli.Scripted")


def test_report_modes():
    full = _invoke(["-e", "3", "-r", "full", "fake", "scripted", "--size", "[1,2]"])
    assert full.exit_code == 0 and full.output.count("s1") == 3
    medians = _invoke(["-e", "3", "fake", "scripted", "--size", "[1,2]"])
    assert medians.exit_code == 0
    # median of 1, 2, 3 ms x size: one row per size (the text result is a
    # group key like the parameters), sorted by time
    rows = [line.split() for line in medians.output.splitlines()
            if line.startswith(("s1", "s2"))]
    assert rows == [["s1", "1", "0.002"], ["s2", "2", "0.004"]]
    best = _invoke(["-e", "3", "-r", "best-median", "fake", "scripted", "--size", "[1,2]"])
    assert best.exit_code == 0 and "0.002" in best.output and "0.004" not in best.output


def test_exit_statuses_and_skips():
    assert _invoke(["fake", "scripted", "--size", "-1"]).exit_code == 1
    assert _invoke(["fake", "scripted", "--mode", "fail"]).exit_code == 1
    invalid = _invoke(["fake", "scripted", "--mode", "invalid-output"])
    assert invalid.exit_code == 2 and "validation error" in invalid.output
    # skipped parameter sets and failed runs: the rest still reports
    skipped = _invoke(["-s", "fake", "scripted", "--size", "[-1,3]"])
    # one parameter set left: every column but the time is constant
    assert skipped.exit_code == 0 and skipped.output.split()[-2:] == ["time", "0.006"]
    nothing = _invoke(["-q", "fake", "scripted", "--mode", "fail"])
    assert nothing.exit_code == 1 and "no data collected" in nothing.output


def test_interrupt_reports_what_ran():
    result = _invoke(["-e", "3", "-r", "full", "fake", "scripted", "--mode", "interrupt-late"])
    assert result.exit_code == 0 and result.output.count("s4") == 1


def test_numpy_benchmark_through_the_command(tmp_path):
    path = tmp_path / "hdiff.csv"
    result = CliRunner().invoke(
        cli.command_group(),
        ["-e", "2", "-o", str(path), "stencils", "numpy", "horizontal-diffusion", "on-the-fly",
         "--domain", "16", "12", "4", "--halo", "2", "2", "0"],
        obj=types.SimpleNamespace())
    assert result.exit_code == 0, result.output
    table = pd.read_csv(path, index_col=0)
    assert len(table) == 2 and (table["bandwidth"] > 0).all()
    assert list(table.loc[0, ["domain-0", "domain-1", "domain-2"]]) == [16, 12, 4]


def test_console_scripts_resolve_to_the_commands():
    project = tomli.loads((ROOT / "pyproject.toml").read_text())
    scripts = project["project"]["scripts"]
    assert set(scripts) == {"sbench", "sbench-analyze", "sbench-mi355x-collection"}
    for target in scripts.values():
        module, _, attribute = target.partition(":")
        entry = getattr(importlib.import_module(module), attribute)
        assert callable(entry)
    sbench = importlib.import_module("stencil_benchmarks_amd.scripts.sbench").main
    run = CliRunner().invoke(cli.command_group(), ["--help"], obj=types.SimpleNamespace())
    assert run.exit_code == 0 and "stencils" in run.output and "stream" in run.output
    assert sbench is cli.main


def test_sbench_main_runs_a_reference_argv():
    import subprocess
    import sys

    argv = ["stencils", "numpy", "horizontal-diffusion", "on-the-fly", "--domain", "8", "8",
            "2", "--verify"]
    done = subprocess.run([sys.executable, "-m", "stencil_benchmarks_amd.scripts.sbench",
                           "-e", "1", *argv], cwd=ROOT, capture_output=True, text=True,
                          timeout=120)
    assert done.returncode == 0, done.stdout + done.stderr
    assert "time" in done.stdout
