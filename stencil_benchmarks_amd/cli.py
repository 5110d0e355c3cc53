"""The ``sbench`` command: one sub-command per registered benchmark.

Interface kept from the reference (stencil_benchmarks/cli.py:47-245), so
existing command lines and scripts work unchanged:

* command path: the benchmark's module path below the collection package
  (``_`` shown as ``-``) and its class name in kebab case, a dash going before
  every run of capitals or digits 1-9 (``OnTheFly`` -> ``on-the-fly``), e.g.
  ``sbench stencils mi355x horizontal-diffusion fused``;
* one option per Parameter: ``--name/--no-name`` for booleans, otherwise
  ``--name`` taking ``nargs`` values in the range syntax of
  :mod:`.tools.cli`; parameters without a default are required;
* global options ``-e/--executions`` (3), ``-r/--report``
  (``all-medians`` default, ``best-median``, ``full``), ``-s`` skip
  parameter sets that raise ParameterError, ``-q`` skip runs that raise
  ExecutionError, ``-o FILE`` write every run as a CSV row;
* exit status 1 for a parameter or execution error (unless skipped) or when
  nothing ran, 2 when verification fails; Ctrl-C reports what ran so far.

Each run becomes one row of the run's results plus the parameter columns of
``pretty_parameters``.  The report groups the rows by every non-numeric
result and parameter column, shows medians, and drops columns that are the
same in every row.
"""

import gc
import sys
import types
import warnings
from typing import Dict, List, Sequence

import click

from . import benchmark
from .tools import cli as cli_tools
from .tools import validation

_CAPS = frozenset("ABCDEFGHIJKLMNOPQRSTUVWXYZ123456789")


def kebab(name: str) -> str:
    """Class name as a command name.

    >>> [kebab(n) for n in ("Fused", "OnTheFly", "KMiddle", "Native")]
    ['fused', 'on-the-fly', 'k-middle', 'native']
    """
    out = []
    for i, ch in enumerate(name):
        if i and ch in _CAPS and (i == 1 or name[i - 1] not in _CAPS):
            out.append("-")
        out.append(ch)
    return "".join(out).lower()


def command_path(bmark) -> List[str]:
    """Sub-command path of a benchmark class."""
    parts = bmark.__module__.split(".")
    if "benchmarks_collection" in parts:
        parts = parts[parts.index("benchmarks_collection") + 1:]
    else:
        parts = parts[2:]
    return parts + [kebab(bmark.__name__)]


class _Stop(Exception):
    """End the command with an exit status after printing a message."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def _numeric(table, column) -> bool:
    import pandas as pd

    return pd.api.types.is_numeric_dtype(table[column])


def summarise(rows: List[Dict], result_names: Sequence[str], mode: str) -> str:
    """The text report of `rows` (one per run) in report `mode`."""
    import pandas as pd

    table = pd.DataFrame(rows)
    if len(table) > 1:
        varying = [c for c in table.columns if table[c].nunique(dropna=False) > 1]
        if varying:
            table = table[varying]
    if mode == "full":
        return table.to_string()
    measured = [c for c in table.columns if c in set(result_names) and _numeric(table, c)]
    keys = [c for c in table.columns if c not in measured]
    if not keys:
        return table[measured].median().to_string()
    medians = table.groupby(keys, dropna=False)[measured].median()
    if "time" not in medians.columns:  # every run took the same time
        return medians.to_string()
    if mode == "best-median":
        return str(medians.loc[[medians["time"].idxmin()]])
    return medians.sort_values("time").to_string()


class _Runs:
    """Runs one benchmark class over the expanded option values."""

    def __init__(self, bmark, settings):
        self.bmark = bmark
        self.settings = settings
        self.rows: List[Dict] = []
        self.result_names: List[str] = []

    def instance(self, values):
        try:
            return self.bmark(**values)
        except benchmark.ParameterError as error:
            if self.settings.skip_invalid_parameters:
                return None
            raise _Stop(1, str(error)) from None

    def execute(self, instance):
        try:
            outcome = instance.run()
        except benchmark.ExecutionError as error:
            if self.settings.skip_execution_failures:
                return
            raise _Stop(1, f"\nexecution error: {error}") from None
        except validation.ValidationError as error:
            raise _Stop(2, f"\nvalidation error: {error}") from None
        outcome = [outcome] if isinstance(outcome, dict) else list(outcome)
        columns = cli_tools.pretty_parameters(instance)
        for result in outcome:
            self.result_names.extend(k for k in result if k not in self.result_names)
            self.rows.append({**result, **columns})

    def all(self, runs):
        with cli_tools.ProgressBar() as progress:
            for values in progress.report(runs):
                instance = self.instance(values)
                if instance is None:
                    continue
                for _ in progress.report(range(self.settings.executions)):
                    self.execute(instance)
                del instance
                gc.collect()


def _command(bmark):
    @click.pass_context
    def command(ctx, **options):
        settings = ctx.obj
        runs = _Runs(bmark, settings)
        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter("always")
            try:
                runs.all(cli_tools.unpack_ranges(**options))
            except KeyboardInterrupt:
                pass
            except _Stop as stop:
                click.echo(str(stop))
                sys.exit(stop.status)
            finally:
                for text in sorted({str(w.message) for w in caught}):
                    click.echo(f"WARNING: {text}")
        if not runs.rows:
            click.echo("no data collected")
            sys.exit(1)
        if settings.output:
            import pandas as pd

            pd.DataFrame(runs.rows).to_csv(settings.output)
        click.echo(summarise(runs.rows, runs.result_names, settings.report))

    for name, parameter in reversed(list(bmark.parameters.items())):
        command = _option(name, parameter)(command)
    return command


def _option(name, parameter):
    flag = "--" + name.replace("_", "-")
    text = parameter.description
    text = text[:1].upper() + text[1:] + "."
    if parameter.dtype is bool:
        return click.option(f"{flag}/--no-{flag[2:]}", default=parameter.default,
                            show_default=True, help=text)
    element = parameter.dtype
    if element is str and parameter.choices:
        element = click.Choice(parameter.choices)
    return click.option(flag, type=cli_tools.range_type(element), nargs=parameter.nargs,
                        required=parameter.default is None, default=parameter.default,
                        show_default=parameter.default is not None, help=text)


_REPORTS = ["best-median", "all-medians", "full"]


def _root_group() -> click.Group:
    @click.group()
    @click.option("--executions", "-e", type=int, default=3, help="Runs per parameter set.")
    @click.option("--report", "-r", type=click.Choice(_REPORTS), default="all-medians",
                  help="Report of the collected runs.")
    @click.option("--skip-invalid-parameters", "-s", is_flag=True,
                  help="Skip parameter sets that a benchmark rejects.")
    @click.option("--skip-execution-failures", "-q", is_flag=True,
                  help="Skip runs that fail.")
    @click.option("--output", "-o", type=click.Path(), help="Write every run to this CSV.")
    @click.pass_context
    def sbench(ctx, **settings):
        ctx.obj.__dict__.update(settings)

    return sbench


def build(commands) -> click.Group:
    """The ``sbench`` group holding `commands` (pairs of command path and
    callback), sub-groups made from the path prefixes."""
    root = _root_group()
    groups: Dict[tuple, click.Group] = {(): root}

    def group_of(prefix):
        if prefix not in groups:
            parent = group_of(prefix[:-1])
            groups[prefix] = parent.group(name=prefix[-1].replace("_", "-"))(
                click.pass_context(lambda ctx: None))
        return groups[prefix]

    for path, callback in commands:
        group_of(tuple(path[:-1])).command(name=path[-1])(callback)
    return root


def command_group() -> click.Group:
    """The ``sbench`` group of every registered benchmark."""
    from . import benchmarks_collection  # noqa: F401  (importing registers them)

    classes = sorted(benchmark.REGISTRY, key=lambda c: (c.__module__, c.__name__))
    return build([(command_path(c), _command(c)) for c in classes])


def main(*args, **kwargs):
    """Entry point of the ``sbench`` console script."""
    return command_group()(*args, obj=types.SimpleNamespace(), **kwargs)
