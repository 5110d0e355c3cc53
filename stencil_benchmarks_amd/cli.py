"""``sbench`` command line: one click command per registered benchmark.

Behaviour of stencil_benchmarks/cli.py:47-245 of the reference:

* the command path is the benchmark's module path below the package's
  ``benchmarks_collection`` plus its class name in kebab case (``OnTheFly``
  -> ``on-the-fly``; the digits 1-9 count as capitals), e.g.
  ``sbench stencils mi355x horizontal-diffusion fused`` (``cli.py:47-50``);
* every Parameter becomes an option (``--x/--no-x`` for bools; range
  syntax for the rest; no default means required) (``cli.py:119-144``);
* global options ``-e/--executions`` (3), ``-r/--report``
  (all-medians | best-median | full), ``-s`` skip invalid parameter sets,
  ``-q`` skip execution failures, ``-o`` CSV output (``cli.py:191-210``);
* exit status 1 on ParameterError / ExecutionError, 2 on ValidationError
  (``cli.py:53-73``); Ctrl-C keeps the results collected so far.
"""

import gc
import re
import sys
import types
import warnings

import click

from . import benchmark
from .tools import cli as cli_tools
from .tools import validation

_KEBAB = re.compile("(?!^)([A-Z1-9]+)")


def command_path(bmark):
    parts = bmark.__module__.split(".")
    if "benchmarks_collection" in parts:
        parts = parts[parts.index("benchmarks_collection") + 1 :]
    else:
        parts = parts[2:]
    return parts + [_KEBAB.sub(r"-\1", bmark.__name__).lower()]


def _instantiate(bmark, skip_invalid, **kwargs):
    try:
        return bmark(**kwargs)
    except benchmark.ParameterError as error:
        if skip_invalid:
            return None
        click.echo(*error.args)
        sys.exit(1)


def _run_instance(instance, skip_failures):
    try:
        results = instance.run()
    except benchmark.ExecutionError as error:
        if skip_failures:
            return None
        click.echo("\nexecution error: " + " ".join(str(a) for a in error.args).strip())
        sys.exit(1)
    except validation.ValidationError as error:
        click.echo("\nvalidation error: " + " ".join(str(a) for a in error.args).strip())
        sys.exit(2)
    if isinstance(results, dict):
        results = [results]
    keys = set(results[0].keys())
    columns = cli_tools.pretty_parameters(instance)
    return [dict(r, **columns) for r in results], keys


def _report(results, result_keys, output, report):
    if not results:
        click.echo("no data collected")
        sys.exit(1)
    import pandas as pd

    table = pd.DataFrame(results)
    if output:
        table.to_csv(output)
    numeric = {k for k in result_keys if pd.api.types.is_numeric_dtype(table.dtypes[k])}
    unique = table.apply(pd.Series.nunique)
    if any(unique > 1):
        table = table.drop(columns=unique[unique <= 1].index)
    if report == "full":
        click.echo(table.to_string())
        return
    group_keys = [c for c in table.columns if c not in numeric]
    if group_keys:
        medians = table.groupby(group_keys).median(numeric_only=True)
        if report == "best-median":
            click.echo(medians.loc[[medians["time"].idxmin()]])
        else:
            click.echo(medians.sort_values(by="time").to_string())
    else:
        click.echo(table.median(numeric_only=True).to_string())


def _options(bmark):
    for name, param in bmark.parameters.items():
        flag = "--" + name.replace("_", "-")
        text = param.description[0].upper() + param.description[1:] + "."
        if param.dtype is bool:
            yield click.option(
                flag + "/" + flag.replace("--", "--no-"),
                default=param.default,
                help=text,
                show_default=True,
            )
        else:
            base = (
                click.Choice(param.choices)
                if param.dtype is str and param.choices
                else param.dtype
            )
            yield click.option(
                flag,
                type=cli_tools.range_type(base),
                nargs=param.nargs,
                help=text,
                required=param.default is None,
                default=param.default,
                show_default=param.default is not None,
            )


def _command(bmark):
    @click.pass_context
    def command(ctx, **kwargs):
        settings = ctx.obj
        expanded = cli_tools.unpack_ranges(**kwargs)
        results, result_keys = [], set()
        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter("always")
            try:
                with cli_tools.ProgressBar() as progress:
                    for kws in progress.report(expanded):
                        instance = _instantiate(bmark, settings.skip_invalid_parameters, **kws)
                        if instance is None:
                            continue
                        for _ in progress.report(range(settings.executions)):
                            outcome = _run_instance(instance, settings.skip_execution_failures)
                            if outcome is None:
                                continue
                            rows, result_keys = outcome
                            results += rows
                        del instance
                        gc.collect()
            except KeyboardInterrupt:
                pass
            for message in sorted({str(w.message) for w in caught}):
                print("WARNING:", message)
        _report(results, result_keys, settings.output, settings.report)

    for option in _options(bmark):
        command = option(command)
    return command


@click.group()
@click.option("--executions", "-e", type=int, default=3)
@click.option(
    "--report",
    "-r",
    default="all-medians",
    type=click.Choice(["best-median", "all-medians", "full"]),
)
@click.option("--skip-invalid-parameters", "-s", is_flag=True)
@click.option("--skip-execution-failures", "-q", is_flag=True)
@click.option("--output", "-o", type=click.Path())
@click.pass_context
def _root(ctx, executions, report, skip_invalid_parameters, skip_execution_failures, output):
    ctx.obj.executions = executions
    ctx.obj.report = report
    ctx.obj.skip_invalid_parameters = skip_invalid_parameters
    ctx.obj.skip_execution_failures = skip_execution_failures
    ctx.obj.output = output


def build(commands):
    """Click group holding `commands` ([(path, callback), ...]) as a tree."""
    tree = {}
    for path, func in commands:
        node = tree
        for part in path[:-1]:
            node = node.setdefault(part, {})
        node[path[-1]] = func

    @click.pass_context
    def passthrough(_ctx):
        pass

    def attach(group, node):
        for name, child in node.items():
            if isinstance(child, dict):
                attach(group.group(name=name.replace("_", "-"))(passthrough), child)
            else:
                group.command(name=name)(child)

    root = click.Group(
        name=_root.name, callback=_root.callback, params=list(_root.params), help=_root.help
    )
    attach(root, tree)
    return root


def command_group():
    """The click group of every registered benchmark (the sbench command)."""
    from . import benchmarks_collection  # noqa: F401  (registers the benchmarks)

    ordered = sorted(benchmark.REGISTRY, key=lambda b: (b.__module__, b.__name__))
    return build([(command_path(b), _command(b)) for b in ordered])


def main(*args, **kwargs):
    return command_group()(*args, **kwargs, obj=types.SimpleNamespace())
