"""``sbench-analyze``: tables and plots from the CSV files that ``sbench -o``
and the collection scripts write.

The commands and their options are those of the reference's
``sbench-analyze`` (stencil_benchmarks/scripts/sbench_analyze.py:46-230), so
the same command lines work:

* ``print CSV``: ``-f/--filter`` (a pandas query on the rows), ``-g/--group``
  (repeatable) or ``-a/--auto-group`` (every non-float column, the ones with
  fewest distinct values first), ``--aggregation`` (``median``), ``-u`` (the
  last group becomes columns), ``-s/--select`` (repeatable), ``--sort``,
  ``-c/--common`` (print the values every row shares instead), ``-o`` (also
  write the table as CSV).  Columns holding one value in every row are set
  aside before grouping: they describe the whole file.
* ``plot CSV -s COLUMN``: one line per group over the last group's values,
  with ``--uniform`` x positions, ``--ylim``, ``--title``, ``-r label=value``
  reference lines, ``--relative-to`` (y as a percentage of a value),
  ``--label-regex /pattern/repl/`` and ``-o FILE`` (``--dpi``); the shared
  values are printed under the plot.

The table is built by a list of steps, each taking and returning a table.

    sbench-analyze print vadv.csv -g name -g domain-0 -u -s bandwidth
"""

import re
import textwrap
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, List, Optional, Sequence

import click
import pandas as pd

pd.set_option("display.max_colwidth", None)


def read_csv(path) -> pd.DataFrame:
    """An sbench CSV: the first column is the row index; only the literal
    strings nan/NaN are missing values (a string column may hold 'none')."""
    return pd.read_csv(path, index_col=0, keep_default_na=False, na_values=["nan", "NaN"])


@dataclass
class Arrangement:
    """The steps from a CSV table to a report, and what they set aside."""

    query: Optional[str] = None
    groups: Sequence[str] = ()
    auto_group: bool = False
    aggregation: str = "median"
    unstack: bool = False
    select: Sequence[str] = ()
    sort: bool = False
    common: pd.Series = field(default_factory=lambda: pd.Series(dtype=object))

    def steps(self) -> List[Callable[[pd.DataFrame], pd.DataFrame]]:
        chosen = [self._set_aside_common]
        if self.query:
            chosen.insert(0, lambda t: t.query(self.query))
        if self.auto_group or self.groups:
            chosen.append(self._group)
        if self.unstack:
            chosen.append(lambda t: t.unstack())
        if self.select:
            chosen.append(self._select)
        if self.sort:
            chosen.append(lambda t: t.sort_values())
        return chosen

    def _set_aside_common(self, table):
        single = [c for c in table.columns if table[c].nunique(dropna=False) <= 1]
        if len(table):
            self.common = table[single].iloc[0].sort_index()
        return table.drop(columns=single)

    def group_columns(self, table) -> List[str]:
        if not self.auto_group:
            return list(self.groups)
        distinct = table.nunique(dropna=False).sort_values(kind="stable")
        return [c for c in distinct.index if table[c].dtype != float]

    def _group(self, table):
        keys = self.group_columns(table)
        return table.groupby(keys).agg(self.aggregation) if keys else table

    def _select(self, table):
        return table[self.select[0]] if len(self.select) == 1 else table[list(self.select)]

    def apply(self, table: pd.DataFrame):
        for step in self.steps():
            table = step(table)
        return table


@click.group()
def main():
    """Analyse sbench CSV output."""


@main.command(name="print")
@click.argument("csv", type=click.Path(exists=True, path_type=Path))
@click.option("--common/--non-common", "-c", default=False,
              help="Print the values shared by every row instead of the table.")
@click.option("--auto-group/--no-auto-group", "-a", default=False,
              help="Group by every non-float column.")
@click.option("--group", "-g", multiple=True, help="Group by this column (repeatable).")
@click.option("--unstack/--dont-unstack", "-u", default=False,
              help="Turn the last group into columns (a pivot table).")
@click.option("--select", "-s", multiple=True, help="Keep this column (repeatable).")
@click.option("--filter", "-f", "query", help="Keep the rows matching this pandas query.")
@click.option("--aggregation", default="median", help="Aggregation of grouped rows.")
@click.option("--sort/--no-sort", default=False, help="Sort the values.")
@click.option("--output", "-o", type=click.File(mode="w"), help="Also write the table as CSV.")
def print_csv(csv, common, auto_group, group, unstack, select, query, aggregation, sort,
              output):
    """Print a report of an sbench CSV file."""
    arrangement = Arrangement(query=query, groups=group, auto_group=auto_group,
                              aggregation=aggregation, unstack=unstack, select=select,
                              sort=sort)
    table = arrangement.apply(read_csv(csv))
    if common:
        table = arrangement.common
    click.echo(table.to_string())
    if output:
        table.to_csv(output)


def _label_rewrites(patterns):
    rewrites = []
    for spec in patterns:
        sep = spec[:1]
        body = spec[1:-1] if len(spec) >= 3 and spec.endswith(sep) else None
        if not body or sep not in body:
            raise click.BadParameter(f"expected /pattern/replacement/, got {spec!r}",
                                     param_hint="--label-regex")
        pattern, replacement = body.split(sep, 1)
        rewrites.append((re.compile(pattern), replacement))
    return rewrites


@main.command()
@click.argument("csv", type=click.Path(exists=True, path_type=Path))
@click.option("--uniform/--non-uniform", help="Equidistant x positions.")
@click.option("--ylim", type=float, nargs=2, help="Y-axis limits.")
@click.option("--title", "-t", help="Plot title (default: the file name).")
@click.option("--auto-group/--no-auto-group", "-a", default=False,
              help="Group by every non-float column.")
@click.option("--group", "-g", multiple=True, help="Group by this column (repeatable).")
@click.option("--select", "-s", required=True, help="Column to plot.")
@click.option("--filter", "-f", "query", help="Keep the rows matching this pandas query.")
@click.option("--aggregation", default="median", help="Aggregation of grouped rows.")
@click.option("--reference", "-r", multiple=True, help="Reference line, as label=value.")
@click.option("--relative-to", type=float, help="Plot y as a percentage of this value.")
@click.option("--ascii/--no-ascii", help="Plot in the terminal (needs drawilleplot).")
@click.option("--label-regex", multiple=True,
              help="Rewrite labels with /pattern/replacement/ (Python regex).")
@click.option("--output", "-o", type=click.Path(path_type=Path), help="Output file.")
@click.option("--dpi", default=300, type=int, help="Output resolution.")
def plot(csv, uniform, ylim, title, auto_group, group, select, query, aggregation, reference,
         relative_to, ascii, label_regex, output, dpi):
    """Plot one column of an sbench CSV file, one line per group."""
    import matplotlib

    if output:
        matplotlib.use("Agg")
    elif ascii:
        matplotlib.use("module://drawilleplot")
    from matplotlib import pyplot as plt
    from matplotlib import ticker

    rewrites = _label_rewrites(label_regex)
    arrangement = Arrangement(query=query, groups=group, auto_group=auto_group,
                              aggregation=aggregation, unstack=True, select=[select])
    table = arrangement.apply(read_csv(csv))
    lines = ([(table.index.name, table)] if isinstance(table, pd.Series)
             else list(table.iterrows()))
    figure, axes = plt.subplots()
    for key, series in lines:
        if isinstance(key, tuple):
            label = ", ".join(f"{n}={v}" for n, v in zip(table.index.names, key))
        else:
            label = str(key)
        for pattern, replacement in rewrites:
            label = pattern.sub(replacement, label)
        xs = range(len(series)) if uniform else series.index
        ys = series.to_numpy() / relative_to if relative_to else series.to_numpy()
        axes.plot(xs, ys, marker="o", label=label)
        if uniform:
            axes.set_xticks(list(xs), [str(v) for v in series.index], rotation=45)
    axes.set_xlabel(table.index.name if isinstance(table, pd.Series) else table.columns.name)
    for number, ref in enumerate(reference):
        label, _, value = ref.partition("=")
        axes.axhline(float(value), color="black", linestyle=(0, (4 + 4 * number, 3)),
                     label=label)
    if len(lines) > 1 or reference:
        axes.legend()
    axes.set_ylabel(select)
    if ylim:
        axes.set_ylim(*ylim)
    if relative_to:
        axes.yaxis.set_major_formatter(ticker.PercentFormatter(1.0))
    axes.set_title(title or csv.name)
    shared = ", ".join(f"{k}: {v}" for k, v in arrangement.common.items())
    figure.text(0.5, 0.005, "\n".join(textwrap.wrap(shared, 130)), fontsize=6,
                ha="center", va="bottom")
    figure.tight_layout(rect=(0, 0.05, 1, 1))
    if output:
        figure.savefig(output, dpi=dpi)
    else:
        plt.show()


if __name__ == "__main__":
    main()
