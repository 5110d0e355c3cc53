"""Tables from sbench CSV files (the `print` command of the reference's
sbench-analyze, stencil_benchmarks/scripts/sbench_analyze.py:46-149).

Reads a CSV written by `sbench ... -o` or a collection script (first column
the row index), drops the columns that hold one value in every row (shown
with --common), optionally filters rows (pandas query), groups by columns
(--group, or --auto-group: every non-float column), aggregates (median by
default), unstacks the last group into columns and selects or sorts.

    python -m stencil_benchmarks_amd.scripts.sbench_analyze print vadv.csv \\
        -g name -g domain-0 -u -s bandwidth
"""

import click
import pandas as pd


def read_csv(path):
    return pd.read_csv(path, index_col=0, na_values=["nan", "NaN"], keep_default_na=False)


def split_common(table):
    """(columns that vary between rows, Series of the single-valued ones)."""
    distinct = table.nunique(dropna=False)
    varying = table[[c for c in table.columns if distinct[c] > 1]]
    common = table[[c for c in table.columns if distinct[c] <= 1]]
    return varying, (common.iloc[0].sort_index() if len(common) else pd.Series(dtype=object))


def auto_groups(table):
    """Non-float columns, fewest distinct values first."""
    order = table.nunique(dropna=False).sort_values().index
    return [c for c in order if table.dtypes[c] != float]


def arrange(table, query=None, groups=(), aggregation="median", unstack=False, select=(),
            sort=False):
    if query:
        table = table.query(query)
    table, common = split_common(table)
    if groups == "auto":
        groups = auto_groups(table)
    if groups:
        table = table.groupby(list(groups)).agg(aggregation)
    if unstack:
        table = table.unstack()
    if select:
        table = table[select[0] if len(select) == 1 else list(select)]
    if sort:
        table = table.sort_values()
    return table, common


@click.group()
def main():
    """Analyse sbench CSV output."""


@main.command(name="print")
@click.argument("csv", type=click.Path(exists=True))
@click.option("--common/--non-common", "-c", default=False,
              help="Print only the values common to every row.")
@click.option("--auto-group/--no-auto-group", "-a", default=False,
              help="Group by every non-float column.")
@click.option("--group", "-g", multiple=True, help="Group by this column.")
@click.option("--unstack/--dont-unstack", "-u", default=False,
              help="Turn the last group into columns (pivot table).")
@click.option("--select", "-s", multiple=True, help="Print only this column.")
@click.option("--filter", "-f", "query", help="Row filter (pandas query expression).")
@click.option("--aggregation", default="median", help="Aggregation of grouped rows.")
@click.option("--sort/--no-sort", default=False, help="Sort the values.")
@click.option("--output", "-o", type=click.File(mode="w"), help="Also write the table as CSV.")
def print_csv(csv, common, auto_group, group, unstack, select, query, aggregation, sort,
              output):
    """Print a report of an sbench CSV file."""
    table, common_values = arrange(read_csv(csv), query, "auto" if auto_group else group,
                                   aggregation, unstack, select, sort)
    if common:
        table = common_values
    click.echo(table.to_string())
    if output:
        table.to_csv(output)


if __name__ == "__main__":
    main()
