"""Parameter sweeps over benchmark configurations into one results table.

Restates the reference's sweep driver (stencil_benchmarks/tools/multirun.py:
42-119): a `Configuration` binds a benchmark class, a name and fixed
parameters; `run_scaling_benchmark` instantiates every configuration at
every domain of a range, runs it `executions` times and returns a pandas
DataFrame with one row per run: the run's results (time, bandwidth, ...),
`name`, and the benchmark's parameters as `pretty_parameters` columns
(tuple parameters split into `<name>-<i>`, plus `sbench-version` and
`benchmark-name`), so the CSV it writes has the reference's columns.
"""

import gc
import re
import warnings
from ast import literal_eval

import pandas as pd

from . import cli


def _kebab(class_name):
    """CamelCase class name -> kebab-case configuration name, the reference's
    naming (a dash before each run of capitals and digits 1-9 after the first
    character: JScanOtf -> j-scan-otf, Copy1D -> copy-1d)."""
    return re.sub(r"(?<=.)([A-Z1-9]+)", r"-\1", class_name).lower()


class Configuration:
    """A benchmark class with fixed parameters, run under a display name.

    Calling it with further keyword arguments (e.g. the domain) builds the
    benchmark and returns a function that runs it once and returns the
    result row.
    """

    def __init__(self, ctor, name=None, **kwargs):
        self.ctor = ctor
        self.name = _kebab(ctor.__name__) if name is None else name
        self.kwargs = kwargs

    def __call__(self, preprocess_args=None, **kwargs):
        arguments = {**self.kwargs, **kwargs}
        if preprocess_args is not None:
            arguments = preprocess_args(**arguments)
        bmark = self.ctor(**arguments)

        def run_once():
            row = dict(bmark.run())
            row["name"] = self.name
            row.update(cli.pretty_parameters(bmark))
            return row

        return run_once


def domains(min_exp=5, max_exp=11, k=80):
    """Square domains 2^min_exp .. 2^(max_exp-1) wide, k levels."""
    for exponent in range(min_exp, max_exp):
        yield 2**exponent, 2**exponent, k


def run_scaling_benchmark(configurations, executions, preprocess_args=None,
                          domain_range=None):
    """Every configuration at every domain, `executions` runs each."""
    domain_range = list(domains() if domain_range is None else domain_range)
    rows = []
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        with cli.ProgressBar() as progress:
            for domain in progress.report(domain_range):
                for config in progress.report(configurations):
                    run = config(preprocess_args=preprocess_args, domain=tuple(domain))
                    rows.extend(run() for _ in progress.report(range(executions)))
                    del run
                    gc.collect()  # release the device fields before the next size
    for message in sorted({str(w.message) for w in caught}):
        print("WARNING:", message)
    return pd.DataFrame(rows)


def truncate_block_size_to_domain(**kwargs):
    """Clip `block_size` (and `storage_block_size`) to the domain extents."""
    if "domain" in kwargs:
        domain = kwargs["domain"]
        if "block_size" in kwargs:
            kwargs["block_size"] = tuple(min(b, d) for b, d in zip(kwargs["block_size"], domain))
        if "storage_block_size" in kwargs:
            kwargs["storage_block_size"] = min(kwargs["storage_block_size"], domain[0])
    return kwargs


def default_kwargs(**defaults):
    """Fixed parameters overridable by `name=value` strings (the collection
    scripts' -o/--option): returns override(options, **more) -> kwargs."""

    def override(options=(), **more):
        kwargs = {**defaults, **more}
        for option in options or ():
            name, value = option.split("=", 1)
            kwargs[name.replace("-", "_")] = literal_eval(value)
        return kwargs

    return override
