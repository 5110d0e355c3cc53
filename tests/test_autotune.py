"""Autotuner logic without a device - CPU only.

Coordinate descent picks the fastest option of the first space with the
default of the second, then the fastest of the second with that choice;
choices persist in the JSON cache (atomic rewrite, unreadable or unwritable
files tolerated); workload keys depend on the shapes, not their order.
"""

import json

import numpy as np

from stencil_benchmarks_amd import autotune


def test_descend_is_coordinate_descent():
    space_a = ({"by": 8}, {"by": 16}, {"by": 32})
    space_b = ({"variant": 4}, {"variant": 3})
    cost = {(8, 4): 5.0, (16, 4): 3.0, (32, 4): 4.0, (8, 3): 0.5, (16, 3): 2.0, (32, 3): 9.0}
    calls = []

    def measure(a, b):
        calls.append((a["by"], b["variant"]))
        return cost[(a["by"], b["variant"])]

    a, b, t, trials = autotune.descend(space_a, space_b, measure, space_b[0])
    # (8, 3) would be the global optimum, but descent fixes a = 16 first
    assert (a, b, t) == ({"by": 16}, {"variant": 3}, 2.0)
    assert calls == [(8, 4), (16, 4), (32, 4), (16, 3)] and len(trials) == 4


def test_cache_roundtrip_and_tolerance(tmp_path):
    path = tmp_path / "sub" / "tune.json"
    cache = autotune.TuneCache(path)
    assert cache.get("k") is None
    cache.put("k", {"hdiff": {"by": 16}, "vadv": {}, "seconds": 1.0})
    assert json.loads(path.read_text())["k"]["hdiff"] == {"by": 16}
    assert autotune.TuneCache(path).get("k")["seconds"] == 1.0
    path.write_text("not json")
    assert autotune.TuneCache(path).entries == {}
    blocked = tmp_path / "file"
    blocked.write_text("")
    autotune.TuneCache(blocked / "tune.json").put("k", {})  # parent is a file: no error


def test_workload_key_ignores_order_and_names_types():
    a = autotune.workload_key("batch", [(32, 40, 50), (64, 64, 80)], np.float32, np.float64)
    b = autotune.workload_key("batch", [(64, 64, 80), (32, 40, 50)], "float32", "float64")
    c = autotune.workload_key("batch", [(64, 64, 80), (32, 40, 51)], "float32", "float64")
    assert a == b != c and "float32|float64" in a


def test_spaces_hold_only_built_combinations():
    built = {(16, 2, 2), (16, 2, 0), (16, 8, 2), (16, 8, 0), (8, 1, 2), (32, 2, 2), (16, 4, 2),
             (16, 1, 2)}  # (tile rows, planes, cache flags) of the batched hdiff kernel
    for h in autotune.BATCH_HDIFF_SPACE:
        flags = (h.get("unroll", 0) & 3) if h.get("unroll") else 2
        assert (h["by"], h["bz"], flags) in built
    assert len(autotune.PLAIN_HDIFF_SPACE) == 3 * 5 * 4
    for v in autotune.BATCH_VADV_SPACE:  # levels per step 4 / 8, prefetch depth 4 / 8
        assert v["variant"] in (3, 4) and v.get("unroll", 0) in (0, 4, 8)
        assert v.get("bx", 0) in (0, 4, 5) and (v["variant"] == 4 or "bx" not in v)


def test_launch_from_sets_fields():
    launch = autotune.launch_from({"variant": 4, "by": 16})
    assert (launch.variant, launch.by, launch.bz) == (4, 16, 0)
