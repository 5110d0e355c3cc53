"""bench.py's host-side pieces, CPU only: the algorithmic byte counts the
roofline divides by, the traffic lookup that must refuse figures measured on
other kernel sources, the host report of the CPU legs and the one-GPU rule."""

import json
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parents[1]


def test_algorithmic_bytes_of_the_measured_workloads():
    # compulsory bytes: every input once with the halo it needs, every output once
    assert bench.hdiff_bytes((512, 512, 160), 8) == 1_011_896_320
    assert bench.vadv_bytes((512, 512, 160), 8) == 2_013_921_280
    assert bench.vadv_bytes((256, 256, 80), 8) == 251_822_080
    assert bench.vadv_bytes((256, 256, 80), 4) == 125_911_040
    assert bench.lap_bytes((512, 512, 160), 8) == 673_715_200
    # halving the item size halves the bytes
    assert bench.hdiff_bytes((64, 48, 20), 4) * 2 == bench.hdiff_bytes((64, 48, 20), 8)


def test_stream_sweep_covers_every_shape_once():
    assert len(bench.STREAM_SWEEP) == 96 == len(set(bench.STREAM_SWEEP))
    assert {s[0] for s in bench.STREAM_SWEEP} == {128, 256, 512, 1024}


def test_profiled_traffic_only_for_the_sources_it_was_measured_on(tmp_path, monkeypatch):
    (tmp_path / "profiles").mkdir()
    kernel = tmp_path / "kernel.hip"
    kernel.write_text("// version 1\n")
    monkeypatch.setattr(bench, "ROOT", tmp_path)
    entry = {"traffic_bytes": 1234, "sources": ["kernel.hip"],
             "sources_sha256": bench.sources_sha(["kernel.hip"])}
    (tmp_path / "profiles" / "traffic.json").write_text(json.dumps({"k": entry}))
    assert bench.profiled_traffic("k") == 1234
    assert bench.profiled_traffic("missing") is None
    kernel.write_text("// version 2\n")
    assert bench.profiled_traffic("k") is None


def test_committed_traffic_entries_name_their_sources():
    entries = json.loads((ROOT / "profiles" / "traffic.json").read_text())
    for key, entry in entries.items():
        if key == "calibration":
            continue
        assert entry["traffic_bytes"] == entry["fetch_bytes"] + entry["write_bytes"], key
        assert entry["traffic_bytes"] >= entry["algorithmic_bytes"], key
        for source in entry["sources"]:
            assert (ROOT / source).is_file(), (key, source)
        assert len(entry["sources_sha256"]) == 64, key


def test_host_description():
    host = bench.host_description()
    assert host["logical_cpus"] >= 1
    assert isinstance(host["model"], str) and host["model"]
    assert host["physical_cores"] is None or 1 <= host["physical_cores"] <= host["logical_cpus"]


def test_more_than_one_gpu_is_refused():
    done = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)
    assert done.returncode == 2
    assert "one GPU" in done.stderr and done.stdout == ""
