"""bench.py keeps its output contract: one JSON line with the headline
metric, the roofline of the dominant kernel and the CPU baseline."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT, say

BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_cpu_baseline_fields(lib, oracle):
    import gpupairhmm_amd as G
    sys.path.insert(0, ROOT)
    import bench
    batch = G.HostBatch.synthetic(0, seed=5, num_records=20)
    per_record = batch.cells() // 20
    cb = bench.cpu_baseline(batch, cells_target=10 * per_record, single_cells=2 * per_record)
    assert cb["unit"] == "GCUPS" and cb["value"] > 0 and cb["cores"] >= 1
    assert cb["kind"] in ("reference", "port")
    assert cb["single_thread"]["value"] > 0
    assert "10 seeded records of 20" in cb["sample"]
    assert "2 seeded records" in cb["single_thread"]["sample"]


def test_seeded_records_reach_target(lib):
    import gpupairhmm_amd as G
    sys.path.insert(0, ROOT)
    import bench
    batch = G.HostBatch.synthetic(2, seed=9, num_records=50)
    recs = bench.seeded_records(batch, batch.cells() // 4, seed=3)
    assert recs == sorted(set(recs)) and 1 <= len(recs) < 50
    sub = batch.subset(recs)
    assert sub.cells() >= batch.cells() // 4
    assert recs == bench.seeded_records(batch, batch.cells() // 4, seed=3)


@pytest.mark.gpu
def test_bench_json_line(lib):
    import gpupairhmm_amd as G
    say("case: bench.py --no-extras --no-cpu --steps 2 --warmup 1 --records 200")
    r = subprocess.run([sys.executable, BENCH, "--no-extras", "--no-cpu", "--steps", "2",
                        "--warmup", "1", "--records", "200"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    assert d["unit"] == "GCUPS" and d["value"] > 0 and d["n_gpus"] == 1
    assert d["steps"] == 2 and d["warmup"] == 1 and d["higher_is_better"] is True
    assert d["ms_per_step"] > 0 and d["vs_baseline"] is None and d["dtype"] == "f32"
    batch = G.HostBatch.synthetic(2, seed=1002, num_records=200)
    assert d["config"]["records"] == 200
    assert d["config"]["alignments"] == batch.num_alignments
    assert d["config"]["cells"] == batch.cells()
    rf = d["roofline"]
    assert rf["unit"] == "TFLOP/s" and rf["peak"] == 157.3
    assert 0 < rf["frac"] < 1 and abs(rf["achieved"] / rf["peak"] - rf["frac"]) < 1e-3
    assert d["parity"]["max_abs_dlog10"] <= 1e-3
