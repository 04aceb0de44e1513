"""The gpuPairHMM command-line drop-in (main.cu:30-94) on the GPU.

Checks the reference's contract: the batching flags change performance only,
never output; the output is the writer's text (output_writer_worker.cu:
110-127) of the engine's scores; errors exit non-zero.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpupairhmm_amd", "gpuPairHMM")


def _run(args, timeout=120):
    say("cli: " + " ".join(args))
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout)


def test_cli_output_independent_of_batching_flags(lib, engine, tmp_path):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=31, num_records=6)
    inp = tmp_path / "in.txt"
    inp.write_text(b.to_text())
    outs = []
    for fb, tb, mode in [(0, 0, "batched"), (1, 1, "plain"), (5000, 300, "batched"),
                         (10000000, 1000000, "plain")]:
        out = tmp_path / f"out_{fb}_{tb}_{mode}.txt"
        r = _run(["--inputfile", str(inp), "--outputfile", str(out), "--fileBatchsize", str(fb),
                  "--transferBatchsize", str(tb), "--mode", mode])
        assert r.returncode == 0, r.stderr
        assert "GCUPS" in r.stdout
        outs.append(out.read_text())
    assert all(o == outs[0] for o in outs[1:])
    want = G.format_scores(b, engine.score_batch(b))
    assert outs[0] == want


def test_cli_readme_example(lib, tmp_path):
    out = tmp_path / "out.txt"
    r = _run(["--inputfile", os.path.join(ROOT, "tests", "golden", "readme_record.txt"),
              "--outputfile", str(out)])
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert lines[0] == "2 2"
    got = np.array([float(x) for line in lines[1:] for x in line.split()])
    want = np.array([-5.97153, -3.1966, -6.34043, -1.66333])
    assert np.all(np.abs(got - want) <= 2e-5)


def test_cli_errors(lib, tmp_path):
    r = _run(["--inputfile", str(tmp_path / "missing.txt"), "--outputfile",
              str(tmp_path / "o.txt")])
    assert r.returncode != 0
    bad = tmp_path / "bad.txt"
    bad.write_text("1 1\nACGT\n")
    r = _run(["--inputfile", str(bad), "--outputfile", str(tmp_path / "o2.txt")])
    assert r.returncode != 0
    r = _run(["--inputfile", str(bad), "--outputfile", ""])
    assert r.returncode != 0
