"""The range-extended FP32 rescue pass: results below the FP32 threshold are
recomputed in FP32 with a power-of-two scale per tile lane (chosen from the
first pass's per-lane magnitudes) before anything goes to FP64.  Checked
against the FP64 rescue alone (range_rescue=False) and the FP64 oracle, on the
seeded workloads and on inputs that exercise the store's capacity and
multi-unit (more than 16 lanes) alignments."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

TOL = 1e-3          # the parity bar, absolute log10
FP32_TOL = 5e-5     # what an FP32 recompute reaches (the FP32 pass: about 1e-5)
SCORE_FLOOR = -75.26  # log10(FLT_MIN) - log10(FLT_MAX/16): the rescue threshold


def _max_delta(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both = np.isneginf(a) & np.isneginf(b)
    with np.errstate(invalid="ignore"):
        d = np.abs(a - b)
    d[both] = 0.0
    return float(np.max(d)) if d.size else 0.0


@pytest.mark.parametrize("cfg,seed,records", [(1, 1001, 60), (2, 1002, 40), (3, 1003, 3)])
def test_range_rescue_agrees_with_fp64(lib, oracle, cfg, seed, records):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(cfg, seed=seed, num_records=records)
    say(f"case: range-extended rescue, config {cfg}, {records} records, "
        f"{b.num_alignments} alignments")
    with G.Engine(small_batch=0) as on:  # the tuned table at any size
        got = on.score_batch(b)
        st = on.stats()
    with G.Engine(small_batch=0, range_rescue=False) as off:
        f64 = off.score_batch(b)
        sf = off.stats()
    assert st.rescued == sf.rescued > 0
    assert sf.f64_alignments == sf.rescued
    # the FP32 recompute takes most of them
    assert st.f64_alignments <= 0.25 * st.rescued
    low = np.flatnonzero(f64 < SCORE_FLOOR + 0.5)
    assert len(low) > 0
    assert _max_delta(got, f64) <= FP32_TOL
    which = low if len(low) <= 3000 else np.sort(
        np.random.default_rng(seed).choice(low, 3000, replace=False))
    ref = oracle.score_subset(b.oracle_arrays(), which, oracle.F64)
    assert _max_delta(got[which], ref) <= FP32_TOL
    # and the rest of the batch is untouched by the rescue path
    hi = np.flatnonzero(f64 >= SCORE_FLOOR + 0.5)
    assert np.array_equal(got[hi].view(np.uint32), f64[hi].view(np.uint32))


@pytest.mark.parametrize("cfg,seed,records", [(2, 1012, 300), (3, 1013, 600)])
def test_range_rescue_matches_fp64_on_large_batches(lib, cfg, seed, records):
    """Every alignment of a large seeded batch (2.4e6 / 7.7e4 alignments; the
    precision-stress workload's long haps make competing alignments whose
    early rows are far weaker than the best prefix): range-extended scores
    within FP32 accuracy of the FP64 rescue's."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(cfg, seed=seed, num_records=records)
    say(f"case: range-extended vs FP64 rescue, config {cfg}, {b.num_alignments} alignments")
    with G.Engine(small_batch=0) as on:
        got = on.score_batch(b)
        st = on.stats()
    with G.Engine(small_batch=0, range_rescue=False) as off:
        f64 = off.score_batch(b)
    d = _max_delta(got, f64)
    print(f"config {cfg}: max |delta| {d:.2e}, rescued {st.rescued}, in FP64 {st.f64_alignments}")
    assert d <= FP32_TOL
    assert st.f64_alignments <= 0.25 * st.rescued


def test_range_rescue_every_alignment(lib, oracle):
    """Every alignment rescued (threshold above every sum): the range-extended
    pass recomputes all of them, scores in the normal range included."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(1, seed=11, num_records=40)
    n = b.num_alignments
    say(f"case: every alignment rescued, {n} alignments")
    with G.Engine(rescue_min_sum=float("inf"), small_batch=0) as e:
        got = e.score_batch(b)
        st = e.stats()
    assert st.rescued == n
    assert st.f64_alignments <= n // 20
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    assert _max_delta(got, ref) <= FP32_TOL


def test_range_rescue_long_reads(lib, oracle):
    """Reads of 300 to 760 bases against unrelated haps (scores far below the
    FP32 floor), in tiles of 16 to 64 lanes: the first pass records the first
    16 lanes' magnitudes, the scales of the lanes past them are extrapolated
    (and checked like any other)."""
    import gpupairhmm_amd as G
    rng = np.random.default_rng(17)
    rs = lambda n: "".join(rng.choice(list("ACGT"), n))
    recs = []
    for L in (300, 390, 500, 640, 760):
        reads = []
        for _ in range(3):
            q = "".join(chr(33 + int(x)) for x in rng.integers(15, 41, L))
            gi = "".join(chr(33 + int(x)) for x in rng.integers(30, 46, L))
            reads.append(" ".join([rs(L), q, gi, gi, "+" * L]))
        haps = [rs(int(h)) for h in rng.integers(L // 2, L + 300, 3)]
        recs.append("\n".join([f"{len(reads)} {len(haps)}"] + reads + haps))
    b = G.HostBatch.parse("\n".join(recs) + "\n")
    say(f"case: long reads, {b.num_alignments} alignments")
    with G.Engine(small_batch=0) as e:
        got = e.score_batch(b)
        st = e.stats()
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    assert np.all(ref < SCORE_FLOOR)
    assert st.rescued == b.num_alignments
    assert st.f64_alignments < st.rescued
    assert _max_delta(got, ref) <= FP32_TOL


def test_small_batches_rescue_in_fp64(lib):
    """Below small_batch the rescue goes to FP64 directly (its time is one
    sweep's latency per stage); above it the range-extended pass takes most."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=1002, num_records=6)
    say(f"case: small batch rescue, {b.num_alignments} alignments")
    with G.Engine() as e:
        e.score_batch(b)
        small = e.stats()
    with G.Engine(small_batch=0) as e:
        e.score_batch(b)
        tuned = e.stats()
    assert small.rescued == tuned.rescued > 0
    assert small.f64_alignments == small.rescued
    assert tuned.f64_alignments < tuned.rescued
