"""Results below the FP32 threshold without the FP64 rescue:

- range_rescue 1 (the default): the FP32 tile kernels keep a power-of-two
  scale per tile lane during the sweep (started at the insertion floor,
  lowered as values grow), so such results are finished in the same pass;
- range_rescue 2: a second FP32 pass recomputes them with per-lane scales
  chosen from the first pass's per-lane magnitudes.

Both are checked against the FP64 rescue alone (range_rescue=False) and the
FP64 oracle, on the seeded workloads and on inputs that exercise long reads
(more than 16 lanes), reads against unrelated haps (insertion-dominated
sums) and every alignment below the threshold."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

TOL = 1e-3          # the parity bar, absolute log10
FP32_TOL = 5e-5     # what an FP32 recompute reaches (the FP32 pass: about 1e-5)
SCORE_FLOOR = -75.26  # log10(FLT_MIN) - log10(FLT_MAX/16): the rescue threshold
INPASS, TWOPASS = 1, 2


def _max_delta(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both = np.isneginf(a) & np.isneginf(b)
    with np.errstate(invalid="ignore"):
        d = np.abs(a - b)
    d[both] = 0.0
    return float(np.max(d)) if d.size else 0.0


def _check_counts(mode, st, below):
    """below: the alignments the FP64-only engine rescued."""
    if mode == TWOPASS:
        assert st.rescued == below > 0
        assert st.f64_alignments <= 0.25 * st.rescued  # the FP32 recompute takes most
    else:
        # finished in the pass; only what the scaling cannot vouch for is
        # rescued (in FP64)
        assert st.scaled_alignments + st.rescued >= below > 0
        assert st.f64_alignments == st.rescued <= 0.05 * below


@pytest.mark.parametrize("mode", [INPASS, TWOPASS])
@pytest.mark.parametrize("cfg,seed,records", [(1, 1001, 60), (2, 1002, 40), (3, 1003, 3)])
def test_range_rescue_agrees_with_fp64(lib, oracle, cfg, seed, records, mode):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(cfg, seed=seed, num_records=records)
    say(f"case: range rescue {mode}, config {cfg}, {records} records, "
        f"{b.num_alignments} alignments")
    with G.Engine(small_batch=0, range_rescue=mode) as on:  # the tuned table at any size
        got = on.score_batch(b)
        st = on.stats()
    with G.Engine(small_batch=0, range_rescue=False) as off:
        f64 = off.score_batch(b)
        sf = off.stats()
    assert sf.f64_alignments == sf.rescued
    _check_counts(mode, st, sf.rescued)
    low = np.flatnonzero(f64 < SCORE_FLOOR + 0.5)
    assert len(low) > 0
    assert _max_delta(got, f64) <= FP32_TOL
    which = low if len(low) <= 3000 else np.sort(
        np.random.default_rng(seed).choice(low, 3000, replace=False))
    ref = oracle.score_subset(b.oracle_arrays(), which, oracle.F64)
    assert _max_delta(got[which], ref) <= FP32_TOL
    # the rest of the batch: the unscaled FP32 results, bit for bit (away
    # from the threshold, where the scaled sums keep contributions the
    # unscaled ones flush)
    hi = np.flatnonzero(f64 >= SCORE_FLOOR + 10.0)
    same = np.mean(got[hi].view(np.uint32) == f64[hi].view(np.uint32))
    print(f"above the threshold: {same:.6f} identical")
    assert same == 1.0 if mode == TWOPASS else same >= 0.999


@pytest.mark.parametrize("mode", [INPASS, TWOPASS])
@pytest.mark.parametrize("cfg,seed,records", [(2, 1012, 300), (3, 1013, 600)])
def test_range_rescue_matches_fp64_on_large_batches(lib, cfg, seed, records, mode):
    """Every alignment of a large seeded batch (2.4e6 / 7.7e4 alignments; the
    precision-stress workload's long haps make competing alignments whose
    early rows are far weaker than the best prefix): scores within FP32
    accuracy of the FP64 rescue's."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(cfg, seed=seed, num_records=records)
    say(f"case: range rescue {mode} vs FP64 rescue, config {cfg}, {b.num_alignments} alignments")
    with G.Engine(small_batch=0, range_rescue=mode) as on:
        got = on.score_batch(b)
        st = on.stats()
    with G.Engine(small_batch=0, range_rescue=False) as off:
        f64 = off.score_batch(b)
        sf = off.stats()
    d = _max_delta(got, f64)
    print(f"config {cfg}: max |delta| {d:.2e}, rescued {st.rescued}, in FP64 {st.f64_alignments}, "
          f"scaled {st.scaled_alignments} (FP64 engine: {sf.rescued})")
    assert d <= FP32_TOL
    _check_counts(mode, st, sf.rescued)


def test_range_rescue_every_alignment(lib, oracle):
    """Every alignment rescued (threshold above every sum): the two-pass form
    recomputes all of them, scores in the normal range included; the
    in-pass form sends the ones it did not scale to FP64."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(1, seed=11, num_records=40)
    n = b.num_alignments
    say(f"case: every alignment rescued, {n} alignments")
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    with G.Engine(rescue_min_sum=float("inf"), small_batch=0, range_rescue=TWOPASS) as e:
        got = e.score_batch(b)
        st = e.stats()
    assert st.rescued == n
    assert st.f64_alignments <= n // 20
    assert _max_delta(got, ref) <= FP32_TOL
    with G.Engine(rescue_min_sum=float("inf"), small_batch=0, range_rescue=INPASS) as e:
        got = e.score_batch(b)
        st = e.stats()
    assert st.scaled_alignments + st.rescued == n
    assert _max_delta(got, ref) <= FP32_TOL


@pytest.mark.parametrize("mode", [INPASS, TWOPASS])
def test_range_rescue_long_reads(lib, oracle, mode):
    """Reads of 300 to 760 bases against unrelated haps (scores far below the
    FP32 floor, sums made mostly of insertion paths), in tiles of 16 to 64
    lanes.  Two-pass: the first pass records the first 16 lanes'
    magnitudes, the scales of the lanes past them are extrapolated (and
    checked like any other)."""
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
    say(f"case: long reads, range rescue {mode}, {b.num_alignments} alignments")
    with G.Engine(small_batch=0, range_rescue=mode) as e:
        got = e.score_batch(b)
        st = e.stats()
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    assert np.all(ref < SCORE_FLOOR)
    if mode == TWOPASS:
        assert st.rescued == b.num_alignments
        assert st.f64_alignments < st.rescued
    else:
        assert st.scaled_alignments + st.rescued == b.num_alignments
        assert st.rescued <= b.num_alignments // 4
    assert _max_delta(got, ref) <= FP32_TOL


def test_reads_overhanging_the_hap(lib, oracle):
    """A read whose first bases lie before its hap's start: its likelihood
    is carried by paths that insert the overhang in the hap's first columns,
    where the deep rows' values sit at the insertion floor -- the values an
    exponent picked after the wavefront reached a lane would have flushed."""
    import gpupairhmm_amd as G
    rng = np.random.default_rng(23)
    rs = lambda n: "".join(rng.choice(list("ACGT"), n))
    recs = []
    for L, over in ((150, 110), (250, 200), (300, 150)):
        hap = rs(L + 100)
        reads = []
        for _ in range(4):
            body = rs(over) + hap[:L - over]
            q = "".join(chr(33 + int(x)) for x in rng.integers(20, 41, L))
            reads.append(" ".join([body, q, "I" * L, "I" * L, "+" * L]))
        recs.append("\n".join([f"{len(reads)} 1"] + reads + [hap]))
    b = G.HostBatch.parse("\n".join(recs) + "\n")
    say(f"case: overhanging reads, {b.num_alignments} alignments")
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    assert np.all(ref < SCORE_FLOOR)
    for sb in (0, None):  # the tuned and the small-batch table
        with G.Engine(small_batch=sb) as e:
            got = e.score_batch(b)
            st = e.stats()
        print(f"small_batch {sb}: scaled {st.scaled_alignments} rescued {st.rescued}")
        assert _max_delta(got, ref) <= FP32_TOL


def test_small_batches_rescue(lib):
    """Small batches run on the one-template table.  In-pass scaling works
    there as on the tuned table; the two-pass form rescues them in FP64
    directly (its second stage would add one sweep's latency)."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=1002, num_records=6)
    say(f"case: small batch rescue, {b.num_alignments} alignments")
    with G.Engine(range_rescue=TWOPASS) as e:
        e.score_batch(b)
        small = e.stats()
    with G.Engine(small_batch=0, range_rescue=TWOPASS) as e:
        e.score_batch(b)
        tuned = e.stats()
    assert small.rescued == tuned.rescued > 0
    assert small.f64_alignments == small.rescued
    assert tuned.f64_alignments < tuned.rescued
    with G.Engine() as e:
        e.score_batch(b)
        inpass = e.stats()
    assert inpass.scaled_alignments > 0
    assert inpass.f64_alignments == inpass.rescued <= 0.05 * small.rescued


@pytest.mark.parametrize("gcp,bq", [(40, (20, 41)), (10, (2, 8)), (25, (2, 41))])
def test_in_pass_scaling_extreme_qualities(lib, oracle, gcp, bq):
    """Gap-continuation quality 40 (eps = 1e-4: the insertion floor falls
    about 13 bits per row, so a 19-row lane spans more than the transfer
    factor can, and the span cap pulls the lanes below along) and base
    qualities down to 2 (emission values near 1/3 and 2/3): reads related and
    unrelated to their haps, every result within FP32 accuracy of the FP64
    oracle, whatever the scaling finishes and whatever it hands to FP64."""
    import gpupairhmm_amd as G
    rng = np.random.default_rng(31 + gcp)
    rs = lambda n: "".join(rng.choice(list("ACGT"), n))
    g = chr(33 + gcp)
    recs = []
    for L in (90, 150, 260):
        hap = rs(L + 120)
        reads = []
        for k in range(6):
            if k % 2:
                body = rs(L)  # unrelated
            else:
                start = int(rng.integers(0, 100))
                body = list(hap[start:start + L])
                for i in rng.choice(L, L // 30, replace=False):
                    body[i] = "ACGT"[("ACGT".index(body[i]) + 1) % 4]
                body = "".join(body)
            q = "".join(chr(33 + int(x)) for x in rng.integers(bq[0], bq[1], L))
            reads.append(" ".join([body, q, "I" * L, "I" * L, g * L]))
        recs.append("\n".join([f"{len(reads)} 2"] + reads + [hap, rs(L + 60)]))
    b = G.HostBatch.parse("\n".join(recs) + "\n")
    say(f"case: extreme qualities gcp {gcp} bq {bq}, {b.num_alignments} alignments")
    # the GPU semantics fix eps at 0.1; the CPU semantics take it from the
    # read's gcp quality
    for sem, osem in ((G.SEM_GPU, oracle.SEM_GPU), (G.SEM_CPU, oracle.SEM_CPU)):
        want = oracle.score_batch(b.oracle_arrays(), oracle.F64, semantics=osem)
        for sb in (0, None):
            with G.Engine(semantics=sem, small_batch=sb) as e:
                got = e.score_batch(b)
                st = e.stats()
            print(f"semantics {sem} small_batch {sb}: scaled {st.scaled_alignments} "
                  f"rescued {st.rescued} max |d| {_max_delta(got, want):.2e}")
            assert _max_delta(got, want) <= FP32_TOL
