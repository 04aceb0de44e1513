"""GPU parity: the HIP engine against the CPU FP64 oracle.

Bar (from the task statement): max |delta log10| <= 1e-3 against the FP64
restatement of the reference's forward algorithm, on the same seeded inputs.
Integer bookkeeping (output placement, chunking, modes, request mixing) must be
bit-exact.  All calls go through the C ABI (ctypes).
"""
import numpy as np
import pytest

from conftest import say

TOL = 1e-3  # absolute, log10
FLOOR = -75.26  # log10(FLT_MIN) - log10(FLT_MAX/16): scores below it are rescued

pytestmark = pytest.mark.gpu


def _max_delta(gpu, ref):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    both_inf = np.isneginf(gpu) & np.isneginf(ref)
    with np.errstate(invalid="ignore"):
        d = np.abs(gpu - ref)
    d[both_inf] = 0.0
    return float(np.max(d)) if d.size else 0.0


def test_readme_known_answer(engine, oracle):
    import gpupairhmm_amd as G
    say("case: README record on the GPU")
    with open("tests/golden/readme_record.txt") as f:
        b = G.HostBatch.parse(f.read())
    got = engine.score_batch(b)
    expected = np.array([-5.97153, -3.1966, -6.34043, -1.66333])
    decimals = np.array([5, 4, 5, 5])
    # the README prints an FP32 result to 6 significant digits: allow its
    # rounding half-unit plus 1e-5 of FP32 noise
    assert np.all(np.abs(got - expected) <= 0.5 * 10.0 ** (-decimals) + 1e-5), got
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    assert _max_delta(got, ref) <= 2e-5
    text = G.format_scores(b, got)
    assert text.splitlines()[0] == "2 2"


@pytest.mark.parametrize("config,records,sample",
                         [(0, 1000, 0), (1, 40, 0), (2, 60, 40000), (3, 20, 0), (4, 6, 0)])
def test_config_parity(engine, oracle, config, records, sample):
    """Every configuration against the FP64 oracle (all alignments, or a
    seeded sample of `sample` alignments of a larger batch)."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(config, seed=1000 + config, num_records=records)
    say(f"case: config {config}, {records} records, {b.num_alignments} alignments")
    got = engine.score_batch(b)
    if sample and sample < b.num_alignments:
        which = np.sort(np.random.default_rng(config).choice(b.num_alignments, sample,
                                                             replace=False))
        ref = oracle.score_subset(b.oracle_arrays(), which, oracle.F64)
        got_s = got[which]
    else:
        ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
        got_s = got
    d = _max_delta(got_s, ref)
    print(f"config {config}: max |delta log10| = {d:.3e} over {len(ref)} alignments")
    assert np.all(np.isfinite(got))
    assert d <= TOL


def test_tuned_table_batched_parity(lib, oracle):
    """A batch above the small-batch size (so the tuned table's persistent
    multi-bin launches, multi-hap segments and cross-record item bins run)
    against the FP64 oracle on a seeded sample."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=2020, num_records=300)
    assert b.num_alignments > (1 << 18)
    say(f"case: tuned table, 300 config-2 records, {b.num_alignments} alignments")
    with G.Engine(small_batch=0) as e:
        got = e.score_batch(b)
    which = np.sort(np.random.default_rng(7).choice(b.num_alignments, 20000, replace=False))
    ref = oracle.score_subset(b.oracle_arrays(), which, oracle.F64)
    d = _max_delta(got[which], ref)
    print(f"tuned table: max |delta log10| = {d:.3e}")
    assert np.all(np.isfinite(got)) and d <= TOL


def _every_class_batch(tiles, seed=31):
    """One record per tile class: reads at the class's bound and just above
    the previous bound, half of them with constant insertion/deletion
    qualities (the rescaled FAST form admits them) and half with qualities
    that jump between rows (e_r > 1/2: the SAFE form), against haps of
    several lengths; plus a read beyond the largest tile."""
    import gpupairhmm_amd as G
    rng = np.random.default_rng(seed)
    lines, prev = [], 0
    bounds = [t[2] for t in tiles] + [tiles[-1][2] + 200]
    for b in bounds:
        reads = []
        for L in sorted({max(1, prev + 1), b}):
            for form in ("fast", "safe"):
                r = "".join(rng.choice(list("ACGT"), L))
                q = "".join(chr(33 + int(x)) for x in rng.integers(10, 41, L))
                if form == "fast":
                    gi = "I" * (L - 1) + "N" if L > 1 else "N"
                else:
                    gi = "".join(chr(33 + int(x)) for x in rng.integers(10, 46, L))
                reads.append(" ".join([r, q, gi, gi, "+" * L]))
        haps = ["".join(rng.choice(list("ACGT"), int(H))) for H in rng.integers(30, 400, 3)]
        lines.append(f"{len(reads)} {len(haps)}")
        lines += reads + haps
        prev = b
    return G.HostBatch.parse("\n".join(lines) + "\n")


@pytest.mark.parametrize("small_batch", [None, 0])
def test_every_tile_class_and_form_in_one_call(lib, oracle, small_batch):
    """Reads of every tile class, in both arithmetic forms, in one batched
    call (the one-template table and the tuned table), against the oracle;
    the plain engine gives the same bits."""
    import gpupairhmm_amd as G
    with G.Engine(small_batch=small_batch) as e:
        b = _every_class_batch(e.tiles())
        say(f"case: every tile class and form, small_batch={small_batch}, "
            f"{b.num_alignments} alignments")
        got = e.score_batch(b)
        st = e.stats()
    with G.Engine(mode=G.MODE_PLAIN, small_batch=small_batch) as p:
        plain = p.score_batch(b)
        pst = p.stats()
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    assert _max_delta(got, ref) <= TOL
    assert np.array_equal(got.view(np.uint32), plain.view(np.uint32))
    # both forms ran: some alignments FAST, some not; the long read went to
    # the multi-tile kernel
    assert 0 < pst.fast_alignments < b.num_alignments
    assert pst.multitile_alignments >= 3
    del st


def test_plain_and_batched_identical(engine, plain_engine):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=77, num_records=4)
    say(f"case: plain vs batched, {b.num_alignments} alignments")
    a = engine.score_batch(b)
    p = plain_engine.score_batch(b)
    assert np.array_equal(a.view(np.uint32), p.view(np.uint32))


def _same_but_rescued(a, b):
    """Scores of one batch on two tile tables: bit-identical well above the
    FP32 rescue floor; near and below it the lanes are scaled differently on
    the two tables (contributions one flushes the other keeps) or a rescue
    finishes an alignment on one table and hands it to FP64 on the other
    (all FP32-accurate)."""
    hi = (a > FLOOR + 10.0) & (b > FLOOR + 10.0)
    return (np.array_equal(a[hi].view(np.uint32), b[hi].view(np.uint32)) and
            _max_delta(a[~hi], b[~hi]) <= 5e-5)


def test_small_batch_table_agrees(lib, engine):
    """Small calls run on the one-template table (16 rows per lane); the
    tuned table gives the same bits above the rescue floor."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=78, num_records=6)
    say(f"case: small-batch table vs tuned table, {b.num_alignments} alignments")
    a = engine.score_batch(b)
    with G.Engine(small_batch=0) as e2:
        t = e2.score_batch(b)
    assert _same_but_rescued(a, t)
    print(f"identical: {np.mean(a.view(np.uint32) == t.view(np.uint32)):.4f}")


@pytest.mark.parametrize("tb", [1, 7, 1000, 0])
def test_transfer_chunking_does_not_change_scores(lib, engine, tb):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(1, seed=5, num_records=12)
    say(f"case: transfer_batchsize {tb}")
    base = engine.score_batch(b)
    with G.Engine(transfer_batchsize=tb) as e2:
        got = e2.score_batch(b)
    assert np.array_equal(base.view(np.uint32), got.view(np.uint32))


def _record(reads, haps, q=30, iq=40, dq=40, gcp=10):
    """Line-record text for reads (strings) x haps (strings)."""
    lines = [f"{len(reads)} {len(haps)}"]
    for r in reads:
        L = len(r)
        bq = "".join(chr(33 + (q if isinstance(q, int) else q[i % len(q)])) for i in range(L))
        ins = chr(33 + iq) * max(L - 1, 0) + (chr(33 + 45) if L else "")
        dl = chr(33 + dq) * max(L - 1, 0) + (chr(33 + 45) if L else "")
        lines.append(" ".join([r, bq, ins, dl, chr(33 + gcp) * L]))
    lines.extend(haps)
    return "\n".join(lines) + "\n"


def test_edge_cases(engine, plain_engine, oracle):
    import gpupairhmm_amd as G
    rng = np.random.default_rng(3)
    rs = lambda n: "".join(rng.choice(list("ACGT"), n)) if n else ""
    recs = []
    # tile boundaries of the default table and reads around them
    bounds = [m for _, _, m in engine.tiles()]
    lens = sorted({1, 2, 3, 5, 17} | {t + d for t in bounds for d in (-1, 0, 1) if t + d > 0})
    for L in lens:
        recs.append(_record([rs(L)], [rs(max(L, 40)), rs(7)]))
    # wildcards and non-ACGT bytes on both sides, hap of length 1
    recs.append(_record(["ACGNNTACGT", "NNNN", "AC-T*GX"], ["ACGTACGTAC", "N", "acgtn", "A"]))
    # multi-tile reads (longer than the largest tile)
    recs.append(_record([rs(1500), rs(1100)], [rs(1600), rs(900)]))
    b = G.HostBatch.parse("".join(recs))
    say(f"case: edge cases, {b.num_alignments} alignments, {len(recs)} records")
    got = engine.score_batch(b)
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    assert _max_delta(got, ref) <= TOL
    # the plain engine counts what each bin launched: the last record's four
    # alignments (reads of 1500 and 1100 > 1023, the largest tile) are
    # multi-tile, and its scores equal the batched engine's
    plain_engine.reset_stats()
    got_plain = plain_engine.score_batch(b)
    assert np.array_equal(got_plain.view(np.uint32), got.view(np.uint32))
    assert plain_engine.stats().multitile_alignments >= 4


def test_record_with_many_haps(engine, plain_engine, oracle):
    """One record with more haps than the binning keeps in LDS (700 > 128):
    hap segments, item runs and -inf runs then read hap lengths from device
    memory; some haps are empty or hold a wildcard."""
    import gpupairhmm_amd as G
    rng = np.random.default_rng(9)
    rs = lambda n: "".join(rng.choice(list("ACGT"), n)) if n else ""
    reads = [rs(int(rng.integers(100, 161))) for _ in range(19)]
    haps = []
    for h in range(700):
        n = int(rng.integers(10, 61))
        seq = rs(n)
        if h % 97 == 5:
            seq = ""
        elif h % 53 == 7:
            seq = seq[:n // 2] + "N" + seq[n // 2 + 1:]
        haps.append(seq)
    b = G.HostBatch.parse(_record(reads, haps))
    say(f"case: one record, 19 reads x 700 haps, {b.num_alignments} alignments")
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    got = engine.score_batch(b)
    assert _max_delta(got, ref) <= TOL
    with G.Engine(small_batch=0) as e2:
        tuned = e2.score_batch(b)
    plain = plain_engine.score_batch(b)
    assert _same_but_rescued(got, tuned)
    assert np.array_equal(got.view(np.uint32), plain.view(np.uint32))


def test_region_likelihoods_match_record_scoring(engine, oracle):
    """The in-process adapter (reads with phred qualities x haps in, a
    log10 likelihood matrix out) gives exactly the scores of the same region
    written as one record, and those match the oracle."""
    import gpupairhmm_amd as G
    rng = np.random.default_rng(21)
    rs = lambda n: "".join(rng.choice(list("ACGTN"), n, p=[.24, .24, .24, .24, .04]))
    reads, lines = [], []
    for _ in range(7):
        L = int(rng.integers(1, 220))
        q = [rng.integers(5, 45, L) for _ in range(4)]
        bases = rs(L)
        reads.append((bases, *q))
        lines.append(" ".join([bases] + ["".join(chr(33 + int(x)) for x in a) for a in q]))
    haps = [rs(int(rng.integers(1, 400))) for _ in range(5)]
    say("case: region likelihoods, 7 reads x 5 haps")
    got = engine.compute_likelihoods(reads, haps)
    b = G.HostBatch.parse("\n".join([f"{len(reads)} {len(haps)}"] + lines + haps) + "\n")
    ref32 = engine.score_batch(b).astype(np.float64).reshape(len(reads), len(haps))
    assert got.shape == (7, 5)
    assert np.array_equal(got, ref32)
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64).reshape(len(reads), len(haps))
    assert _max_delta(got.ravel(), ref.ravel()) <= TOL


def test_empty_reads_and_haps(engine):
    import gpupairhmm_amd as G
    text = "2 2\nACGT IIII IIIN IIIN ++++\n    \nACGTAC\n\n"
    # an empty read line is "    " (5 empty fields): length 0
    b = G.HostBatch.parse(text)
    say("case: empty read and empty hap")
    got = engine.score_batch(b)
    assert got.shape == (4,)
    assert np.isfinite(got[0])          # read 0 x hap 0
    assert np.isneginf(got[1])          # read 0 x empty hap
    assert np.isneginf(got[2]) and np.isneginf(got[3])  # empty read


def _low_scoring_batch(seed=11, n_reads=3, read_len=300, hap_len=3000):
    import gpupairhmm_amd as G
    rng = np.random.default_rng(seed)
    hap = "".join(rng.choice(list("ACGT"), hap_len))
    # low base quality and random reads against the hap: log10 likelihood
    # far below the FP32 range of the reference's scaling (~ -75)
    reads = ["".join(rng.choice(list("ACGT"), read_len)) for _ in range(n_reads)]
    return G.HostBatch.parse(_record(reads, [hap], q=[3, 4, 5, 6]))


def test_fp64_rescue(lib, oracle):
    """FP32 sums below the rescue threshold are recomputed in FP64 on the GPU
    (the FP64 tile kernel of the alignment's bin)."""
    import gpupairhmm_amd as G
    b = _low_scoring_batch()
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    say(f"case: FP64 rescue, oracle scores {ref}")
    assert np.all(ref < -80)
    with G.Engine(fp64_rescue=True, range_rescue=False) as e:
        got = e.score_batch(b)
        st = e.stats()
    assert st.rescued == st.f64_alignments == b.num_alignments
    assert _max_delta(got, ref) <= TOL
    with G.Engine() as e:  # in-pass scaling first
        got = e.score_batch(b)
        st = e.stats()
    assert st.rescued + st.scaled_alignments == b.num_alignments
    assert _max_delta(got, ref) <= TOL
    with G.Engine(fp64_rescue=False) as e:
        raw = e.score_batch(b)
    assert _max_delta(raw, ref) > TOL  # FP32 alone cannot represent these


@pytest.mark.parametrize("config,records", [(1, 200), (3, 30)])
def test_fp32_accurate_above_rescue_threshold(lib, oracle, config, records):
    """The default rescue threshold is FLT_MIN (the reference's FP32 floor,
    log10 sum ~ -75.3 in its scaled units).  Without any rescue the FP32
    kernels must already be accurate for every alignment whose sum stays
    above it -- with two decades of margin -- and fail below it, which is
    what the rescue is for."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(config, seed=3000 + config, num_records=records)
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    say(f"case: FP32 accuracy above the threshold, config {config}")
    with G.Engine(fp64_rescue=False) as e:
        raw = e.score_batch(b)
    floor = np.log10(G.FLT_MIN) - np.log10(2.1267646664908054e37)  # ~ -75.26
    above = ref >= floor + 2.0
    assert above.sum() > 0.5 * len(ref)
    assert _max_delta(raw[above], ref[above]) <= 1e-4
    with G.Engine() as e:
        got = e.score_batch(b)
        st = e.stats()
    assert _max_delta(got, ref) <= TOL
    assert 0 < st.rescued + st.scaled_alignments < len(ref)


@pytest.mark.parametrize("read_len,hap_len", [(800, 2000), (1300, 1500)])
def test_fp64_rescue_very_low_scores(lib, oracle, read_len, hap_len):
    """Scores far below the FP32 range (-500 and lower): the FP64 tile pass
    (start 2^1000/H) and, past FP64's own range or for reads longer than the
    largest tile, the FP64 wave kernel with power-of-two exponent tracking."""
    import gpupairhmm_amd as G
    b = _low_scoring_batch(seed=12, n_reads=4, read_len=read_len, hap_len=hap_len)
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    say(f"case: very low scores (read {read_len}), oracle scores {ref}")
    assert np.all(ref < -300)
    for rr in (False, True):  # FP64 alone; in-pass scaling first
        with G.Engine(range_rescue=rr) as e:
            got = e.score_batch(b)
            st = e.stats()
        assert st.rescued + st.scaled_alignments == b.num_alignments
        assert np.all(np.isfinite(got))
        assert _max_delta(got, ref) <= TOL


def test_device_api_matches_host_api(engine):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(1, seed=9, num_records=6)
    say("case: device-resident batch")
    host = engine.score_batch(b)
    db = engine.upload(b)
    dev = engine.score_resident(db)
    db.free()
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))


@pytest.mark.parametrize("cfg,records", [(1, 10000), (2, 120)])
def test_host_pieces_match_device(lib, cfg, records):
    """A host batch of about 1e6 alignments is staged, copied and scored in
    pieces on two call contexts and streams; every piece scores on the whole
    batch's tile table, so the scores equal one device-resident call's bit
    for bit (pageable and pinned host arrays alike)."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(cfg, seed=70 + cfg, num_records=records)
    assert b.num_alignments > 750000  # four pieces
    say(f"case: host pieces vs device call, config {cfg}, {b.num_alignments} alignments")
    with G.Engine() as e:
        db = e.upload(b)
        dev = e.score_resident(db)
        db.free()
        host = e.score_batch(b)
        pb = e.pinned_copy(b)
        pinned = e.score_batch(pb)
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))
    assert np.array_equal(pinned.view(np.uint32), dev.view(np.uint32))


def _serve(srv, reqs, tags=None):
    """Submit every request, then collect all results: {tag: scores}."""
    tags = list(range(len(reqs))) if tags is None else tags
    for t, r in zip(tags, reqs):
        assert srv.submit(r, t)
    got = {}
    while len(got) < len(reqs):
        res = srv.poll(timeout_ms=20000)
        assert res is not None, "server timed out"
        tag, scores, lat = res
        assert lat >= 0
        got[tag] = scores
    return got


@pytest.mark.parametrize("streams,max_batch_requests", [(2, 4096), (1, 4096), (2, 1), (3, 2)])
def test_server_alone_equals_mix(engine, streams, max_batch_requests):
    """Each request's scores are bit-identical scored alone (phmm_score_batch)
    and served in a mix, whatever the merge and however many launches
    overlap on the device."""
    import gpupairhmm_amd as G
    reqs = [G.HostBatch.synthetic(4, seed=500 + i, num_records=1) for i in range(8)]
    say(f"case: serving alone vs mix (streams {streams}, max_batch_requests {max_batch_requests})")
    alone = [engine.score_batch(r) for r in reqs]
    srv = G.Server(engine, streams=streams, max_batch_requests=max_batch_requests)
    try:
        got = _serve(srv, reqs)
        st = srv.stats()
    finally:
        srv.close()
    for i in range(len(reqs)):
        assert np.array_equal(alone[i].view(np.uint32), got[i].view(np.uint32))
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 1 <= st["max_concurrent"] <= streams
    if max_batch_requests == 1:  # every request is its own launch
        assert st["launches"] == len(reqs)


def test_server_launches_overlap_and_merge(engine):
    """A burst larger than one launch: requests merge, two launches are on
    the device at once, and every score matches the request scored alone."""
    import gpupairhmm_amd as G
    reqs = [G.HostBatch.synthetic(4, seed=600 + i, num_records=1) for i in range(40)]
    alone = [engine.score_batch(r) for r in reqs]
    cap = max(r.num_alignments for r in reqs) * 3
    srv = G.Server(engine, streams=2, max_batch_alignments=cap)
    say(f"case: serving burst of {len(reqs)} requests, {cap} alignments per launch")
    try:
        got = _serve(srv, reqs)
        st = srv.stats()
    finally:
        srv.close()
    for i in range(len(reqs)):
        assert np.array_equal(alone[i].view(np.uint32), got[i].view(np.uint32))
    assert st["launches"] < len(reqs)           # merged
    assert st["max_concurrent"] == 2            # overlapped
    assert st["bytes_in_use"] == st["scratch_bytes"]  # every arena block released


def test_server_small_arena_stream(engine):
    """A stream of requests through an arena that holds only a few of them:
    submits are refused while it is full and admitted again as results come
    back, arena blocks wrap around, and requests staged while a copy is in
    flight are copied together (one copy per run contiguous in staging and
    arena, split where the arena wraps).  Every result equals the request
    scored alone."""
    import gpupairhmm_amd as G
    reqs = [G.HostBatch.synthetic(4, seed=800 + i, num_records=1) for i in range(48)]
    alone = [engine.score_batch(r) for r in reqs]
    caps = dict(max_batch_alignments=1 << 16, max_batch_bytes=8 << 20, max_batch_requests=8)
    probe = G.Server(engine, **caps)
    try:
        fixed = probe.stats()["scratch_bytes"]
    finally:
        probe.close()
    big = max(reqs, key=lambda b: b.num_alignments)
    budget = fixed + 4 * (16 * big.num_alignments + (1 << 16))
    srv = G.Server(engine, hbm_budget_bytes=budget, **caps)
    say(f"case: serving stream of {len(reqs)} requests through a "
        f"{(budget - fixed) >> 10} KiB arena")
    got, nxt, refused = {}, 0, 0
    try:
        while len(got) < len(reqs):
            while nxt < len(reqs) and srv.submit(reqs[nxt], nxt):
                nxt += 1
            if nxt < len(reqs):
                refused += 1
            res = srv.poll(timeout_ms=20000)
            assert res is not None, "server timed out"
            got[res[0]] = res[1]
        st = srv.stats()
    finally:
        srv.close()
    for i in range(len(reqs)):
        assert np.array_equal(alone[i].view(np.uint32), got[i].view(np.uint32)), i
    assert refused > 0 and st["refused"] == refused  # the arena was full at times
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["bytes_in_use"] == st["scratch_bytes"]


def test_server_budget_refusal(engine):
    """A budget that holds the launch contexts and one request but not two:
    the second request is refused with PHMM_EBUSY and admitted again once the
    first one completed."""
    import gpupairhmm_amd as G
    # a large request (about 30k alignments of long reads and haps), so it is
    # still on the device when the second submit follows
    r = max((G.HostBatch.synthetic(4, seed=700 + i, num_records=1) for i in range(12)),
            key=lambda b: b.cells())
    caps = dict(max_batch_alignments=1 << 16, max_batch_bytes=8 << 20, max_batch_requests=8)
    probe = G.Server(engine, **caps)
    try:
        fixed = probe.stats()["scratch_bytes"]
        assert probe.submit(r, 0)
        block = probe.stats()["bytes_in_use"] - fixed
        assert probe.poll(timeout_ms=20000) is not None
    finally:
        probe.close()
    assert block >= r.num_alignments  # the request's image is in the arena
    budget = fixed + block + block // 2
    srv = G.Server(engine, hbm_budget_bytes=budget, **caps)
    say(f"case: serving budget refusal (budget {budget}, request block {block})")
    try:
        st = srv.stats()
        assert st["scratch_bytes"] == fixed and st["arena_bytes"] < 2 * block
        assert srv.submit(r, 1)
        assert not srv.submit(r, 2)  # PHMM_EBUSY: the arena holds one request
        assert srv.stats()["refused"] == 1
        res = srv.poll(timeout_ms=20000)
        assert res is not None and res[0] == 1
        assert srv.submit(r, 3)  # admitted after the first completed
        res = srv.poll(timeout_ms=20000)
        assert res is not None and res[0] == 3
        assert srv.stats()["bytes_in_use"] == fixed
    finally:
        srv.close()
    # a request larger than one launch is refused outright
    big = G.HostBatch.synthetic(4, seed=1, num_records=8)
    srv = G.Server(engine, max_batch_alignments=1000)
    try:
        with pytest.raises(G.PairHMMError) as e:
            srv.submit(big, 9)
        assert e.value.code == G.ERANGE
    finally:
        srv.close()
