"""Per-position gap continuation (PHMM_SEM_GCP) on the GPU, against the FP64
oracle: row r takes eps_r = ph2pr[gcp[r-1]] and beta_r = 1 - eps_r (the
reference's align_host recurrence, gpuworker.cuh:1043-1094, with every row's
own gcp where align_host takes the read's first, gpuworker.cuh:1009,
1048-1049); literal-'N' wildcard (gpuworker.cuh:1063).

Records mix reads whose gcp varies along the read (gcp 2..40, so some rows
have eps > 1/2 and those reads take the SAFE form), constant-gcp reads,
reads holding N and IUPAC bytes, a read beyond the largest tile and
low-quality reads whose FP32 sums underflow (FP64 rescue with per-row eps)."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _seq(rng, n, alphabet):
    return "".join(rng.choice(list(alphabet), n))


def _batch(seed):
    import gpupairhmm_amd as G
    rng = np.random.default_rng(seed)
    lines = []
    for rec in range(8):
        reads, haps = [], []
        nh = int(rng.integers(1, 6))
        for k in range(int(rng.integers(1, 12))):
            if rec == 7 and k == 0:
                L = int(rng.integers(1100, 1300))  # beyond the largest tile
            else:
                L = int(rng.integers(8, 300))
            alphabet = "ACGTN" if k % 4 != 3 else "ACGTNRYacgn*"
            r = _seq(rng, L, alphabet)
            lo = 2 if rec % 3 == 2 else 10  # low base qualities: FP64 rescue
            q = "".join(chr(33 + int(x)) for x in rng.integers(lo, 41, L))
            gi = "".join(chr(33 + int(x)) for x in rng.integers(20, 46, L))
            gd = "".join(chr(33 + int(x)) for x in rng.integers(20, 46, L))
            if k % 3 == 0:
                gcp = chr(33 + int(rng.integers(2, 40))) * L
            else:
                gcp = "".join(chr(33 + int(x)) for x in rng.integers(2, 41, L))
            reads.append(" ".join([r, q, gi, gd, gcp]))
        for k in range(nh):
            H = int(rng.integers(20, 700))
            haps.append(_seq(rng, H, "ACGT" if k < 3 else "ACGTNRYt"))
        lines.append(f"{len(reads)} {len(haps)}")
        lines += reads + haps
    return G.HostBatch.parse("\n".join(lines) + "\n")


def _max_delta(got, ref):
    got = np.asarray(got, dtype=np.float64)
    both = np.isneginf(got) & np.isneginf(ref)
    with np.errstate(invalid="ignore"):
        d = np.abs(got - ref)
    d[both] = 0
    return float(np.max(d))


@pytest.mark.parametrize("seed", [51, 52])
def test_gcp_semantics_against_oracle(lib, oracle, seed):
    import gpupairhmm_amd as G
    b = _batch(seed)
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64, semantics=oracle.SEM_GCP)
    cpu = oracle.score_batch(b.oracle_arrays(), oracle.F64, semantics=oracle.SEM_CPU)
    assert _max_delta(ref, cpu) > 1e-2  # the batch tells the two semantics apart
    outs = []
    for mode in (G.MODE_PLAIN, G.MODE_BATCHED):
        say(f"case: gcp semantics seed {seed}, mode {mode}, {b.num_alignments} alignments")
        with G.Engine(mode=mode, semantics=G.SEM_GCP) as e:
            got = e.score_batch(b)
            st = e.stats()
        assert np.all(np.isfinite(got))
        d = _max_delta(got, ref)
        print(f"gcp seed {seed} mode {mode}: max |delta log10| {d:.2e}, "
              f"rescued {st.rescued}, fast {st.fast_alignments}")
        assert d <= 1e-3
        outs.append(got)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_gcp_device_path_and_rescue(lib, oracle):
    """Device-resident batched calls (captured graph) on a batch whose FP32
    sums underflow, with varying gcp: the FP64 tiles with per-row eps."""
    import gpupairhmm_amd as G
    base = G.HostBatch.synthetic(3, seed=53, num_records=2)
    arrays = {k: base.array(k).copy() for k in (
        "read_data", "hap_data", "base_quals", "ins_quals", "del_quals", "gcp_quals",
        "read_lengths", "hap_lengths", "read_offsets", "hap_offsets", "num_reads_per_group",
        "num_haps_per_group", "reads_prefix", "haps_prefix")}
    rng = np.random.default_rng(54)
    arrays["gcp_quals"] = rng.integers(2, 41, arrays["gcp_quals"].size).astype(np.uint8)
    b = G.HostBatch.from_arrays(**arrays)
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64, semantics=oracle.SEM_GCP)
    say(f"case: gcp device path, {b.num_alignments} alignments")
    with G.Engine(semantics=G.SEM_GCP) as e:
        db = e.upload(b)
        got = e.score_resident(db)
        again = e.score_resident(db)
        st = e.stats()
        db.free()
    assert st.rescued > 0
    assert _max_delta(got, ref) <= 1e-3
    assert np.array_equal(got.view(np.uint32), again.view(np.uint32))


def test_gcp_equals_cpu_semantics_for_constant_gcp(lib):
    """With a constant gcp per read the per-row kernels compute what the
    first-gcp kernels compute, bit for bit (same coefficients, same order;
    both with the FP64 rescue: the per-row kernels have neither in-pass
    scaling nor a range-extended pass)."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=55, num_records=6)
    say(f"case: gcp vs cpu semantics, constant gcp, {b.num_alignments} alignments")
    with G.Engine(semantics=G.SEM_GCP) as e:
        a = e.score_batch(b)
    with G.Engine(semantics=G.SEM_CPU, range_rescue=False) as e:
        c = e.score_batch(b)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
