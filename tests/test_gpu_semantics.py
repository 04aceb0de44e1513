"""Both input semantics of the reference on the GPU, against the FP64 oracle:

- PHMM_SEM_GPU: the reference GPU kernel -- eps = 0.1, beta = 0.9 fixed
  (pairhmm_kernels.cuh:1195-1196) and every byte other than A/C/G/T is the N
  wildcard (utility_kernels.cuh:16-20);
- PHMM_SEM_CPU: the reference align_host -- eps/beta from the read's first
  gcp quality (gpuworker.cuh:1009, 1048-1049) and only a literal 'N' is a
  wildcard, every other byte matching only itself (gpuworker.cuh:1063).

Records mix plain ACGTN alignments (tile kernels) with reads and haps that
hold other bytes (IUPAC codes, lowercase, punctuation), reads beyond the
largest tile, and gap-continuation qualities other than '+'."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _seq(rng, n, alphabet):
    return "".join(rng.choice(list(alphabet), n))


def _batch(seed=21):
    import gpupairhmm_amd as G
    rng = np.random.default_rng(seed)
    lines = []
    for rec in range(6):
        reads, haps = [], []
        for k in range(5):
            if rec == 5 and k == 0:
                L = int(rng.integers(1100, 1400))   # beyond the largest tile
            else:
                L = int(rng.integers(20, 260))
            alphabet = "ACGT" if k < 2 else ("ACGTN" if k == 2 else "ACGTNRYacgn*")
            r = _seq(rng, L, alphabet)
            q = "".join(chr(33 + int(x)) for x in rng.integers(10, 41, L))
            gi = "".join(chr(33 + int(x)) for x in rng.integers(25, 46, L))
            gcp = chr(33 + int(rng.integers(5, 40))) * L
            reads.append(" ".join([r, q, gi, gi, gcp]))
        for k in range(3):
            H = int(rng.integers(150, 500))
            haps.append(_seq(rng, H, "ACGT" if k < 2 else "ACGTNRYt"))
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


@pytest.mark.parametrize("semantics", [0, 1])
def test_semantics_against_oracle(lib, oracle, semantics):
    import gpupairhmm_amd as G
    b = _batch()
    ref = oracle.score_batch(b.oracle_arrays(), oracle.F64, semantics=semantics)
    outs = []
    for mode in (G.MODE_PLAIN, G.MODE_BATCHED):
        say(f"case: semantics {semantics}, mode {mode}, {b.num_alignments} alignments")
        with G.Engine(mode=mode, semantics=semantics) as e:
            got = e.score_batch(b)
            st = e.stats()
        assert np.all(np.isfinite(got))
        assert _max_delta(got, ref) <= 1e-3
        outs.append(got)
        if mode == G.MODE_PLAIN:
            # haps with a byte other than A/C/G/T (and, under CPU semantics,
            # reads with a byte other than A/C/G/T/N) go to the exact kernel
            assert st.multitile_alignments > b.num_alignments // 4
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_semantics_differ_where_the_reference_paths_differ(lib, oracle):
    """On this input the two reference paths disagree (gcp and wildcards),
    and each engine mode follows its own reference."""
    import gpupairhmm_amd as G
    b = _batch(seed=22)
    gpu = oracle.score_batch(b.oracle_arrays(), oracle.F64, semantics=0)
    cpu = oracle.score_batch(b.oracle_arrays(), oracle.F64, semantics=1)
    assert _max_delta(gpu, cpu) > 0.1
    with G.Engine(semantics=1) as e:
        got = e.score_batch(b)
    assert _max_delta(got, cpu) <= 1e-3
    assert _max_delta(got, gpu) > 0.1
