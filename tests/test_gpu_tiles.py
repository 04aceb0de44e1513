"""Every compiled tile shape against the FP64 oracle: each rows-per-lane
instantiation (multiples of 4 from 4 to 32, and 13 and 19, whose last LDS
vector is partly unused) with several lane counts, both arithmetic forms (the
rescaled 5-op form and the 6-op form the read pass picks when a read's
coefficients would leave the rescaled range), and the FP64 tile kernel."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

ROWS = (4, 8, 12, 13, 16, 19, 20, 24, 28, 32)


def _batch(G, max_len, seed):
    rng = np.random.default_rng(seed)
    lines = []
    for rec in range(3):
        reads = []
        for k in range(6):
            L = int(rng.integers(max(1, max_len // 2), max_len + 1)) if k else max_len
            r = "".join(rng.choice(list("ACGT"), L))
            q = "".join(chr(33 + int(x)) for x in rng.integers(10, 41, L))
            if k % 2:  # alternating gap qualities: the 6-op form
                gi = "".join(chr(33 + (10 if i % 2 else 40)) for i in range(L))
            else:
                gi = "I" * (L - 1) + "N"
            reads.append(" ".join([r, q, gi, gi, "+" * L]))
        haps = ["".join(rng.choice(list("ACGT"), int(rng.integers(max_len, max_len + 120))))
                for _ in range(3)]
        # one hap that contains the first read
        haps.append("ACGT" + reads[0].split()[0] + "TTGCA")
        lines.append(f"{len(reads)} {len(haps)}")
        lines += reads + haps
    return G.HostBatch.parse("\n".join(lines) + "\n")


@pytest.mark.parametrize("rows", ROWS)
def test_tile_shapes(lib, oracle, rows):
    import gpupairhmm_amd as G
    for lanes in (1, 3, 7, 16, 64):
        max_len = lanes * rows - 1
        if max_len < 1 or max_len > 1023:
            continue
        b = _batch(G, max_len, seed=rows * 100 + lanes)
        say(f"case: tile {lanes}x{rows}, reads up to {max_len}, {b.num_alignments} alignments")
        ref = oracle.score_batch(b.oracle_arrays(), oracle.F64)
        with G.Engine(mode=G.MODE_PLAIN, small_batch=0) as e:  # this table, at any size
            e.set_tiles([(lanes, rows, max_len)])
            got = e.score_batch(b)
            st = e.stats()
        with np.errstate(invalid="ignore"):
            d = np.abs(got.astype(np.float64) - ref)
        d[np.isneginf(got) & np.isneginf(ref)] = 0
        assert np.max(d) <= 1e-3, (lanes, rows, float(np.max(d)))
        assert 0 < st.fast_alignments < b.num_alignments  # both forms ran
