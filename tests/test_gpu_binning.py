"""The read pass's choice of arithmetic form, checked against a numpy
restatement of its rule: a read takes the rescaled 5-op form when every row
has 0 < alpha <= 1 with delta, sigma > 0, every row below the first has
eps*delta_{r-1} <= 1/2 * (delta_r*alpha_{r-1}) (float arithmetic), and the
product of alpha over the read (double) is at least 1e-30.  Plain mode counts
the alignments of rescaled-form bins, so the count must match exactly."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _batch(G, rng):
    lines, reads_meta = [], []
    for rec in range(40):
        nr, nh = 6, 2
        lines.append(f"{nr} {nh}")
        for k in range(nr):
            L = int(rng.integers(20, 300))
            seq = "".join(rng.choice(list("ACGT"), L))
            bq = rng.integers(10, 41, L)
            mode = int(rng.integers(0, 4))
            if mode == 0:      # constant high gap qualities: rescaled form
                iq = np.full(L, 40)
                dq = np.full(L, 40)
            elif mode == 1:    # jumpy insertion qualities: e_r > 1/2 somewhere
                iq = rng.integers(5, 60, L)
                dq = np.full(L, 40)
            elif mode == 2:    # gap probabilities summing past 1: alpha <= 0
                iq = np.full(L, 40)
                dq = np.full(L, 40)
                iq[L // 2] = 1
                dq[L // 2] = 1
            else:              # gentle variation
                iq = rng.integers(38, 42, L)
                dq = rng.integers(38, 42, L)
            q = lambda a: "".join(chr(33 + int(x)) for x in a)
            lines.append(" ".join([seq, q(bq), q(iq), q(dq), "+" * L]))
            reads_meta.append((rec, iq, dq))
        for h in range(nh):
            lines.append("".join(rng.choice(list("ACGT"), 300)))
    return G.HostBatch.parse("\n".join(lines) + "\n"), reads_meta


def _fast(P, iq, dq):
    f32 = np.float32
    eps = f32(0.1)
    delta = P[np.minimum(iq, 127)].astype(f32)
    sigma = P[np.minimum(dq, 127)].astype(f32)
    alpha = (f32(1) - (delta + sigma)).astype(f32)
    if not np.all((alpha > 0) & (alpha <= 1) & (delta > 0) & (sigma > 0)):
        return False
    num = (eps * delta[:-1]).astype(f32)
    den = (delta[1:] * alpha[:-1]).astype(f32)
    if not np.all(num <= (f32(0.5) * den).astype(f32)):
        return False
    return float(np.prod(alpha.astype(np.float64))) >= 1e-30


def test_fast_form_classification(lib, oracle):
    import gpupairhmm_amd as G
    rng = np.random.default_rng(11)
    b, meta = _batch(G, rng)
    P = oracle.ph2pr()
    want = sum(2 for _, iq, dq in meta if _fast(P, iq, dq))  # 2 haps per record
    say(f"case: form classification, {b.num_alignments} alignments, {want} expected rescaled")
    assert 0 < want < b.num_alignments
    with G.Engine(mode=G.MODE_PLAIN) as e:
        e.score_batch(b)
        st = e.stats()
    assert st.fast_alignments == want
