"""Pins the oracle's per-position gap-continuation semantics (SEM_GCP): row r
(1-based) takes eps_r = ph2pr[gcp[r-1]] and beta_r = 1 - eps_r where the
reference's align_host takes one eps and beta from the read's first gcp
quality (gpuworker.cuh:1009, 1048-1049), indexed like the row's insertion
and deletion qualities (gpuworker.cuh:1059-1061).

No reference code computes this variant, so it is pinned by:
  * brute-force enumeration of every state path with per-row transitions
    (no dynamic programming), on tiny inputs with varying gcp;
  * the brute force's power: a dynamic program with each plausible index
    slip (I chain, D chain or gap-to-match taking a neighbouring row's gcp)
    disagrees with it;
  * closed forms for R = 1 and H = 1, which isolate beta_1 and the I chain's
    eps_3 .. eps_R;
  * bit-exact equality with the reference CPU semantics (itself pinned
    bit-exact against the compiled align_host) whenever a read's gcp
    qualities are all equal, in both oracle precisions.
"""
import math

import numpy as np
import pytest


def P(q):
    return float(np.float32(10.0 ** (-min(int(q), 127) / 10.0)))


def _case(rng, R, H, alphabet="ACGTN", glo=2, ghi=40):
    read = "".join(rng.choice(list(alphabet), R))
    hap = "".join(rng.choice(list(alphabet), H))
    bq = rng.integers(5, 41, R)
    iq = rng.integers(10, 46, R)
    dq = rng.integers(10, 46, R)
    gq = rng.integers(glo, ghi + 1, R)
    return hap, read, bq, iq, dq, gq


def _brute(oracle, hap, read, bq, iq, dq, gq):
    eps = [P(g) for g in gq]
    beta = [1.0 - e for e in eps]
    return oracle.brute_force(hap.encode(), read.encode(), bq, iq, dq, eps, beta, literal_n=True)


@pytest.mark.parametrize("seed", range(16))
def test_gcp_matches_brute_force_paths(oracle, seed):
    rng = np.random.default_rng(500 + seed)
    R, H = int(rng.integers(1, 5)), int(rng.integers(1, 5))
    hap, read, bq, iq, dq, gq = _case(rng, R, H)
    got = oracle.align_gcp(hap, read, bq, iq, dq, gq, oracle.F64, oracle.SEM_GCP)
    want = _brute(oracle, hap, read, bq, iq, dq, gq)
    # the brute force takes exact phred values for the base and indel
    # qualities, the oracle the FP32 table: agreement to ~1e-8
    assert abs(got - want) < 1e-7, (R, H, list(gq), got, want)


def _dp(read, hap, bq, iq, dq, gq, slip=None):
    """A direct double-precision forward DP with per-row gap continuation, in
    which one index can be planted wrong: 'I' (the I chain takes eps_{r-1}),
    'D' (the D chain takes eps_{r+1}), 'B' (gap-to-match takes beta_{r-1})."""
    R, H = len(read), len(hap)
    eps = [P(g) for g in gq]

    def e(r):  # 1-based, clamped at the read's ends
        return eps[min(max(r, 1), R) - 1]
    M = [[0.0] * (H + 1) for _ in range(R + 1)]
    I = [[0.0] * (H + 1) for _ in range(R + 1)]
    D = [[0.0] * (H + 1) for _ in range(R + 1)]
    for c in range(H + 1):
        D[0][c] = 1.0 / H
    for r in range(1, R + 1):
        p, de, si = P(bq[r - 1]), P(iq[r - 1]), P(dq[r - 1])
        al = 1.0 - (de + si)
        ei = e(r - 1) if slip == "I" else e(r)
        ed = e(r + 1) if slip == "D" else e(r)
        b = 1.0 - (e(r - 1) if slip == "B" else e(r))
        for c in range(1, H + 1):
            x, h = read[r - 1], hap[c - 1]
            lam = 1 - p if (x == h or x == "N" or h == "N") else p / 3
            M[r][c] = lam * (al * M[r - 1][c - 1] + b * (I[r - 1][c - 1] + D[r - 1][c - 1]))
            I[r][c] = de * M[r - 1][c] + ei * I[r - 1][c]
            D[r][c] = si * M[r][c - 1] + ed * D[r][c - 1]
    return math.log10(sum(M[R][c] + I[R][c] for c in range(1, H + 1)))


def test_brute_force_detects_planted_gcp_index_slips(oracle):
    # a read that is the hap with a two-base deletion after row 2 (the D
    # chain continues on row 2) and gaps made likely by low indel qualities;
    # gcp alternates between long and short gaps, so every neighbouring
    # row's gcp differs
    hap, read = "ACGGTA", "ACTA"
    bq = np.array([30, 30, 30, 30])
    iq = np.array([12, 11, 13, 12])
    dq = np.array([11, 12, 10, 13])
    gq = np.array([30, 3, 30, 10])
    bf = _brute(oracle, hap, read, bq, iq, dq, gq)
    assert abs(_dp(read, hap, bq, iq, dq, gq) - bf) < 1e-7
    for slip in ("I", "D", "B"):
        assert abs(_dp(read, hap, bq, iq, dq, gq, slip) - bf) > 1e-4, slip
    # the oracle itself, fed gcp shifted by one row or collapsed to the
    # first value (the reference CPU semantics), disagrees as well
    shifted = np.concatenate([gq[1:], gq[-1:]])
    a = oracle.align_gcp(hap, read, bq, iq, dq, shifted, oracle.F64, oracle.SEM_GCP)
    c = oracle.align_gcp(hap, read, bq, iq, dq, gq, oracle.F64, oracle.SEM_CPU)
    assert abs(a - bf) > 1e-4 and abs(c - bf) > 1e-4


def test_gcp_closed_form_single_read_row(oracle):
    """R = 1: score = log10(beta_1 * sum_c lambda(1, c) / H), beta_1 from gcp[0]."""
    rng = np.random.default_rng(23)
    for _ in range(20):
        H = int(rng.integers(1, 30))
        hap, read, bq, iq, dq, gq = _case(rng, 1, H, alphabet="ACGT")
        p = P(bq[0])
        lam = [(1 - p) if read[0] == h else p / 3 for h in hap]
        want = math.log10((1.0 - P(gq[0])) * sum(lam) / H)
        got = oracle.align_gcp(hap, read, bq, iq, dq, gq, oracle.F64, oracle.SEM_GCP)
        assert abs(got - want) < 1e-12


def test_gcp_closed_form_single_hap_base(oracle):
    """H = 1: the only paths match read base 1 and then insert, so the score
    is log10(eps_3 ... eps_R * delta_2 * beta_1 * lambda(1, 1))."""
    rng = np.random.default_rng(29)
    for R in range(2, 12):
        hap, read, bq, iq, dq, gq = _case(rng, R, 1, alphabet="ACGT")
        p = P(bq[0])
        lam = (1 - p) if read[0] == hap[0] else p / 3
        prod = 1.0
        for r in range(3, R + 1):
            prod *= P(gq[r - 1])
        want = math.log10(prod * P(iq[1]) * (1.0 - P(gq[0])) * lam)
        got = oracle.align_gcp(hap, read, bq, iq, dq, gq, oracle.F64, oracle.SEM_GCP)
        assert abs(got - want) < 1e-10, (R, got, want)


@pytest.mark.parametrize("precision", [0, 1])
def test_gcp_equals_cpu_semantics_for_constant_gcp(oracle, precision):
    rng = np.random.default_rng(31 + precision)
    for _ in range(60):
        R, H = int(rng.integers(1, 50)), int(rng.integers(1, 70))
        hap, read, bq, iq, dq, _ = _case(rng, R, H, alphabet="ACGTNRy*")
        g = int(rng.integers(2, 45))
        gq = np.full(R, g)
        a = oracle.align_gcp(hap, read, bq, iq, dq, gq, precision, oracle.SEM_GCP)
        b = oracle.align(hap, read, bq, iq, dq, g, precision, oracle.SEM_CPU)
        assert a == b or (math.isinf(a) and math.isinf(b)), (a, b)


def test_gcp_needs_the_read_gcp_array(oracle):
    # the scalar form cannot serve per-position gcp
    assert math.isnan(oracle.align("ACGT", "ACGT", [30] * 4, [40] * 4, [40] * 4, 10,
                                   oracle.F64, oracle.SEM_GCP))


def test_gcp_batch_driver_reads_each_row_gcp(oracle, lib):
    """The batch driver passes every read's own gcp array: batch scores equal
    per-alignment scores, and differ from the first-gcp semantics."""
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(0, seed=41, num_records=3)
    arrays = {k: b.array(k).copy() for k in ("read_data", "hap_data", "base_quals", "ins_quals",
                                             "del_quals", "gcp_quals", "read_lengths",
                                             "hap_lengths", "read_offsets", "hap_offsets",
                                             "num_reads_per_group", "num_haps_per_group",
                                             "reads_prefix", "haps_prefix")}
    rng = np.random.default_rng(43)
    arrays["gcp_quals"] = rng.integers(2, 41, arrays["gcp_quals"].size).astype(np.uint8)
    v = G.HostBatch.from_arrays(**arrays)
    got = oracle.score_batch(v.oracle_arrays(), oracle.F64, oracle.SEM_GCP)
    cpu = oracle.score_batch(v.oracle_arrays(), oracle.F64, oracle.SEM_CPU)
    ro, rl, ho, hl = (arrays[k] for k in ("read_offsets", "read_lengths", "hap_offsets",
                                          "hap_lengths"))
    k = 0
    for g in range(v.num_groups):
        r0, nr = arrays["reads_prefix"][g], arrays["num_reads_per_group"][g]
        h0, nh = arrays["haps_prefix"][g], arrays["num_haps_per_group"][g]
        for r in range(r0, r0 + nr):
            sl = slice(ro[r], ro[r] + rl[r])
            for h in range(h0, h0 + nh):
                hap = bytes(arrays["hap_data"][ho[h]: ho[h] + hl[h]]).decode()
                read = bytes(arrays["read_data"][sl]).decode()
                want = oracle.align_gcp(hap, read, arrays["base_quals"][sl], arrays["ins_quals"][sl],
                                        arrays["del_quals"][sl], arrays["gcp_quals"][sl],
                                        oracle.F64, oracle.SEM_GCP)
                assert got[k] == want
                k += 1
    assert k == len(got)
    assert np.max(np.abs(got - cpu)) > 1e-3
