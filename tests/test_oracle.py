"""Pins the CPU oracle (oracle/pairhmm_oracle.c) before anything is checked
against it.

Pins, each chosen so a plausible slip (a dropped term, a wrong row index,
swapped I/D, a wrong constant) fails at least one:
  * the README known-answer record (README.md:45-61), the reference's only
    golden vector;
  * bit-exact agreement of the float-cell restatement with the reference's
    own align_host compiled from its source (oracle/_ref, built when the
    reference tree is present);
  * brute-force enumeration of every state path on tiny inputs (no DP);
  * closed forms for R=1 and for H=1;
  * the phred table against 10^(-q/10) and against the reference's
    generate_ph2pr;
  * per-cell exponent epochs (extreme scores, rows wider than double's
    range) against an arbitrary-exponent Decimal evaluation.
"""
import itertools
import math
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

F32_EPS = float(np.float32(0.1))
F32_BETA = float(np.float32(0.9))


def _q(s):
    return np.frombuffer(s.encode(), dtype=np.uint8).astype(np.int64) - 33


README = [
    ("CTGTGTCCATGTCAGAGCAATGGCCCAAGTCTGGGCCTGGG", "888888888888'88888888'8888888888888'8'888",
     "IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIN", "IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIN",
     "+++++++++++++++++++++++++++++++++++++++++"),
    ("ATGTCAGAGCAATGGCCCAAGTCTGGGTCTGGG", "<AFFFKKKKKKKKKKKKKKKKKKKKKKKKKKKK",
     "IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIN", "IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIN",
     "+++++++++++++++++++++++++++++++++"),
]
README_HAPS = ["CTGTGTCCATGTCAGAGCAACGGCCCAAGTCTGGGTCTGGG",
               "CTGTGTCCATGTCAGAGCAATGGCCCAAGTCTGGGTCTGGG"]
README_SCORES = [[-5.97153, -3.1966], [-6.34043, -1.66333]]
README_DECIMALS = [[5, 4], [5, 5]]


@pytest.mark.parametrize("precision", [0, 1])
def test_readme_known_answer(oracle, precision):
    for i, r in enumerate(README):
        for j, h in enumerate(README_HAPS):
            got = oracle.align(h, r[0], _q(r[1]), _q(r[2]), _q(r[3]), int(_q(r[4])[0]),
                               precision, oracle.SEM_CPU)
            tol = 0.5 * 10.0 ** (-README_DECIMALS[i][j]) + 1e-5
            assert abs(got - README_SCORES[i][j]) <= tol, (i, j, got)


def _random_case(rng, R, H, alphabet="ACGT", qlo=5, qhi=40):
    read = "".join(rng.choice(list(alphabet), R))
    hap = "".join(rng.choice(list("ACGT"), H))
    bq = rng.integers(qlo, qhi + 1, R)
    iq = rng.integers(10, 46, R)
    dq = rng.integers(10, 46, R)
    return hap, read, bq, iq, dq


def test_float_cells_match_reference_align_host_bit_exact(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("reference align_host not built (reference tree absent)")
    rng = np.random.default_rng(7)
    for _ in range(200):
        R, H = int(rng.integers(1, 60)), int(rng.integers(1, 80))
        hap, read, bq, iq, dq = _random_case(rng, R, H, "ACGTN")
        gcp = int(rng.integers(5, 40))
        a = oracle.align(hap, read, bq, iq, dq, gcp, oracle.F32CELLS, oracle.SEM_CPU)
        b = oracle.ref_align(hap, read, bq, iq, dq, gcp)
        assert (a == b) or (math.isinf(a) and math.isinf(b)), (a, b)
    # bytes other than A/C/G/T/N in reads and haps: align_host compares raw
    # bytes with a literal 'N' wildcard (gpuworker.cuh:1063)
    for _ in range(100):
        R, H = int(rng.integers(1, 50)), int(rng.integers(1, 60))
        read = "".join(rng.choice(list("ACGTNRYacgn*"), R))
        hap = "".join(rng.choice(list("ACGTNRYacgn*"), H))
        bq = rng.integers(5, 41, R)
        iq = rng.integers(10, 46, R)
        dq = rng.integers(10, 46, R)
        gcp = int(rng.integers(5, 40))
        a = oracle.align(hap, read, bq, iq, dq, gcp, oracle.F32CELLS, oracle.SEM_CPU)
        b = oracle.ref_align(hap, read, bq, iq, dq, gcp)
        assert (a == b) or (math.isinf(a) and math.isinf(b)), (read, hap, a, b)
    # the reference's own phred table
    import ctypes
    t = np.zeros(128, dtype=np.float32)
    oracle.ref_lib().ref_ph2pr(t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert np.array_equal(t, oracle.ph2pr())


def test_ph2pr_closed_form(oracle):
    t = oracle.ph2pr()
    want = np.array([np.float32(10.0 ** (-q / 10.0)) for q in range(128)], dtype=np.float32)
    assert np.array_equal(t, want)
    assert t[0] == 1.0 and t[10] == np.float32(0.1) and t[20] == np.float32(0.01)


@pytest.mark.parametrize("seed", range(12))
def test_brute_force_paths(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    R, H = int(rng.integers(1, 5)), int(rng.integers(1, 5))
    hap, read, bq, iq, dq = _random_case(rng, R, H, "ACGTN")
    got = oracle.align(hap, read, bq, iq, dq, 10, oracle.F64, oracle.SEM_GPU)
    want = oracle.brute_force(hap.encode(), read.encode(), bq, iq, dq, F32_EPS, F32_BETA)
    assert abs(got - want) < 1e-7, (R, H, got, want)


def test_brute_force_detects_a_planted_index_slip(oracle):
    """The brute force is sensitive: a DP that uses delta of row r instead of
    row r+1 for M->I (a plausible slip) disagrees with it."""
    rng = np.random.default_rng(5)
    hap, read, bq, iq, dq = _random_case(rng, 4, 3)
    P = [10.0 ** (-q / 10.0) for q in range(128)]

    def dp(slip):
        R, H = len(read), len(hap)
        C = 1.0 / H
        M = [[0.0] * (H + 1) for _ in range(R + 1)]
        I = [[0.0] * (H + 1) for _ in range(R + 1)]
        D = [[0.0] * (H + 1) for _ in range(R + 1)]
        for c in range(H + 1):
            D[0][c] = C
        for r in range(1, R + 1):
            p, de, si = P[bq[r - 1]], P[iq[r - 1]], P[dq[r - 1]]
            de_used = P[iq[r - 2]] if (slip and r >= 2) else de
            al = 1.0 - (de + si)
            for c in range(1, H + 1):
                lam = 1 - p if read[r - 1] == hap[c - 1] else p / 3
                M[r][c] = lam * (al * M[r - 1][c - 1] + F32_BETA * (I[r - 1][c - 1] + D[r - 1][c - 1]))
                I[r][c] = de_used * M[r - 1][c] + F32_EPS * I[r - 1][c]
                D[r][c] = si * M[r][c - 1] + F32_EPS * D[r][c - 1]
        return math.log10(sum(M[R][c] + I[R][c] for c in range(1, H + 1)))

    bf = oracle.brute_force(hap.encode(), read.encode(), bq, iq, dq, F32_EPS, F32_BETA)
    assert abs(dp(False) - bf) < 1e-7
    assert abs(dp(True) - bf) > 1e-4


def test_closed_form_single_read_row(oracle):
    """R = 1: I is zero on row 1, so the score is log10(beta * sum_c lambda(1,c) / H)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        H = int(rng.integers(1, 30))
        hap, read, bq, iq, dq = _random_case(rng, 1, H)
        p = float(np.float32(10.0 ** (-bq[0] / 10.0)))
        lam = [(1 - p) if read[0] == h else p / 3 for h in hap]
        want = math.log10(F32_BETA * sum(lam) / H)
        got = oracle.align(hap, read, bq, iq, dq, 10, oracle.F64)
        assert abs(got - want) < 1e-12


def test_closed_form_single_hap_base(oracle):
    """H = 1: the only paths match the first read base and then insert, so the
    score is log10(eps^(R-2) * delta_2 * beta * lambda(1,1)) for R >= 2."""
    rng = np.random.default_rng(4)
    for R in range(2, 12):
        hap, read, bq, iq, dq = _random_case(rng, R, 1)
        P = lambda q: float(np.float32(10.0 ** (-q / 10.0)))
        p = P(bq[0])
        lam = (1 - p) if read[0] == hap[0] else p / 3
        want = math.log10(F32_EPS ** (R - 2) * P(iq[1]) * F32_BETA * lam)
        got = oracle.align(hap, read, bq, iq, dq, 10, oracle.F64)
        assert abs(got - want) < 1e-10, (R, got, want)


def test_n_wildcard_and_semantics(oracle):
    rng = np.random.default_rng(9)
    hap, read, bq, iq, dq = _random_case(rng, 20, 30)
    # an all-N hap matches every read base: equal to the hap that equals the read
    nhap = "N" * 30
    same = read + "".join(rng.choice(list("ACGT"), 10))
    a = oracle.align(nhap, read, bq, iq, dq, 10, oracle.F64)
    b = oracle.align(same, read, bq, iq, dq, 10, oracle.F64)
    assert abs(a - b) < 1e-12 or a > b  # N matches everywhere, so >= any real hap
    # GPU and CPU semantics agree for ACGTN input with gcp '+' (10) ...
    g = oracle.align(hap, read, bq, iq, dq, 10, oracle.F64, oracle.SEM_GPU)
    c = oracle.align(hap, read, bq, iq, dq, 10, oracle.F64, oracle.SEM_CPU)
    assert abs(g - c) < 1e-6
    # ... and differ when gcp is not 10 (the GPU path ignores gcp)
    c2 = oracle.align(hap, read, bq, iq, dq, 20, oracle.F64, oracle.SEM_CPU)
    assert abs(g - c2) > 1e-3
    # non-ACGT bytes: wildcard on the GPU, literal mismatch on the CPU
    g3 = oracle.align("AC*T" * 5, read, bq, iq, dq, 10, oracle.F64, oracle.SEM_GPU)
    c3 = oracle.align("AC*T" * 5, read, bq, iq, dq, 10, oracle.F64, oracle.SEM_CPU)
    assert g3 > c3


def test_empty_inputs(oracle):
    assert oracle.align("", "ACGT", [30] * 4, [40] * 4, [40] * 4) == float("-inf")
    assert oracle.align("ACGT", "", [], [], []) == float("-inf")


def _decimal_forward(read, hap, bq, iq, dq, eps, beta):
    """The forward algorithm in Decimal (unbounded exponent), reference
    constants and GPU-kernel wildcard rule for ACGT input."""
    getcontext().prec = 40
    P = [Decimal(float(np.float32(10.0 ** (-q / 10.0)))) for q in range(128)]
    eps, beta = Decimal(eps), Decimal(beta)
    C = Decimal(float(np.float32(np.finfo(np.float32).max) / np.float32(16)))
    R, H = len(read), len(hap)
    Mp = [Decimal(0)] * (H + 1)
    Ip = [Decimal(0)] * (H + 1)
    Dp = [C / H] * (H + 1)
    for r in range(1, R + 1):
        p, de, si = P[int(bq[r - 1])], P[int(iq[r - 1])], P[int(dq[r - 1])]
        al = 1 - (de + si)
        Mc, Ic, Dc = [Decimal(0)] * (H + 1), [Decimal(0)] * (H + 1), [Decimal(0)] * (H + 1)
        for c in range(1, H + 1):
            lam = (1 - p) if read[r - 1] == hap[c - 1] else p / 3
            Mc[c] = lam * (al * Mp[c - 1] + beta * (Ip[c - 1] + Dp[c - 1]))
            Ic[c] = de * Mp[c] + eps * Ip[c]
            Dc[c] = si * Mc[c - 1] + eps * Dc[c - 1]
        Mp, Ip, Dp = Mc, Ic, Dc
    s = sum(Mp[c] + Ip[c] for c in range(1, H + 1))
    return float(s.log10() - C.log10())


def test_extreme_scores_match_arbitrary_exponent_evaluation(oracle):
    """Scores far below double's range: the oracle (a per-cell exponent
    epoch) agrees with a Decimal evaluation (unbounded exponent)."""
    rng = np.random.default_rng(21)
    # a read far longer than the hap: every path inserts > 1000 bases at
    # eps = 0.1 each, so log10 < -1000 (beyond double's range of ~ -308)
    R, H = 1100, 30
    read = "".join(rng.choice(list("ACGT"), R))
    hap = "".join(rng.choice(list("ACGT"), H))
    q = np.full(R, 93)
    got = oracle.align(hap, read, q, q, q, 10, oracle.F64)
    assert got < -1000
    want = _decimal_forward(read, hap, q, q, q, F32_EPS, F32_BETA)
    assert abs(got - want) < 1e-9 * abs(want)


def test_wide_row_range_matches_arbitrary_exponent_evaluation(oracle):
    """A row whose values span more than double's whole range: a decaying
    deletion carry-over at the left next to the path that dominates later
    columns.  One scale per row (or per wave lane) would flush the path's
    cells; scaling per row lost 11 decades here (-1618.73 for -1607.65)."""
    rng = np.random.default_rng(0)
    R, H = 600, 150
    read = "".join(rng.choice(list("ACGT"), R))
    hap = "".join(rng.choice(list("ACGT"), H))
    bq = rng.integers(10, 41, R)
    iq = rng.integers(25, 46, R)
    dq = rng.integers(25, 46, R)
    gcp = 38  # CPU semantics: eps = ph2pr[38], beta = 1 - eps
    got = oracle.align(hap, read, bq, iq, dq, gcp, oracle.F64, oracle.SEM_CPU)
    eps = float(np.float32(10.0 ** (-gcp / 10.0)))
    want = _decimal_forward(read, hap, bq, iq, dq, eps, 1.0 - eps)
    assert abs(got - (-1607.6463)) < 1e-3
    assert abs(got - want) < 1e-9 * abs(want)


def test_batch_driver_maps_outputs_read_major(oracle, lib):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=3, num_records=2)
    full = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    nr, nh = b.array("num_reads_per_group"), b.array("num_haps_per_group")
    rp, hp = b.array("reads_prefix"), b.array("haps_prefix")
    rl, hl = b.array("read_lengths"), b.array("hap_lengths")
    ro, ho = b.array("read_offsets"), b.array("hap_offsets")
    rd, hd = b.array("read_data"), b.array("hap_data")
    bq, iq, dq = b.array("base_quals"), b.array("ins_quals"), b.array("del_quals")
    rng = np.random.default_rng(0)
    off = 0
    for g in range(len(nr)):
        for _ in range(5):
            r, h = int(rng.integers(nr[g])), int(rng.integers(nh[g]))
            ri, hi = rp[g] + r, hp[g] + h
            want = oracle.align(hd[ho[hi]:ho[hi] + hl[hi]], rd[ro[ri]:ro[ri] + rl[ri]],
                                bq[ro[ri]:ro[ri] + rl[ri]], iq[ro[ri]:ro[ri] + rl[ri]],
                                dq[ro[ri]:ro[ri] + rl[ri]], 10, oracle.F64)
            assert full[off + r * nh[g] + h] == want
        off += int(nr[g]) * int(nh[g])
