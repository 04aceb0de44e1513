"""The drop-in boundary on the CPU: the line-record parser, the batch layout,
the score writer, file batching, and the synthetic workloads.

Reference behaviour followed: file_parser_worker.cu:31-218 (fields sliced at
k*(L+1), qualities minus 33, padding 4-L%4 with 'N', offsets as prefix sums
of padded lengths, a batch closes at the first record where the running
alignment count reaches fileBatchsize) and output_writer_worker.cu:110-127
("R H" then R lines of H "%.5f").
"""
import os

import numpy as np
import pytest

import gpupairhmm_amd as G


def _readme_text():
    with open(os.path.join(os.path.dirname(__file__), "golden", "readme_record.txt")) as f:
        return f.read()


def test_parse_readme_record_layout(lib):
    b = G.HostBatch.parse(_readme_text())
    assert b.num_groups == 1 and b.num_alignments == 4
    assert list(b.array("read_lengths")) == [41, 33]
    assert list(b.array("hap_lengths")) == [41, 41]
    # padding 4 - L%4 (1..4): 41 -> 44, 33 -> 36
    assert list(b.array("read_offsets")) == [0, 44, 80]
    assert list(b.array("hap_offsets")) == [0, 44, 88]
    rd = b.array("read_data")
    assert bytes(rd[:41]).decode() == "CTGTGTCCATGTCAGAGCAATGGCCCAAGTCTGGGCCTGGG"
    assert bytes(rd[41:44]) == b"NNN"
    bq = b.array("base_quals")
    assert bq[0] == ord("8") - 33 and bq[44] == ord("<") - 33
    assert bq[41] == ord("N") - 33  # quality padding is 'N' - 33 = 45
    assert b.array("ins_quals")[40] == ord("N") - 33 and b.array("ins_quals")[39] == 40
    assert b.array("gcp_quals")[0] == 10
    assert list(b.array("reads_prefix")) == [0, 2] and list(b.array("haps_prefix")) == [0, 2]


def test_padding_lengths(lib):
    for L in range(1, 13):
        text = f"1 1\n{'A' * L} {'I' * L} {'I' * L} {'I' * L} {'+' * L}\n{'C' * L}\n"
        b = G.HostBatch.parse(text)
        pad = 4 - L % 4
        assert b.array("read_offsets")[1] == L + pad
        assert b.array("hap_offsets")[1] == L + pad
        assert 1 <= pad <= 4


def test_text_round_trip(lib):
    b = G.HostBatch.synthetic(2, seed=9, num_records=3)
    b2 = G.HostBatch.parse(b.to_text())
    for k in G._VIEW_ARRAYS:
        assert np.array_equal(b.array(k), b2.array(k)), k


def test_writer_format(lib):
    b = G.HostBatch.parse(_readme_text())
    scores = np.array([-5.971538, -3.1966, float("-inf"), -1.663333], dtype=np.float32)
    text = G.format_scores(b, scores)
    assert text == "2 2\n-5.97154 -3.19660\n-inf -1.66333\n"


def test_malformed_input_is_an_error(lib):
    bad = [
        "x y\nACGT IIII IIII IIII ++++\nACGT\n",           # header not numeric
        "1 1\nACGT\n",                                      # read without fields
        "1 1\nACGT IIII IIII\nACGT\n",                      # too few fields
        "2 1\nACGT IIII IIII IIII ++++\n",                  # truncated record
        "-1 1\n",                                           # negative count
    ]
    for text in bad:
        with pytest.raises(G.PairHMMError) as e:
            G.HostBatch.parse(text)
        assert e.value.code == G.EPARSE


def test_blank_lines_between_records_are_skipped(lib):
    t = _readme_text()
    b = G.HostBatch.parse(t + "\n" + t)
    assert b.num_groups == 2 and b.num_alignments == 8


def test_file_batching_semantics(lib, tmp_path):
    # records of 8x4 = 32 alignments; a batch closes when the running count
    # reaches the threshold (at record granularity)
    b = G.HostBatch.synthetic(0, seed=1, num_records=10)
    path = tmp_path / "in.txt"
    path.write_text(b.to_text())
    sizes = lambda fb: [x.num_groups for x in G.parse_file(str(path), fb)]
    assert sizes(0) == [10]
    assert sizes(-5) == [10]
    assert sizes(1) == [1] * 10
    assert sizes(32) == [1] * 10
    assert sizes(33) == [2] * 5
    assert sizes(100) == [4, 4, 2]
    with pytest.raises(G.PairHMMError) as e:
        G.parse_file(str(tmp_path / "missing.txt"), 0)
    assert e.value.code == G.EIO


def test_synthetic_workloads_are_seeded(lib):
    a = G.HostBatch.synthetic(1, seed=42, num_records=5)
    b = G.HostBatch.synthetic(1, seed=42, num_records=5)
    c = G.HostBatch.synthetic(1, seed=43, num_records=5)
    assert np.array_equal(a.array("read_data"), b.array("read_data"))
    assert not np.array_equal(a.array("read_data"), c.array("read_data"))


@pytest.mark.parametrize("config,R,H,rl,hl", [
    (0, (8, 8), (4, 4), (100, 100), (120, 120)),
    (1, (25, 25), (4, 4), (150, 150), (150, 350)),
    (2, (1, 500), (1, 64), (36, 300), (37, 1000)),
    (3, (16, 16), (8, 8), (250, 250), (2000, 5000)),
])
def test_synthetic_shapes(lib, config, R, H, rl, hl):
    b = G.HostBatch.synthetic(config, seed=7, num_records=12)
    nr, nh = b.array("num_reads_per_group"), b.array("num_haps_per_group")
    assert nr.min() >= R[0] and nr.max() <= R[1]
    assert nh.min() >= H[0] and nh.max() <= H[1]
    L, HL = b.array("read_lengths"), b.array("hap_lengths")
    assert L.min() >= rl[0] and L.max() <= rl[1]
    assert HL.min() >= hl[0] and HL.max() <= hl[1]
    q = b.array("base_quals")
    lo, hi = (2, 20) if config == 3 else (10, 40)
    real = np.concatenate([q[o:o + l] for o, l in zip(b.array("read_offsets"), L)])
    assert real.min() >= lo and real.max() <= hi
    assert set(np.unique(b.array("gcp_quals"))) <= {10, 45}


def test_cells_match_reference_definition(lib):
    b = G.HostBatch.synthetic(2, seed=1, num_records=3)
    nr, nh = b.array("num_reads_per_group"), b.array("num_haps_per_group")
    rp, hp = b.array("reads_prefix"), b.array("haps_prefix")
    L, HL = b.array("read_lengths").astype(np.int64), b.array("hap_lengths").astype(np.int64)
    want = sum(int(L[rp[g]:rp[g + 1]].sum()) * int(HL[hp[g]:hp[g + 1]].sum())
               for g in range(len(nr)))
    assert b.cells() == want


def test_subset_preserves_records(lib, oracle):
    b = G.HostBatch.synthetic(1, seed=2, num_records=4)
    s = b.subset([1, 3])
    assert s.num_groups == 2
    full = oracle.score_batch(b.oracle_arrays(), oracle.F64)
    sub = oracle.score_batch(s.oracle_arrays(), oracle.F64)
    assert np.array_equal(sub, np.concatenate([full[100:200], full[300:400]]))
