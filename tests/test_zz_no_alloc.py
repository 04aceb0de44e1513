"""The allocation-free entry point (phmm_score_device_noalloc) against
phmm_score_device, through the C ABI: bit-identical scores on two calls in a
row, the second on a smaller batch (the array keeps the first call's
contents), on both engine modes; batches that need more scratch than the
array holds are refused with PHMM_ERANGE before any launch."""
import ctypes

import numpy as np
import pytest

import gpupairhmm_amd as G
from conftest import say

pytestmark = pytest.mark.gpu


def _scores(eng, db, fn):
    d = eng.alloc(4 * max(db.num_alignments, 1))
    try:
        fn(db, d)
        eng.synchronize()
        return eng.download(d, db.num_alignments)
    finally:
        eng.free(d)


@pytest.mark.parametrize("mode", [G.MODE_BATCHED, G.MODE_PLAIN])
def test_noalloc_matches_score_device(lib, mode):
    big = G.HostBatch.synthetic(2, seed=31, num_records=12)
    small = G.HostBatch.synthetic(1, seed=32, num_records=40)
    assert small.num_alignments < big.num_alignments
    with G.Engine(mode=mode) as eng:
        for name, b in (("big", big), ("small", small)):
            say(f"case: noalloc mode={mode} {name} batch, {b.num_alignments} alignments")
            db = eng.upload(b)
            want = _scores(eng, db, eng.score_device)
            got = _scores(eng, db, eng.score_device_noalloc)
            db.free()
            assert np.array_equal(want.view(np.uint32), got.view(np.uint32)), name
            assert np.all(np.isfinite(got))


def test_noalloc_two_calls_in_a_row_second_smaller(lib):
    first = G.HostBatch.synthetic(3, seed=33, num_records=3)
    second = G.HostBatch.synthetic(0, seed=34, num_records=2)
    assert second.num_alignments < first.num_alignments
    with G.Engine() as eng:
        dbs = [eng.upload(first), eng.upload(second)]
        ds = [eng.alloc(4 * b.num_alignments) for b in (first, second)]
        say("case: noalloc two calls in a row, second smaller")
        for db, d in zip(dbs, ds):
            eng.score_device_noalloc(db, d)
        eng.synchronize()
        got = [eng.download(d, db.num_alignments) for db, d in zip(dbs, ds)]
        want = [_scores(eng, db, eng.score_device) for db in dbs]
        for g, w in zip(got, want):
            assert np.array_equal(g.view(np.uint32), w.view(np.uint32))
        for db, d in zip(dbs, ds):
            eng.free(d)
            db.free()


def test_noalloc_refuses_batches_beyond_the_array(lib):
    big = G.HostBatch.synthetic(1, seed=35, num_records=10000)  # 1e6 alignments
    with G.Engine() as eng:
        db = eng.upload(big)
        d = eng.alloc(4 * big.num_alignments)
        say("case: noalloc refusal of a 1e6-alignment batch")
        rc = G.load_library().phmm_score_device_noalloc(eng._h, ctypes.byref(db.view),
                                                       ctypes.c_void_p(d), None)
        assert rc == G.ERANGE
        msg = G.load_library().phmm_engine_last_error(eng._h).decode()
        assert "scratch" in msg and str(G.NOALLOC_SCRATCH_BYTES) in msg
        # the engine still works after the refusal
        small = G.HostBatch.synthetic(0, seed=36, num_records=4)
        db2 = eng.upload(small)
        want = _scores(eng, db2, eng.score_device)
        got = _scores(eng, db2, eng.score_device_noalloc)
        assert np.array_equal(want.view(np.uint32), got.view(np.uint32))
        db2.free()
        eng.free(d)
        db.free()
