"""The batched engine's captured-graph replay against kernel-by-kernel
launches and the plain path: bit-identical scores across a sequence of
different batches (each replay reads a new batch descriptor), across scratch
growth and tile-table changes (both force a new capture), with the forward
stage timed separately, and with calls alternating between streams."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _batches():
    import gpupairhmm_amd as G
    # small, large (scratch grows), small again, another config
    return [G.HostBatch.synthetic(1, seed=3, num_records=5),
            G.HostBatch.synthetic(2, seed=4, num_records=6),
            G.HostBatch.synthetic(0, seed=5, num_records=40),
            G.HostBatch.synthetic(3, seed=6, num_records=2),
            G.HostBatch.synthetic(1, seed=7, num_records=3)]


def _bits(x):
    return np.asarray(x, dtype=np.float32).view(np.uint32)


# small_batch: None = default (every batch here on the one-template table),
# 0 = the tuned table throughout, 1000 = calls switch between the two
@pytest.mark.parametrize("timed,small", [(False, None), (True, None), (False, 0), (False, 1000)])
def test_graph_replay_matches_eager_and_plain(lib, timed, small):
    import gpupairhmm_amd as G
    bs = _batches()
    with G.Engine(mode=G.MODE_PLAIN, small_batch=small) as plain, \
            G.Engine(mode=G.MODE_BATCHED, graphs=False, small_batch=small) as eager, \
            G.Engine(mode=G.MODE_BATCHED, graphs=True, time_kernels=timed,
                     small_batch=small) as graph:
        for rep in range(2):
            for i, b in enumerate(bs):
                say(f"case: graphs timed={timed} small_batch={small} rep {rep} batch {i} "
                    f"({b.num_alignments} alignments)")
                want = plain.score_batch(b)
                assert np.array_equal(_bits(eager.score_batch(b)), _bits(want))
                assert np.array_equal(_bits(graph.score_batch(b)), _bits(want))
        st = graph.stats()
        assert st.graph_launches >= 2 * len(bs)
        # captures happen only on new scratch sizes, not per call
        assert 1 <= st.graph_captures < st.graph_launches


def test_graph_recaptured_after_tile_change(lib):
    import gpupairhmm_amd as G
    b = G.HostBatch.synthetic(2, seed=8, num_records=4)
    with G.Engine(mode=G.MODE_PLAIN) as plain, G.Engine(mode=G.MODE_BATCHED) as graph:
        base = graph.score_batch(b)
        assert np.array_equal(_bits(base), _bits(plain.score_batch(b)))
        table = [(8, 16, 127), (16, 16, 255), (20, 16, 319)]
        say("case: graph after set_tiles")
        plain.set_tiles(table)
        graph.set_tiles(table)
        got = graph.score_batch(b)
        assert np.array_equal(_bits(got), _bits(plain.score_batch(b)))


def _hip():
    # the HIP runtime the engine library is linked against
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def test_graph_device_calls_on_two_streams(lib):
    import ctypes
    import gpupairhmm_amd as G
    bs = _batches()[:3]
    with G.Engine(mode=G.MODE_PLAIN) as plain, G.Engine(mode=G.MODE_BATCHED) as graph:
        hip = _hip()
        streams = [ctypes.c_void_p(), ctypes.c_void_p()]
        for s in streams:
            assert hip.hipStreamCreateWithFlags(ctypes.byref(s), 1) == 0  # non-blocking
        dbs = [graph.upload(b) for b in bs]
        outs = [graph.alloc(4 * b.num_alignments) for b in bs]
        say("case: graph replays alternating between two streams")
        for k in range(3):
            for i, (db, d) in enumerate(zip(dbs, outs)):
                graph.score_device(db, d, stream=streams[(k + i) % 2].value)
        graph.synchronize()
        for s in streams:
            hip.hipStreamDestroy(s)
        for b, db, d in zip(bs, dbs, outs):
            got = graph.download(d, b.num_alignments)
            assert np.array_equal(_bits(got), _bits(plain.score_batch(b)))
        for db, d in zip(dbs, outs):
            graph.free(d)
            db.free()
