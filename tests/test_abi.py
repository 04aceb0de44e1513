"""The C ABI library on the CPU: it loads, exports every function that
include/gpupairhmm.h declares, and rejects bad arguments with status codes
(no compute call needs a GPU here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gpupairhmm_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "gpupairhmm.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phmm_[a-z0-9_]+)\s*\(", text)) - {"phmm_batch_callback"})


def test_library_exports_every_declared_symbol(lib):
    declared = _declared()
    assert len(declared) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", G.library_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (phmm_[a-z0-9_]+)$", out, flags=re.M))
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    assert sorted(declared) == sorted(G.EXPORTED_SYMBOLS)
    for s in declared:
        assert hasattr(lib, s)


def test_abi_version_and_defaults(lib):
    assert lib.phmm_abi_version() == 1
    o = G.Options()
    lib.phmm_options_default(ctypes.byref(o))
    assert o.transfer_batchsize == 1000000  # common.cuh:19
    assert o.mode == G.MODE_BATCHED and o.semantics == G.SEM_GPU
    assert o.fp64_rescue == 1 and o.rescue_min_sum == np.float32(G.FLT_MIN)


def test_struct_sizes_match_header(lib):
    # sizes fixed by the header's field lists (LP64)
    assert ctypes.sizeof(G.Options) == 8 + 4 * 4 + 4 * 2 + 6 * 4
    assert ctypes.sizeof(G.BatchView) == 4 * 4 + 3 * 8 + 14 * 8
    assert ctypes.sizeof(G.TileConfig) == 12


def test_region_likelihoods_reject_bad_arguments(lib):
    rr = (G.RegionRead * 1)()
    hh = (G.RegionHap * 1)()
    out = (ctypes.c_double * 1)()
    assert lib.phmm_compute_likelihoods(None, rr, 1, hh, 1, out) == G.EINVAL
    assert b"phmm_compute_likelihoods" in lib.phmm_last_error()


def test_null_arguments_are_rejected(lib):
    assert lib.phmm_engine_create(0, None, None) == G.EINVAL
    assert lib.phmm_score_batch(None, None, None) == G.EINVAL
    assert lib.phmm_score_device(None, None, None, None) == G.EINVAL
    assert lib.phmm_host_batch_parse(None, 5, None) == G.EINVAL
    assert lib.phmm_synth_generate(9, 0, 1, ctypes.byref(ctypes.c_void_p())) == G.EINVAL
    assert lib.phmm_server_create(None, None, None) == G.EINVAL
    assert b"null" in lib.phmm_last_error() or b"config" in lib.phmm_last_error()


def test_engine_create_without_gpu_fails_cleanly(lib):
    h = ctypes.c_void_p()
    rc = lib.phmm_engine_create(0, None, ctypes.byref(h))
    if rc == G.OK:  # a GPU is present: clean up
        lib.phmm_engine_destroy(h)
    else:
        assert rc == G.EDEVICE
        assert lib.phmm_last_error()


def test_validation_of_host_views(lib):
    b = G.HostBatch.synthetic(0, seed=1, num_records=2)
    v = G.BatchView.from_buffer_copy(b.view)
    v.num_alignments += 1  # inconsistent with the groups
    rc = lib.phmm_host_batch_to_text(ctypes.byref(v), ctypes.byref(ctypes.c_void_p()),
                                     ctypes.byref(ctypes.c_size_t()))
    assert rc == G.EINVAL
    assert b"num_alignments" in lib.phmm_last_error()


def test_package_fails_loudly_without_library(tmp_path, monkeypatch):
    monkeypatch.setattr(G, "PKG_DIR", str(tmp_path))
    monkeypatch.setattr(G, "_LIB", None)
    with pytest.raises(G.PairHMMError):
        G.load_library()
