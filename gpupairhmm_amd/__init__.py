"""gpupairhmm_amd -- Python front end of the MI355X PairHMM engine.

A thin ctypes layer over ``libgpupairhmm.so`` (C ABI declared in
``include/gpupairhmm.h``).  The compute runs in the library's HIP kernels;
this module only marshals batches and results.  It mirrors the reference's
host objects (asbschmidt/gpuPairHMM):

* :class:`HostBatch`   -- ``BatchOfSequenceGroups`` (common.cuh:109-201), built by
  the line-record parser (file_parser_worker.cu:88-218) or the synthetic
  workload generator;
* :class:`Engine`      -- ``GpuPairHMMWorker`` (gpuworker.cuh:381-973):
  :meth:`Engine.score_batch` is ``processBatch_overlapped_float_coalesced_smem``;
  :meth:`Engine.score_device` is the device-pointer dispatch of
  ``MultiConfigGpuPairHMM`` (gpuworker.cuh:182-375);
* :func:`format_scores` -- the writer's text (output_writer_worker.cu:110-127);
* :class:`Server`      -- continuous-batching request loop (no reference
  counterpart).

Errors raise :class:`PairHMMError` carrying the library's status code and
message.  Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

__all__ = [
    "PairHMMError", "HostBatch", "DeviceBatch", "Engine", "Server", "load_library",
    "library_path", "format_scores", "build", "MODE_PLAIN", "MODE_BATCHED", "SEM_GPU",
    "SEM_CPU", "SEM_GCP", "EXPORTED_SYMBOLS",
]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libgpupairhmm.so"

OK, EINVAL, ENOMEM, EDEVICE, EIO, EPARSE, EBUSY, EAGAIN, ERANGE = range(9)
MODE_PLAIN, MODE_BATCHED = 0, 1
SEM_GPU, SEM_CPU, SEM_GCP = 0, 1, 2

# Every symbol include/gpupairhmm.h declares.
EXPORTED_SYMBOLS = (
    "phmm_abi_version", "phmm_last_error", "phmm_options_default", "phmm_engine_create",
    "phmm_engine_destroy", "phmm_engine_last_error", "phmm_engine_stats",
    "phmm_engine_reset_stats", "phmm_engine_set_tiles", "phmm_engine_get_tiles",
    "phmm_score_batch", "phmm_score_device", "phmm_engine_synchronize",
    "phmm_device_batch_upload", "phmm_device_batch_view", "phmm_device_batch_free",
    "phmm_device_alloc", "phmm_device_free", "phmm_memcpy_d2h", "phmm_memcpy_h2d",
    "phmm_host_batch_parse", "phmm_parse_file", "phmm_host_batch_view", "phmm_host_batch_free",
    "phmm_synth_generate", "phmm_host_batch_to_text", "phmm_format_scores", "phmm_free_text",
    "phmm_batch_cells", "phmm_server_options_default", "phmm_server_create",
    "phmm_server_destroy", "phmm_server_submit", "phmm_server_poll", "phmm_server_stats_get",
    "phmm_compute_likelihoods", "phmm_score_device_noalloc", "phmm_host_alloc_pinned",
    "phmm_host_free_pinned",
)
# Device scratch of the allocation-free entry point (PHMM_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 64 << 20


class PairHMMError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


# ---------------------------------------------------------------------------
# C structures (must match include/gpupairhmm.h)
# ---------------------------------------------------------------------------
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)


# Smallest normal float: the default rescue threshold (the FP32 sum in the
# reference's units below which a result is recomputed in FP64).
FLT_MIN = 1.1754943508222875e-38


class Options(ctypes.Structure):
    _fields_ = [
        ("transfer_batchsize", ctypes.c_int64), ("mode", ctypes.c_int32),
        ("semantics", ctypes.c_int32), ("fp64_rescue", ctypes.c_int32),
        ("rescue_min_sum", ctypes.c_float), ("verbose", ctypes.c_int32),
        ("time_kernels", ctypes.c_int32), ("graphs", ctypes.c_int32),
        ("small_batch", ctypes.c_int32), ("f64_rows", ctypes.c_int32),
        ("range_rescue", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2),
    ]


class BatchView(ctypes.Structure):
    _fields_ = [
        ("num_groups", ctypes.c_int32), ("num_reads", ctypes.c_int32),
        ("num_haps", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("read_bytes", ctypes.c_int64), ("hap_bytes", ctypes.c_int64),
        ("num_alignments", ctypes.c_int64),
        ("read_data", _u8p), ("hap_data", _u8p), ("base_quals", _u8p), ("ins_quals", _u8p),
        ("del_quals", _u8p), ("gcp_quals", _u8p),
        ("read_lengths", _i32p), ("hap_lengths", _i32p), ("read_offsets", _i32p),
        ("hap_offsets", _i32p), ("num_reads_per_group", _i32p), ("num_haps_per_group", _i32p),
        ("reads_prefix", _i32p), ("haps_prefix", _i32p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("alignments", ctypes.c_int64), ("cells", ctypes.c_int64), ("rescued", ctypes.c_int64),
        ("fast_alignments", ctypes.c_int64), ("multitile_alignments", ctypes.c_int64),
        ("engine_seconds", ctypes.c_double), ("kernel_seconds", ctypes.c_double),
        ("graph_launches", ctypes.c_int64), ("graph_captures", ctypes.c_int64),
        ("rescue_seconds", ctypes.c_double), ("f64_alignments", ctypes.c_int64),
        ("scaled_alignments", ctypes.c_int64), ("reserved", ctypes.c_int64 * 3),
    ]


class TileConfig(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_int32), ("rows", ctypes.c_int32),
                ("max_read_len", ctypes.c_int32)]


class RegionRead(ctypes.Structure):
    _fields_ = [("bases", _u8p), ("base_quals", _u8p), ("ins_quals", _u8p), ("del_quals", _u8p),
                ("gcp_quals", _u8p), ("length", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RegionHap(ctypes.Structure):
    _fields_ = [("bases", _u8p), ("length", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64), ("max_inflight", ctypes.c_int32),
        ("max_batch_requests", ctypes.c_int32),
        ("max_batch_alignments", ctypes.c_int64), ("max_batch_bytes", ctypes.c_int64),
        ("streams", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("latency_batch_alignments", ctypes.c_int64),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_int64), ("refused", ctypes.c_int64), ("completed", ctypes.c_int64),
        ("launches", ctypes.c_int64), ("inflight", ctypes.c_int64), ("queued", ctypes.c_int64),
        ("max_queue_depth", ctypes.c_int64), ("bytes_in_use", ctypes.c_int64),
        ("cells", ctypes.c_int64), ("scratch_bytes", ctypes.c_int64),
        ("arena_bytes", ctypes.c_int64), ("failed", ctypes.c_int64),
        ("max_concurrent", ctypes.c_int64), ("reserved", ctypes.c_int64 * 4),
    ]


BATCH_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(BatchView), ctypes.c_void_p)

# ---------------------------------------------------------------------------
# library loading
# ---------------------------------------------------------------------------
_LIB = None


def library_path() -> str:
    return os.path.join(PKG_DIR, LIB_NAME)


def build(jobs: int = 16) -> None:
    """Compile libgpupairhmm.so and the gpuPairHMM tool for gfx950 (hipcc)."""
    import subprocess
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", os.path.join(PKG_DIR, "csrc")], check=True)


def load_library():
    """Load the engine library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise PairHMMError(EINVAL, f"{path} is not built; run gpupairhmm_amd.build() "
                                   "(or __graft_entry__.build()) first")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    sig = {
        "phmm_abi_version": (ctypes.c_int, []),
        "phmm_last_error": (ctypes.c_char_p, []),
        "phmm_options_default": (None, [ctypes.POINTER(Options)]),
        "phmm_engine_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Options),
                                              ctypes.POINTER(vp)]),
        "phmm_engine_destroy": (None, [vp]),
        "phmm_engine_last_error": (ctypes.c_char_p, [vp]),
        "phmm_engine_stats": (ctypes.c_int, [vp, ctypes.POINTER(Stats)]),
        "phmm_engine_reset_stats": (ctypes.c_int, [vp]),
        "phmm_engine_set_tiles": (ctypes.c_int, [vp, ctypes.POINTER(TileConfig), ctypes.c_int]),
        "phmm_engine_get_tiles": (ctypes.c_int, [vp, ctypes.POINTER(TileConfig), ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int)]),
        "phmm_score_batch": (ctypes.c_int, [vp, ctypes.POINTER(BatchView),
                                            ctypes.POINTER(ctypes.c_float)]),
        "phmm_score_device": (ctypes.c_int, [vp, ctypes.POINTER(BatchView), vp, vp]),
        "phmm_score_device_noalloc": (ctypes.c_int, [vp, ctypes.POINTER(BatchView), vp, vp]),
        "phmm_compute_likelihoods": (ctypes.c_int, [vp, ctypes.POINTER(RegionRead), ctypes.c_int32,
                                                    ctypes.POINTER(RegionHap), ctypes.c_int32,
                                                    ctypes.POINTER(ctypes.c_double)]),
        "phmm_engine_synchronize": (ctypes.c_int, [vp]),
        "phmm_device_batch_upload": (ctypes.c_int, [vp, ctypes.POINTER(BatchView),
                                                    ctypes.POINTER(vp)]),
        "phmm_device_batch_view": (ctypes.c_int, [vp, ctypes.POINTER(BatchView)]),
        "phmm_device_batch_free": (None, [vp]),
        "phmm_device_alloc": (ctypes.c_int, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
        "phmm_device_free": (ctypes.c_int, [vp, vp]),
        "phmm_host_alloc_pinned": (ctypes.c_int, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
        "phmm_host_free_pinned": (ctypes.c_int, [vp, vp]),
        "phmm_memcpy_d2h": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t]),
        "phmm_memcpy_h2d": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t]),
        "phmm_host_batch_parse": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t,
                                                 ctypes.POINTER(vp)]),
        "phmm_parse_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, BATCH_CALLBACK, vp]),
        "phmm_host_batch_view": (ctypes.c_int, [vp, ctypes.POINTER(BatchView)]),
        "phmm_host_batch_free": (None, [vp]),
        "phmm_synth_generate": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int32,
                                               ctypes.POINTER(vp)]),
        "phmm_host_batch_to_text": (ctypes.c_int, [ctypes.POINTER(BatchView), ctypes.POINTER(vp),
                                                   ctypes.POINTER(ctypes.c_size_t)]),
        "phmm_format_scores": (ctypes.c_int, [ctypes.POINTER(BatchView),
                                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(vp),
                                              ctypes.POINTER(ctypes.c_size_t)]),
        "phmm_free_text": (None, [vp]),
        "phmm_batch_cells": (ctypes.c_int64, [ctypes.POINTER(BatchView)]),
        "phmm_server_options_default": (None, [ctypes.POINTER(ServerOptions)]),
        "phmm_server_create": (ctypes.c_int, [vp, ctypes.POINTER(ServerOptions),
                                              ctypes.POINTER(vp)]),
        "phmm_server_destroy": (None, [vp]),
        "phmm_server_submit": (ctypes.c_int, [vp, ctypes.POINTER(BatchView), ctypes.c_uint64]),
        "phmm_server_poll": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_double)]),
        "phmm_server_stats_get": (ctypes.c_int, [vp, ctypes.POINTER(ServerStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _err(code: int, handle=None) -> PairHMMError:
    lib = load_library()
    if handle is not None:
        msg = lib.phmm_engine_last_error(handle)
    else:
        msg = lib.phmm_last_error()
    return PairHMMError(code, (msg or b"").decode(errors="replace"))


def _check(code: int, handle=None) -> None:
    if code != OK:
        raise _err(code, handle)


# ---------------------------------------------------------------------------
# host batches
# ---------------------------------------------------------------------------
_VIEW_ARRAYS = {
    "read_data": ("read_bytes", np.uint8), "hap_data": ("hap_bytes", np.uint8),
    "base_quals": ("read_bytes", np.uint8), "ins_quals": ("read_bytes", np.uint8),
    "del_quals": ("read_bytes", np.uint8), "gcp_quals": ("read_bytes", np.uint8),
    "read_lengths": ("num_reads", np.int32), "hap_lengths": ("num_haps", np.int32),
    "read_offsets": ("num_reads+1", np.int32), "hap_offsets": ("num_haps+1", np.int32),
    "num_reads_per_group": ("num_groups", np.int32), "num_haps_per_group": ("num_groups", np.int32),
    "reads_prefix": ("num_groups+1", np.int32), "haps_prefix": ("num_groups+1", np.int32),
}

# oracle-style short names for the same arrays
_SHORT = {
    "num_reads": "num_reads_per_group", "num_haps": "num_haps_per_group",
    "read_offsets": "read_offsets", "hap_offsets": "hap_offsets",
}


def _len_of(view: BatchView, spec: str) -> int:
    if spec.endswith("+1"):
        return int(getattr(view, spec[:-2])) + 1
    return int(getattr(view, spec))


class HostBatch:
    """A batch of sequence groups in the reference's padded SoA layout.

    Build one with :meth:`parse`, :meth:`synthetic` or :meth:`from_arrays`.
    Array attributes are numpy views (zero-copy) of the batch's memory.
    """

    def __init__(self, handle=None, arrays=None):
        self._handle = handle
        self._arrays = arrays
        self._view = BatchView()
        if handle is not None:
            _check(load_library().phmm_host_batch_view(handle, ctypes.byref(self._view)))
        else:
            self._fill_view_from_arrays()

    # -- constructors -------------------------------------------------------
    @classmethod
    def parse(cls, text: str | bytes) -> "HostBatch":
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = ctypes.c_void_p()
        _check(load_library().phmm_host_batch_parse(data, len(data), ctypes.byref(h)))
        return cls(handle=h)

    @classmethod
    def synthetic(cls, config: int, seed: int = 0, num_records: int = -1) -> "HostBatch":
        h = ctypes.c_void_p()
        _check(load_library().phmm_synth_generate(int(config), int(seed), int(num_records),
                                                  ctypes.byref(h)))
        return cls(handle=h)

    @classmethod
    def from_arrays(cls, **arrays) -> "HostBatch":
        """Wrap caller-made numpy arrays named like the view fields."""
        arr = {k: np.ascontiguousarray(v, dtype=_VIEW_ARRAYS[k][1]) for k, v in arrays.items()}
        return cls(arrays=arr)

    def _fill_view_from_arrays(self):
        a = self._arrays
        v = self._view
        v.num_groups = len(a["num_reads_per_group"])
        v.num_reads = len(a["read_lengths"])
        v.num_haps = len(a["hap_lengths"])
        v.read_bytes = len(a["read_data"])
        v.hap_bytes = len(a["hap_data"])
        v.num_alignments = int(np.sum(a["num_reads_per_group"].astype(np.int64) *
                                      a["num_haps_per_group"]))
        for k, (_, dt) in _VIEW_ARRAYS.items():
            ct = ctypes.c_uint8 if dt == np.uint8 else ctypes.c_int32
            setattr(v, k, a[k].ctypes.data_as(ctypes.POINTER(ct)))

    def __del__(self):
        if getattr(self, "_handle", None) is not None and _LIB is not None:
            _LIB.phmm_host_batch_free(self._handle)
            self._handle = None

    # -- accessors ----------------------------------------------------------
    @property
    def view(self) -> BatchView:
        return self._view

    def array(self, name: str) -> np.ndarray:
        name = _SHORT.get(name, name)
        if self._arrays is not None:
            return self._arrays[name]
        spec, dt = _VIEW_ARRAYS[name]
        n = _len_of(self._view, spec)
        ptr = getattr(self._view, name)
        if n == 0 or not ptr:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(ptr, shape=(n,)).view(dt)

    def __getattr__(self, name):
        if name in _VIEW_ARRAYS or name in _SHORT:
            return self.array(name)
        raise AttributeError(name)

    @property
    def num_alignments(self) -> int:
        return int(self._view.num_alignments)

    @property
    def num_groups(self) -> int:
        return int(self._view.num_groups)

    def cells(self) -> int:
        """Sum of readLen*hapLen over all alignments (gpuworker.cuh:495-520)."""
        return int(load_library().phmm_batch_cells(ctypes.byref(self._view)))

    def oracle_arrays(self) -> dict:
        """The arrays under the names oracle.score_batch expects."""
        return dict(
            num_reads=self.array("num_reads_per_group"), num_haps=self.array("num_haps_per_group"),
            reads_prefix=self.array("reads_prefix"), haps_prefix=self.array("haps_prefix"),
            read_data=self.array("read_data"), hap_data=self.array("hap_data"),
            base_quals=self.array("base_quals"), ins_quals=self.array("ins_quals"),
            del_quals=self.array("del_quals"), gcp_quals=self.array("gcp_quals"),
            read_lengths=self.array("read_lengths"), hap_lengths=self.array("hap_lengths"),
            read_offsets=self.array("read_offsets"), hap_offsets=self.array("hap_offsets"))

    def to_text(self) -> str:
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        lib = load_library()
        _check(lib.phmm_host_batch_to_text(ctypes.byref(self._view), ctypes.byref(p),
                                           ctypes.byref(n)))
        try:
            return ctypes.string_at(p, n.value).decode()
        finally:
            lib.phmm_free_text(p)

    def subset(self, groups) -> "HostBatch":
        """A new batch holding the listed groups (in the given order)."""
        groups = [int(g) for g in groups]
        nr, nh = self.array("num_reads_per_group"), self.array("num_haps_per_group")
        rp, hp = self.array("reads_prefix"), self.array("haps_prefix")
        ro, ho = self.array("read_offsets"), self.array("hap_offsets")
        rl, hl = self.array("read_lengths"), self.array("hap_lengths")
        parts = {k: [] for k in ("read_data", "base_quals", "ins_quals", "del_quals", "gcp_quals",
                                 "hap_data")}
        r_len, h_len, n_r, n_h = [], [], [], []
        for g in groups:
            r0, r1, h0, h1 = rp[g], rp[g + 1], hp[g], hp[g + 1]
            rb0 = ro[r0]
            rb1 = ro[r1 - 1] + rl[r1 - 1] + (4 - rl[r1 - 1] % 4) if r1 > r0 else rb0
            hb0 = ho[h0]
            hb1 = ho[h1 - 1] + hl[h1 - 1] + (4 - hl[h1 - 1] % 4) if h1 > h0 else hb0
            for k in ("read_data", "base_quals", "ins_quals", "del_quals", "gcp_quals"):
                parts[k].append(self.array(k)[rb0:rb1])
            parts["hap_data"].append(self.array("hap_data")[hb0:hb1])
            r_len.append(rl[r0:r1])
            h_len.append(hl[h0:h1])
            n_r.append(nr[g])
            n_h.append(nh[g])
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
        rlens, hlens = cat(r_len, np.int32), cat(h_len, np.int32)
        pad = lambda x: x + (4 - x % 4)
        arrays = {k: cat(v, np.uint8) for k, v in parts.items()}
        arrays["read_lengths"], arrays["hap_lengths"] = rlens, hlens
        arrays["read_offsets"] = np.concatenate([[0], np.cumsum(pad(rlens))]).astype(np.int32)
        arrays["hap_offsets"] = np.concatenate([[0], np.cumsum(pad(hlens))]).astype(np.int32)
        arrays["num_reads_per_group"] = np.array(n_r, dtype=np.int32)
        arrays["num_haps_per_group"] = np.array(n_h, dtype=np.int32)
        arrays["reads_prefix"] = np.concatenate([[0], np.cumsum(n_r)]).astype(np.int32)
        arrays["haps_prefix"] = np.concatenate([[0], np.cumsum(n_h)]).astype(np.int32)
        return HostBatch.from_arrays(**arrays)


def parse_file(path: str, file_batchsize: int = 0):
    """Yield the batches the reference's parser would form (copies)."""
    out = []
    keep = []

    def cb(view_p, _user):
        v = view_p.contents
        hb = HostBatch(arrays=_copy_view(v))
        out.append(hb)
        return 0

    cfun = BATCH_CALLBACK(cb)
    keep.append(cfun)
    _check(load_library().phmm_parse_file(path.encode(), int(file_batchsize), cfun, None))
    return out


def _copy_view(v: BatchView) -> dict:
    arrays = {}
    for k, (spec, dt) in _VIEW_ARRAYS.items():
        n = _len_of(v, spec)
        ptr = getattr(v, k)
        arrays[k] = (np.ctypeslib.as_array(ptr, shape=(n,)).view(dt).copy()
                     if n and ptr else np.zeros(n, dtype=dt))
    return arrays


def format_scores(batch: HostBatch, scores: np.ndarray) -> str:
    s = np.ascontiguousarray(scores, dtype=np.float32)
    p = ctypes.c_void_p()
    n = ctypes.c_size_t()
    lib = load_library()
    _check(lib.phmm_format_scores(ctypes.byref(batch.view),
                                  s.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  ctypes.byref(p), ctypes.byref(n)))
    try:
        return ctypes.string_at(p, n.value).decode()
    finally:
        lib.phmm_free_text(p)


# ---------------------------------------------------------------------------
# engine
# ---------------------------------------------------------------------------
@dataclass
class EngineStats:
    alignments: int
    cells: int
    rescued: int
    fast_alignments: int
    multitile_alignments: int
    engine_seconds: float
    kernel_seconds: float
    graph_launches: int = 0
    graph_captures: int = 0
    rescue_seconds: float = 0.0
    f64_alignments: int = 0
    scaled_alignments: int = 0


class _PinnedBlocks:
    """Page-locked host blocks of a batch, returned with it."""

    def __init__(self, engine, blocks):
        self.engine, self.blocks = engine, blocks

    def __del__(self):
        if _LIB is not None and getattr(self.engine, "_h", None) is not None:
            for p in self.blocks:
                _LIB.phmm_host_free_pinned(self.engine._h, ctypes.c_void_p(p))


class DeviceBatch:
    """A batch copied to engine-owned device memory."""

    def __init__(self, engine: "Engine", handle, host: HostBatch):
        self.engine = engine
        self._handle = handle
        self.host = host
        self.view = BatchView()
        _check(load_library().phmm_device_batch_view(handle, ctypes.byref(self.view)))

    @property
    def num_alignments(self) -> int:
        return int(self.view.num_alignments)

    def free(self):
        if self._handle is not None:
            load_library().phmm_device_batch_free(self._handle)
            self._handle = None

    def __del__(self):
        if _LIB is not None:
            self.free()


class Engine:
    """One engine on one GPU (``GpuPairHMMWorker``)."""

    def __init__(self, device: int = 0, mode: int = MODE_BATCHED, semantics: int = SEM_GPU,
                 transfer_batchsize: int = 1000000, fp64_rescue: bool = True,
                 rescue_min_sum: float = FLT_MIN, time_kernels: bool = False,
                 graphs: bool = True, small_batch: int | None = None, f64_rows: int = 0,
                 range_rescue: int | bool = True):
        lib = load_library()
        o = Options()
        lib.phmm_options_default(ctypes.byref(o))
        if small_batch is not None:
            o.small_batch = int(small_batch)
        o.transfer_batchsize = int(transfer_batchsize)
        o.mode = int(mode)
        o.semantics = int(semantics)
        o.fp64_rescue = 1 if fp64_rescue else 0
        o.rescue_min_sum = float(rescue_min_sum)
        o.time_kernels = 1 if time_kernels else 0
        o.graphs = 1 if graphs else 0
        o.f64_rows = int(f64_rows)
        # True / 1: per-lane scaling inside the FP32 pass; 2: a second,
        # range-extended FP32 pass; False / 0: FP64 only
        o.range_rescue = int(range_rescue)
        h = ctypes.c_void_p()
        _check(lib.phmm_engine_create(int(device), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self.options = o

    def close(self):
        if getattr(self, "_h", None) is not None:
            load_library().phmm_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        if _LIB is not None:
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            raise _err(rc, self._h)

    def score_batch(self, batch: HostBatch) -> np.ndarray:
        out = np.empty(batch.num_alignments, dtype=np.float32)
        self._chk(load_library().phmm_score_batch(
            self._h, ctypes.byref(batch.view), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def compute_likelihoods(self, reads, haps) -> np.ndarray:
        """One region's log10 likelihood matrix, shape (len(reads), len(haps)).

        reads: sequence of (bases, base_quals, ins_quals, del_quals, gcp_quals),
        bases as bytes/str and qualities as phred values (not ASCII); haps:
        sequence of bytes/str."""
        def u8(x):
            if isinstance(x, str):
                x = x.encode()
            return np.ascontiguousarray(np.frombuffer(x, dtype=np.uint8) if isinstance(x, bytes)
                                        else np.asarray(x, dtype=np.uint8))
        keep = []
        rr = (RegionRead * max(len(reads), 1))()
        for i, (bases, bq, iq, dq, gq) in enumerate(reads):
            arrs = [u8(bases), u8(bq), u8(iq), u8(dq), u8(gq)]
            L = len(arrs[0])
            if any(len(a) != L for a in arrs[1:]):
                raise ValueError(f"read {i}: qualities must match the read length")
            keep.append(arrs)
            ptrs = [a.ctypes.data_as(_u8p) for a in arrs]
            rr[i] = RegionRead(*ptrs, L, 0)
        hh = (RegionHap * max(len(haps), 1))()
        for i, h in enumerate(haps):
            a = u8(h)
            keep.append(a)
            hh[i] = RegionHap(a.ctypes.data_as(_u8p), len(a), 0)
        out = np.empty((len(reads), len(haps)), dtype=np.float64)
        self._chk(load_library().phmm_compute_likelihoods(
            self._h, rr, len(reads), hh, len(haps),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def upload(self, batch: HostBatch) -> DeviceBatch:
        h = ctypes.c_void_p()
        self._chk(load_library().phmm_device_batch_upload(self._h, ctypes.byref(batch.view),
                                                          ctypes.byref(h)))
        return DeviceBatch(self, h, batch)

    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self._chk(load_library().phmm_device_alloc(self._h, int(nbytes), ctypes.byref(p)))
        return p.value

    def free(self, ptr: int) -> None:
        self._chk(load_library().phmm_device_free(self._h, ctypes.c_void_p(ptr)))

    def score_device(self, dbatch: DeviceBatch, d_scores: int, stream: int | None = None) -> None:
        self._chk(load_library().phmm_score_device(self._h, ctypes.byref(dbatch.view),
                                                   ctypes.c_void_p(d_scores),
                                                   ctypes.c_void_p(stream or 0)))

    def score_device_noalloc(self, dbatch: DeviceBatch, d_scores: int,
                             stream: int | None = None) -> None:
        """phmm_score_device without any device allocation: the call's
        scratch comes from the library's per-GPU array."""
        self._chk(load_library().phmm_score_device_noalloc(self._h, ctypes.byref(dbatch.view),
                                                           ctypes.c_void_p(d_scores),
                                                           ctypes.c_void_p(stream or 0)))

    def synchronize(self) -> None:
        self._chk(load_library().phmm_engine_synchronize(self._h))

    def download(self, d_ptr: int, n: int, dtype=np.float32) -> np.ndarray:
        out = np.empty(n, dtype=dtype)
        self._chk(load_library().phmm_memcpy_d2h(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                                 ctypes.c_void_p(d_ptr), out.nbytes))
        return out

    def pinned_copy(self, batch: HostBatch) -> HostBatch:
        """A copy of `batch` whose arrays live in page-locked host memory
        (freed with the returned batch): phmm_score_batch copies them to the
        device without staging."""
        lib = load_library()
        arrays, blocks = {}, []
        for name in _VIEW_ARRAYS:
            a = batch.array(name)
            p = ctypes.c_void_p()
            self._chk(lib.phmm_host_alloc_pinned(self._h, max(a.nbytes, 1), ctypes.byref(p)))
            blocks.append(p.value)
            buf = (ctypes.c_uint8 * max(a.nbytes, 1)).from_address(p.value)
            dst = np.frombuffer(buf, dtype=a.dtype, count=a.size)
            dst[:] = a
            arrays[name] = dst
        out = HostBatch.from_arrays(**arrays)
        out._pinned = _PinnedBlocks(self, blocks)
        return out

    def score_resident(self, dbatch: DeviceBatch) -> np.ndarray:
        """Score a device batch and copy the scores back (convenience)."""
        n = max(dbatch.num_alignments, 1)
        d = self.alloc(4 * n)
        try:
            self.score_device(dbatch, d)
            self.synchronize()
            return self.download(d, dbatch.num_alignments)
        finally:
            self.free(d)

    def stats(self) -> EngineStats:
        s = Stats()
        self._chk(load_library().phmm_engine_stats(self._h, ctypes.byref(s)))
        return EngineStats(s.alignments, s.cells, s.rescued, s.fast_alignments,
                           s.multitile_alignments, s.engine_seconds, s.kernel_seconds,
                           s.graph_launches, s.graph_captures, s.rescue_seconds,
                           s.f64_alignments, s.scaled_alignments)

    def reset_stats(self) -> None:
        self._chk(load_library().phmm_engine_reset_stats(self._h))

    def tiles(self) -> list[tuple[int, int, int]]:
        """The tile table: (lanes, rows, max_read_len) rows."""
        buf = (TileConfig * 64)()
        n = ctypes.c_int()
        self._chk(load_library().phmm_engine_get_tiles(self._h, buf, 64, ctypes.byref(n)))
        return [(buf[i].lanes, buf[i].rows, buf[i].max_read_len) for i in range(n.value)]

    def set_tiles(self, tiles) -> None:
        arr = (TileConfig * len(tiles))(*[TileConfig(int(a), int(b), int(c)) for a, b, c in tiles])
        self._chk(load_library().phmm_engine_set_tiles(self._h, arr, len(tiles)))


class Server:
    """Continuous-batching request loop on an engine."""

    def __init__(self, engine: Engine, hbm_budget_bytes: int = 64 << 30,
                 max_inflight: int = 1 << 20, max_batch_requests: int = 4096,
                 max_batch_alignments: int | None = None, max_batch_bytes: int | None = None,
                 streams: int | None = None, latency_batch_alignments: int | None = None):
        lib = load_library()
        o = ServerOptions()
        lib.phmm_server_options_default(ctypes.byref(o))
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        o.max_inflight = int(max_inflight)
        o.max_batch_requests = int(max_batch_requests)
        if max_batch_alignments is not None:
            o.max_batch_alignments = int(max_batch_alignments)
        if max_batch_bytes is not None:
            o.max_batch_bytes = int(max_batch_bytes)
        if streams is not None:
            o.streams = int(streams)
        if latency_batch_alignments is not None:
            o.latency_batch_alignments = int(latency_batch_alignments)
        h = ctypes.c_void_p()
        _check(lib.phmm_server_create(engine._h, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self.engine = engine
        self._buf = None

    def submit(self, request: HostBatch, tag: int) -> bool:
        """True when admitted, False when the budget refused it."""
        rc = load_library().phmm_server_submit(self._h, ctypes.byref(request.view), int(tag))
        if rc == EBUSY:
            return False
        _check(rc)
        return True

    def poll(self, timeout_ms: int = 0, capacity: int = 1 << 16):
        """(tag, scores, latency_seconds) of a finished request, or None.
        A request whose launch failed raises PairHMMError."""
        lib = load_library()
        if self._buf is None or len(self._buf) < capacity:
            self._buf = np.empty(capacity, dtype=np.float32)
        buf = self._buf
        tag = ctypes.c_uint64()
        n = ctypes.c_int64()
        lat = ctypes.c_double()
        rc = lib.phmm_server_poll(self._h, int(timeout_ms), ctypes.byref(tag),
                                  buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  len(buf), ctypes.byref(n), ctypes.byref(lat))
        if rc == EAGAIN:
            return None
        if rc == ERANGE and n.value > len(buf):
            return self.poll(timeout_ms, int(n.value))
        _check(rc)
        return int(tag.value), buf[: n.value].copy(), float(lat.value)

    def stats(self) -> dict:
        s = ServerStats()
        _check(load_library().phmm_server_stats_get(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_ if k != "reserved"}

    def close(self):
        if getattr(self, "_h", None) is not None:
            load_library().phmm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if _LIB is not None:
            self.close()
