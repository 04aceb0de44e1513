"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, and only as the checker.  The engine never
calls it.

It wraps ``liboracle.so`` (the C restatement of the reference's
``align_host``, gpuworker.cuh:1022-1098, and ``processBatchCPU``,
gpuworker.cuh:975-1020) and, when it has been built, ``_ref/libref_align_host.so``
(the reference's own ``align_host`` compiled from its source; see Makefile).
It also holds a brute-force path enumerator used to pin the recurrence on
tiny inputs.
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_align_host.so")

F64, F32CELLS = 0, 1
SEM_GPU, SEM_CPU, SEM_GCP = 0, 1, 2

FLT_MAX = float(np.finfo(np.float32).max)
CONSTANT = FLT_MAX / 16.0

_lib = None
_ref = None


def build(with_ref: bool | None = None) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference")
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        _lib = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.oracle_align.restype = ctypes.c_double
        _lib.oracle_align.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, u8p, u8p, u8p,
                                      ctypes.c_uint8, ctypes.c_int, ctypes.c_int]
        _lib.oracle_align_gcp.restype = ctypes.c_double
        _lib.oracle_align_gcp.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, u8p, u8p, u8p,
                                          u8p, ctypes.c_int, ctypes.c_int]
        _lib.oracle_ph2pr.argtypes = [ctypes.POINTER(ctypes.c_float)]
        batch_args = [ctypes.c_int, ip, ip, ip, ip, u8p, u8p, u8p, u8p, u8p, u8p, ip, ip, ip, ip]
        _lib.oracle_score_batch.restype = ctypes.c_int
        _lib.oracle_score_batch.argtypes = batch_args + [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        _lib.oracle_score_subset.restype = ctypes.c_int
        _lib.oracle_score_subset.argtypes = batch_args + [
            ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    return _lib


def ref_lib():
    """The reference's own align_host, or None when it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = ctypes.CDLL(REF_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        _ref.ref_align_host.restype = ctypes.c_double
        _ref.ref_align_host.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int, u8p, u8p, u8p,
                                        ctypes.c_uint8]
        _ref.ref_ph2pr.argtypes = [ctypes.POINTER(ctypes.c_float)]
        ip = ctypes.POINTER(ctypes.c_int)
        _ref.ref_score_batch.restype = ctypes.c_int
        _ref.ref_score_batch.argtypes = [ctypes.c_int, ip, ip, ip, ip, u8p, u8p, u8p, u8p, u8p,
                                         u8p, ip, ip, ip, ip, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.c_int64, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
    return _ref


def _u8(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray, str)):
        a = a.encode() if isinstance(a, str) else a
        return np.frombuffer(bytes(a), dtype=np.uint8).copy()
    return np.ascontiguousarray(a, dtype=np.uint8)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def ph2pr() -> np.ndarray:
    out = np.zeros(128, dtype=np.float32)
    lib().oracle_ph2pr(_ptr(out, ctypes.c_float))
    return out


def align(hap, read, base_q, ins_q, del_q, gcp0: int = 10,
          precision: int = F64, semantics: int = SEM_GPU) -> float:
    """One alignment.  Qualities are phred values (33 already subtracted)."""
    h, r = _u8(hap), _u8(read)
    bq, iq, dq = _u8(base_q), _u8(ins_q), _u8(del_q)
    return lib().oracle_align(_ptr(h, ctypes.c_uint8), len(h), _ptr(r, ctypes.c_uint8), len(r),
                              _ptr(bq, ctypes.c_uint8), _ptr(iq, ctypes.c_uint8),
                              _ptr(dq, ctypes.c_uint8), int(gcp0), precision, semantics)


def align_gcp(hap, read, base_q, ins_q, del_q, gcp_q, precision: int = F64,
              semantics: int = SEM_GCP) -> float:
    """One alignment with the read's gcp qualities (one per base), the form
    that serves the per-position gcp semantics."""
    h, r = _u8(hap), _u8(read)
    bq, iq, dq, gq = _u8(base_q), _u8(ins_q), _u8(del_q), _u8(gcp_q)
    if len(gq) < len(r):
        raise ValueError("one gcp quality per read base")
    return lib().oracle_align_gcp(_ptr(h, ctypes.c_uint8), len(h), _ptr(r, ctypes.c_uint8),
                                  len(r), _ptr(bq, ctypes.c_uint8), _ptr(iq, ctypes.c_uint8),
                                  _ptr(dq, ctypes.c_uint8), _ptr(gq, ctypes.c_uint8), precision,
                                  semantics)


def ref_align(hap, read, base_q, ins_q, del_q, gcp0: int = 10) -> float:
    """The reference's align_host (float cells, double sum, CPU semantics)."""
    rl = ref_lib()
    if rl is None:
        raise RuntimeError("oracle/_ref not built")
    h, r = _u8(hap), _u8(read)
    bq, iq, dq = _u8(base_q), _u8(ins_q), _u8(del_q)
    return rl.ref_align_host(_ptr(h, ctypes.c_uint8), len(h), _ptr(r, ctypes.c_uint8), len(r),
                             _ptr(bq, ctypes.c_uint8), _ptr(iq, ctypes.c_uint8),
                             _ptr(dq, ctypes.c_uint8), int(gcp0))


def ref_score_batch(b, which=None, nthreads: int = 0) -> np.ndarray:
    """The reference's align_host over a batch (or the listed alignments),
    OpenMP over alignments like processBatchCPU."""
    rl = ref_lib()
    if rl is None:
        raise RuntimeError("oracle/_ref not built")
    args, keep = _batch_args(b)
    if which is None:
        n = int(np.sum(keep["nR"].astype(np.int64) * keep["nH"]))
        wp = None
    else:
        w = np.ascontiguousarray(which, dtype=np.int64)
        keep["which"] = w
        n = len(w)
        wp = _ptr(w, ctypes.c_int64)
    out = np.empty(n, dtype=np.float64)
    rc = rl.ref_score_batch(*args, wp, n, nthreads, _ptr(out, ctypes.c_double))
    if rc != 0:
        raise RuntimeError("ref_score_batch failed")
    return out


def _batch_args(b):
    """Arguments for the batch entry points from a mapping of numpy arrays
    named like the reference's BatchOfSequenceGroups fields."""
    i32 = lambda k: np.ascontiguousarray(b[k], dtype=np.int32)
    u8 = lambda k: np.ascontiguousarray(b[k], dtype=np.uint8)
    keep = dict(
        nR=i32("num_reads"), nH=i32("num_haps"), rp=i32("reads_prefix"), hp=i32("haps_prefix"),
        rd=u8("read_data"), hd=u8("hap_data"), bq=u8("base_quals"), iq=u8("ins_quals"),
        dq=u8("del_quals"), gq=u8("gcp_quals"), rl=i32("read_lengths"), hl=i32("hap_lengths"),
        ro=i32("read_offsets"), ho=i32("hap_offsets"))
    ip = lambda k: _ptr(keep[k], ctypes.c_int)
    up = lambda k: _ptr(keep[k], ctypes.c_uint8)
    args = [len(keep["nR"]), ip("nR"), ip("nH"), ip("rp"), ip("hp"), up("rd"), up("hd"), up("bq"),
            up("iq"), up("dq"), up("gq"), ip("rl"), ip("hl"), ip("ro"), ip("ho")]
    return args, keep


def score_batch(b, precision: int = F64, semantics: int = SEM_GPU, nthreads: int = 0) -> np.ndarray:
    args, keep = _batch_args(b)
    total = int(np.sum(keep["nR"].astype(np.int64) * keep["nH"]))
    out = np.empty(total, dtype=np.float64)
    rc = lib().oracle_score_batch(*args, precision, semantics, nthreads, _ptr(out, ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"oracle_score_batch failed ({rc})")
    return out


def score_subset(b, which, precision: int = F64, semantics: int = SEM_GPU,
                 nthreads: int = 0) -> np.ndarray:
    args, keep = _batch_args(b)
    w = np.ascontiguousarray(which, dtype=np.int64)
    out = np.empty(len(w), dtype=np.float64)
    rc = lib().oracle_score_subset(*args, _ptr(w, ctypes.c_int64), len(w), precision, semantics,
                                   nthreads, _ptr(out, ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"oracle_score_subset failed ({rc})")
    return out


# ---------------------------------------------------------------------------
# Brute force: explicit enumeration of every state path (tiny inputs only).
# ---------------------------------------------------------------------------

def brute_force(hap: bytes, read: bytes, base_q, ins_q, del_q, eps=0.1, beta=0.9,
                literal_n: bool = False) -> float:
    """Sum, over every path through the pair HMM, of the product of its
    transition and emission probabilities -- no dynamic programming.

    eps and beta are numbers, or sequences of one value per read row (row r
    at index r-1: the per-position gcp model).  literal_n: only the byte 'N'
    is a wildcard (the reference CPU rule), else every byte other than
    A/C/G/T.

    The model (the one the reference's forward recurrence computes,
    gpuworker.cuh:1043-1094): a path starts in the deletion state above hap
    column c0 (c0 = 0..H-1) with weight 1/H and enters the match state at
    (row 1, column c0+1) with probability beta.  Moves, for rows r and
    columns c (1-based):
      M(r,c) -> M(r+1,c+1)  alpha_{r+1} * emit(r+1,c+1)
      M(r,c) -> I(r+1,c)    delta_{r+1}
      M(r,c) -> D(r,c+1)    sigma_r
      I(r,c) -> M(r+1,c+1)  beta_{r+1} * emit ;  I(r,c) -> I(r+1,c)  eps_{r+1}
      D(r,c) -> M(r+1,c+1)  beta_{r+1} * emit ;  D(r,c) -> D(r,c+1)  eps_r
    (the start enters row 1 with beta_1).  A path ends in M or I on the last
    read row.  Returns log10 of the sum (the reference's score with its
    scaling constant cancelled).
    """
    R, H = len(read), len(hap)
    ep = list(eps) if hasattr(eps, "__len__") else [eps] * R
    bt = list(beta) if hasattr(beta, "__len__") else [beta] * R
    if len(ep) < R or len(bt) < R:
        raise ValueError("one eps and beta per read row")
    P = [10.0 ** (-q / 10.0) for q in range(128)]
    p = [P[min(q, 127)] for q in base_q]
    dl = [P[min(q, 127)] for q in ins_q]
    sg = [P[min(q, 127)] for q in del_q]
    al = [1.0 - (dl[i] + sg[i]) for i in range(R)]

    def code(b):
        return {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}.get(b, 4)

    def emit(r, c):  # 1-based
        x, h = read[r - 1], hap[c - 1]
        if literal_n:
            m = x == h or x == ord("N") or h == ord("N")
        else:
            x, h = code(x), code(h)
            m = x == h or x == 4 or h == 4
        return 1.0 - p[r - 1] if m else p[r - 1] / 3.0

    total = 0.0
    # explicit depth-first enumeration of paths: (state, r, c, weight)
    stack = [("M", 1, c0 + 1, bt[0] * emit(1, c0 + 1) / H) for c0 in range(H)]
    while stack:
        s, r, c, w = stack.pop()
        if r == R and s in ("M", "I"):
            total += w
        if s == "M":
            if r < R and c < H:
                stack.append(("M", r + 1, c + 1, w * al[r] * emit(r + 1, c + 1)))
            if r < R:
                stack.append(("I", r + 1, c, w * dl[r]))
            if c < H:
                stack.append(("D", r, c + 1, w * sg[r - 1]))
        elif s == "I":
            if r < R and c < H:
                stack.append(("M", r + 1, c + 1, w * bt[r] * emit(r + 1, c + 1)))
            if r < R:
                stack.append(("I", r + 1, c, w * ep[r]))
        else:  # D
            if r < R and c < H:
                stack.append(("M", r + 1, c + 1, w * bt[r] * emit(r + 1, c + 1)))
            if c < H:
                stack.append(("D", r, c + 1, w * ep[r - 1]))
    return math.log10(total) if total > 0 else float("-inf")
