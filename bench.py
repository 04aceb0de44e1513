#!/usr/bin/env python3
"""Benchmark: GCUPS of the MI355X PairHMM engine.

Default (N=1): the configuration-1 workload -- 10,000 records x 25 reads
(150 bp) x 4 haplotypes (U[150,350] bp) = 1e6 alignments, 3.75e10 cells --
scored from device-resident inputs with the batched engine (FP32 bulk, FP64
rescue on the GPU).  A step is one full pass over the batch.  GCUPS =
sum(readLen*hapLen) / seconds (the reference's definition,
gpuworker.cuh:495-520).

The JSON line also carries:
  roofline     -- forward kernels: 11 FLOP/cell (algorithmic) / event-timed
                  duration vs the MI355X FP32 vector peak (157.3 TFLOP/s);
  cpu_baseline -- the reference's own align_host (compiled from its source,
                  OpenMP over alignments like processBatchCPU) on a bounded
                  sample of the same workload on this host's cores;
  parity       -- max |delta log10| vs the CPU FP64 oracle on a 1% sample;
  extras       -- batched vs plain per-record launches on the mixed-length
                  workload (configuration 2), the precision-stress workload
                  (configuration 3), and the serving loop (configuration 4).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector (packed), AMD public spec
FLOP_PER_CELL = 11            # M: add mul fma mul (5), I: mul fma (3), D: mul fma (3)
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_config1.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(eng, dbatch, d_scores, steps, warmup):
    for _ in range(warmup):
        eng.score_device(dbatch, d_scores)
    eng.synchronize()
    eng.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.score_device(dbatch, d_scores)
    eng.synchronize()
    return time.perf_counter() - t0


def parity_sample(eng, dbatch, d_scores, batch, frac=0.01, seed=0):
    from oracle import oracle as O
    scores = eng.download(d_scores, batch.num_alignments)
    rng = np.random.default_rng(seed)
    n = batch.num_alignments
    which = np.sort(rng.choice(n, size=max(1, int(n * frac)), replace=False))
    ref = O.score_subset(batch.oracle_arrays(), which, O.F64)
    got = scores[which].astype(np.float64)
    both = np.isneginf(got) & np.isneginf(ref)
    with np.errstate(invalid="ignore"):
        d = np.abs(got - ref)
    d[both] = 0.0
    return float(np.max(d)), int(len(which))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _time_cpu(sub, O, threads):
    t0 = time.perf_counter()
    if O.ref_lib() is not None:
        O.ref_score_batch(sub.oracle_arrays(), nthreads=threads)
    else:
        O.score_batch(sub.oracle_arrays(), O.F64, nthreads=threads)
    return time.perf_counter() - t0


def cpu_baseline(batch, records_frac=1.0, single_frac=0.01):
    """The reference's align_host (compiled from its source; the FP64 oracle
    when that library is absent) on every host thread over a bounded sample
    of the workload -- by default the whole 1e6-alignment batch, about 8 s
    on 16 threads -- and on one thread over the first 1% of its records."""
    from oracle import oracle as O
    n_groups = batch.num_groups
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    kind = "reference" if O.ref_lib() is not None else "port"
    k = max(1, int(n_groups * records_frac))
    sub = batch.subset(range(k))
    dt = _time_cpu(sub, O, cores)
    k1 = max(1, int(n_groups * single_frac))
    sub1 = batch.subset(range(k1))
    dt1 = _time_cpu(sub1, O, 1)
    return {"value": round(sub.cells() / dt / 1e9, 4), "unit": "GCUPS", "cores": cores,
            "kind": kind,
            "sample": f"first {k} of {n_groups} records ({sub.num_alignments} alignments, "
                      f"{sub.cells():.3e} cells), {dt:.1f} s",
            "single_thread": {"value": round(sub1.cells() / dt1 / 1e9, 4), "unit": "GCUPS",
                              "sample": f"first {k1} records ({sub1.num_alignments} alignments, "
                                        f"{sub1.cells():.3e} cells), {dt1:.1f} s"},
            "cpu_model": _cpu_model()}


def bench_mixed(G, records, seed):
    """Configuration 2: batched (one device-resident call) vs plain per-record
    launches (the reference with --fileBatchsize 1) on the same records."""
    batch = G.HostBatch.synthetic(2, seed=seed, num_records=records)
    cells = batch.cells()
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        db = eng.upload(batch)
        d = eng.alloc(4 * batch.num_alignments)
        timed_steps(eng, db, d, 1, 1)
        dt_b = timed_steps(eng, db, d, 2, 0) / 2
        eng.free(d)
        db.free()
    per_record = [batch.subset([g]) for g in range(batch.num_groups)]
    with G.Engine(mode=G.MODE_PLAIN) as eng:
        eng.score_batch(per_record[0])
        t0 = time.perf_counter()
        for rb in per_record:
            eng.score_batch(rb)
        dt_p = time.perf_counter() - t0
    return {"records": records, "alignments": batch.num_alignments, "cells": cells,
            "batched_gcups": round(cells / dt_b / 1e9, 2),
            "plain_per_record_gcups": round(cells / dt_p / 1e9, 2),
            "speedup": round(dt_p / dt_b, 2)}


def bench_full_mixed(G, records, seed):
    batch = G.HostBatch.synthetic(2, seed=seed, num_records=records)
    cells = batch.cells()
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        db = eng.upload(batch)
        d = eng.alloc(4 * batch.num_alignments)
        timed_steps(eng, db, d, 1, 1)
        dt = timed_steps(eng, db, d, 3, 0) / 3
        st = eng.stats()
        eng.free(d)
        db.free()
    return {"records": records, "alignments": batch.num_alignments, "cells": cells,
            "batched_gcups": round(cells / dt / 1e9, 2), "seconds": round(dt, 3),
            "rescued": st.rescued // 3}


def bench_precision(G, records, seed):
    """Configuration 3: FP32 bulk with FP64 recompute of the alignments whose
    FP32 sum falls below FLT_MIN (reference units); rescued fraction and
    parity on a 2% sample."""
    batch = G.HostBatch.synthetic(3, seed=seed, num_records=records)
    cells = batch.cells()
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        db = eng.upload(batch)
        d = eng.alloc(4 * batch.num_alignments)
        timed_steps(eng, db, d, 1, 1)
        dt = timed_steps(eng, db, d, 5, 0) / 5
        st = eng.stats()
        delta, n = parity_sample(eng, db, d, batch, frac=0.02, seed=3)
        eng.free(d)
        db.free()
    return {"records": records, "alignments": batch.num_alignments, "cells": cells,
            "gcups": round(cells / dt / 1e9, 2),
            "rescued_fraction": round(st.rescued / max(5 * batch.num_alignments, 1), 5),
            "max_abs_dlog10_sample": delta, "sample_alignments": n}


def bench_config1_mode(G, batch, mode_name, steps=10, warmup=3):
    """The same configuration-1 batch through the other engine mode."""
    mode = G.MODE_PLAIN if mode_name == "plain" else G.MODE_BATCHED
    with G.Engine(mode=mode) as eng:
        db = eng.upload(batch)
        d = eng.alloc(4 * batch.num_alignments)
        dt = timed_steps(eng, db, d, steps, warmup)
        eng.free(d)
        db.free()
    return {"mode": mode_name, "gcups": round(batch.cells() * steps / dt / 1e9, 2),
            "ms_per_step": round(dt / steps * 1e3, 3)}


def bench_engine_time(G, batch, calls=5):
    """The reference's engine time (gpuworker.cuh:562, 966-969): one
    phmm_score_batch call from host buffers -- staging, H2D, compute and D2H
    -- per pass over the batch."""
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        eng.score_batch(batch)
        t0 = time.perf_counter()
        for _ in range(calls):
            eng.score_batch(batch)
        dt = (time.perf_counter() - t0) / calls
    return {"gcups": round(batch.cells() / dt / 1e9, 2), "ms_per_call": round(dt * 1e3, 3),
            "calls": calls, "path": "host buffers: staging + H2D + compute + D2H"}


def bench_config0(G, seed=1000, steps=20):
    """Configuration 0, the reference's CPU-runnable case (1,000 records x 8
    reads of 100 bp x 4 haps of 120 bp): the batched engine from
    device-resident inputs, and align_host on every host thread."""
    from oracle import oracle as O
    batch = G.HostBatch.synthetic(0, seed=seed)
    cells = batch.cells()
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        db = eng.upload(batch)
        d = eng.alloc(4 * batch.num_alignments)
        dt = timed_steps(eng, db, d, steps, 3) / steps
        eng.free(d)
        db.free()
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    dc = _time_cpu(batch, O, cores)
    return {"records": batch.num_groups, "alignments": batch.num_alignments, "cells": cells,
            "gpu_gcups": round(cells / dt / 1e9, 2), "gpu_ms_per_call": round(dt * 1e3, 3),
            "cpu_gcups": round(cells / dc / 1e9, 4), "cpu_cores": cores,
            "cpu_kind": "reference" if O.ref_lib() is not None else "port"}


def bench_serving(G, rates, seconds, pool_records, seed):
    """Poisson arrivals of configuration-4 requests (one record each)."""
    pool = [G.HostBatch.synthetic(4, seed=seed + i, num_records=1) for i in range(pool_records)]
    out = []
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        # warm-up, as a server would before taking traffic: single requests,
        # then one burst of the whole pool, whose merged launches reach the
        # per-launch cap, so the engine's buffers and graphs for both tables
        # exist before the timed runs
        srv = G.Server(eng, hbm_budget_bytes=64 << 30)
        for i in range(8):
            srv.submit(pool[i % len(pool)], i)
            srv.poll(timeout_ms=10000)
        for i, req in enumerate(pool):
            srv.submit(req, i)
        for _ in pool:
            if srv.poll(timeout_ms=10000) is None:
                break
        srv.close()
        for rate in rates:
            srv = G.Server(eng, hbm_budget_bytes=64 << 30)
            rng = np.random.default_rng(seed + rate)
            lat, done, cells_done, submitted, refused = [], 0, 0, 0, 0
            t0 = time.perf_counter()
            t_next = t0
            i = 0
            while True:
                now = time.perf_counter()
                if now - t0 >= seconds:
                    break
                while t_next <= now:
                    req = pool[i % len(pool)]
                    if srv.submit(req, i):
                        submitted += 1
                    else:
                        refused += 1
                    i += 1
                    t_next += rng.exponential(1.0 / rate)
                res = srv.poll(timeout_ms=0)
                while res is not None:
                    tag, sc, l = res
                    lat.append(l)
                    cells_done += pool[tag % len(pool)].cells()
                    done += 1
                    res = srv.poll(timeout_ms=0)
            while done < submitted:
                res = srv.poll(timeout_ms=10000)
                if res is None:
                    break
                tag, sc, l = res
                lat.append(l)
                cells_done += pool[tag % len(pool)].cells()
                done += 1
            elapsed = time.perf_counter() - t0
            st = srv.stats()
            srv.close()
            lat_ms = np.array(lat) * 1e3
            out.append({"rate_req_s": rate, "submitted": submitted, "refused": refused,
                        "completed": done, "gcups": round(cells_done / elapsed / 1e9, 1),
                        "p50_ms": round(float(np.percentile(lat_ms, 50)), 2) if done else None,
                        "p99_ms": round(float(np.percentile(lat_ms, 99)), 2) if done else None,
                        "max_queue_depth": st["max_queue_depth"], "launches": st["launches"]})
            log(f"serving {rate} req/s: {out[-1]}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--records", type=int, default=10000, help="configuration-1 records")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="batched", choices=["batched", "plain"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(alignments shard by record with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import gpupairhmm_amd as G
    mode = G.MODE_BATCHED if args.mode == "batched" else G.MODE_PLAIN
    log("generating configuration 1 ...")
    batch = G.HostBatch.synthetic(1, seed=1001, num_records=args.records)
    cells = batch.cells()
    with G.Engine(mode=mode, time_kernels=True) as eng:
        db = eng.upload(batch)
        d_scores = eng.alloc(4 * batch.num_alignments)
        log(f"timing {args.steps} steps ({batch.num_alignments} alignments, {cells:.3e} cells)")
        dt = timed_steps(eng, db, d_scores, args.steps, args.warmup)
        st = eng.stats()
        delta, n_par = parity_sample(eng, db, d_scores, batch)
        eng.free(d_scores)
        db.free()
    value = cells * args.steps / dt / 1e9
    kernel_avg = st.kernel_seconds / args.steps if st.kernel_seconds > 0 else None
    achieved = cells * FLOP_PER_CELL / kernel_avg / 1e12 if kernel_avg else None
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "achieved": round(achieved, 2) if achieved else None,
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4) if achieved else None,
                "traffic": traffic,
                "kernel": "k_fwd_tile_set (FP32 forward stage of one step, event-timed)",
                "kernel_ms_per_launch": round(kernel_avg * 1e3, 3) if kernel_avg else None,
                "flop_per_cell": FLOP_PER_CELL,
                "note": "compute-bound on FP32 VALU (a 3-term recurrence, no MFMA); peak is MI355X's "
                        "FP32 157.3 TFLOP/s"}
    result = {
        "metric": "GCUPS on 1 MI355X, batched vs plain; p50/p99 region latency at \u22641e-3 log10 err",
        "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded, configuration 1 generator)",
        "config": {"workload": "config1: 10,000 records x 25 reads (150 bp) x 4 haps "
                               "(U[150,350] bp) = 1e6 alignments; FP32 + FP64 rescue",
                   "records": args.records, "alignments": batch.num_alignments, "cells": cells,
                   "mode": args.mode},
        "roofline": roofline,
        "parity": {"max_abs_dlog10": delta, "sample_alignments": n_par, "oracle": "FP64"},
        "rescued_fraction": round(st.rescued / max(batch.num_alignments * args.steps, 1), 5),
    }
    if not args.no_cpu:
        log("CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(batch)
    if not args.no_extras:
        other = "plain" if args.mode == "batched" else "batched"
        log(f"configuration 1: {other} path ...")
        result[f"config1_{other}"] = bench_config1_mode(G, batch, other)
        log("configuration 1: engine time from host buffers ...")
        result["config1_engine_time"] = bench_engine_time(G, batch)
        log("configuration 0 ...")
        result["config0"] = bench_config0(G)
        log("configuration 2: batched vs plain per-record ...")
        result["mixed_batched_vs_plain"] = bench_mixed(G, records=150, seed=1002)
        log("configuration 2: batched, 2000 records ...")
        result["mixed_batched"] = bench_full_mixed(G, records=2000, seed=1002)
        log("configuration 3: precision stress ...")
        result["precision_stress"] = bench_precision(G, records=500, seed=1003)
        log("configuration 4: serving ...")
        result["serving"] = bench_serving(G, rates=[1000, 5000, 10000, 20000], seconds=1.0,
                                          pool_records=400, seed=1004)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
