#!/usr/bin/env python3
"""Benchmark: GCUPS of the MI355X PairHMM engine, batched vs plain.

Headline (N=1): the mixed-length workload -- 20,000 seeded records with
R ~ U[1,500] reads and H ~ U[1,64] haplotypes, log-normal read lengths
(median 150, clipped to 36-300) and haplotypes of read length + log-normal
extra (median 100, total <= 1,000): about 1.63e8 alignments and 6.8e12
cells -- scored from device-resident inputs by the batched engine (one
call: binning, persistent FP32 tile launches, and the rescue of the
alignments whose FP32 sum underflows FLT_MIN in the reference's units: an
FP32 recompute with a power-of-two scale per tile lane, FP64 for what it
cannot accept).  A
step is one full pass over the batch.  GCUPS = sum(readLen*hapLen) /
seconds (the reference's definition, gpuworker.cuh:495-520).

The JSON line also carries:
  batched_vs_plain -- a seeded sample of the same records scored as one
                  batched call and as plain per-record calls (the reference
                  with --fileBatchsize 1: every record its own batch,
                  file_parser_worker.cu:55-58), both under the same span:
                  device-resident (inputs uploaded before timing) and from
                  host buffers (staging + H2D + compute + D2H, the
                  reference's engine time, gpuworker.cuh:562, 966-969);
  roofline     -- the FP32 forward stage of a step (the k_fwd_tile_set
                  launches): 11 FLOP/cell (algorithmic) / event-timed
                  duration vs the MI355X FP32 vector peak (157.3 TFLOP/s);
  cpu_baseline -- the reference's own align_host (compiled from its source,
                  OpenMP over alignments like processBatchCPU) on a bounded
                  seeded sample of the same records on this host's cores;
  parity       -- max |delta log10| vs the CPU FP64 oracle on a seeded
                  sample, for every configuration reported;
  extras       -- the CPU-runnable workload (configuration 0), the
                  fixed-shape workload (configuration 1), the
                  precision-stress workload at full size (configuration 3)
                  and the serving loop (configuration 4).
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
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_config2.json")
METRIC = "GCUPS on 1 MI355X, batched vs plain; p50/p99 region latency at ≤1e-3 log10 err"
WORKLOADS = {
    0: "config0: 1,000 records x 8 reads (100 bp) x 4 haps (120 bp)",
    1: "config1: 10,000 records x 25 reads (150 bp) x 4 haps (U[150,350] bp) = 1e6 alignments",
    2: "config2: 20,000 records, R~U[1,500], H~U[1,64], read length lognormal(median 150) "
       "in [36,300], hap = read + lognormal(median 100), <= 1,000",
    3: "config3: 50,000 records x 16 reads (250 bp) x 8 haps (U[2000,5000] bp), "
       "Q 2-20, 5% mismatches",
    4: "config4: Poisson arrivals of single config-2 records",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


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


def max_delta(got, ref) -> float:
    got = np.asarray(got, dtype=np.float64)
    both = np.isneginf(got) & np.isneginf(ref)
    with np.errstate(invalid="ignore"):
        d = np.abs(got - ref)
    d[both] = 0.0
    return float(np.max(d)) if d.size else 0.0


def parity_of(scores, batch, n_sample, seed):
    """max |delta log10| of `scores` (all alignments of `batch`) against the
    FP64 oracle on a seeded sample of n_sample alignments (all when fewer)."""
    from oracle import oracle as O
    n = batch.num_alignments
    rng = np.random.default_rng(seed)
    if n_sample >= n:
        which = np.arange(n)
    else:
        which = np.sort(rng.choice(n, size=n_sample, replace=False))
    ref = O.score_subset(batch.oracle_arrays(), which, O.F64, nthreads=host_cores())
    return {"max_abs_dlog10": max_delta(scores[which], ref), "sample_alignments": int(len(which)),
            "oracle": "FP64", "seed": seed}


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


def seeded_records(batch, cells_target, seed):
    """Records of `batch` in a seeded random order, taken until their cells
    reach cells_target (at least one), returned sorted."""
    rng = np.random.default_rng(seed)
    rp = batch.array("reads_prefix")
    hp = batch.array("haps_prefix")
    rl = batch.array("read_lengths").astype(np.int64)
    hl = batch.array("hap_lengths").astype(np.int64)
    picked, cells = [], 0
    for g in rng.permutation(batch.num_groups):
        picked.append(int(g))
        cells += int(rl[rp[g]:rp[g + 1]].sum()) * int(hl[hp[g]:hp[g + 1]].sum())
        if cells >= cells_target:
            break
    return sorted(picked)


def cpu_baseline(batch, cells_target=5e10, single_cells=1.2e9, seed=7):
    """The reference's align_host (compiled from its source; the FP64 oracle
    when that library is absent) on every host thread over a seeded sample
    of the workload's records of about cells_target cells (~10 s on 16
    threads), and on one thread over a smaller sample."""
    from oracle import oracle as O
    cores = host_cores()
    kind = "reference" if O.ref_lib() is not None else "port"
    recs = seeded_records(batch, cells_target, seed)
    sub = batch.subset(recs)
    dt = _time_cpu(sub, O, cores)
    recs1 = seeded_records(batch, single_cells, seed + 1)
    sub1 = batch.subset(recs1)
    dt1 = _time_cpu(sub1, O, 1)
    return {"value": round(sub.cells() / dt / 1e9, 4), "unit": "GCUPS", "cores": cores,
            "kind": kind,
            "sample": f"{len(recs)} seeded records of {batch.num_groups} ({sub.num_alignments} "
                      f"alignments, {sub.cells():.3e} cells), {dt:.1f} s on {cores} threads",
            "single_thread": {"value": round(sub1.cells() / dt1 / 1e9, 4), "unit": "GCUPS",
                              "sample": f"{len(recs1)} seeded records ({sub1.num_alignments} "
                                        f"alignments, {sub1.cells():.3e} cells), {dt1:.1f} s"},
            "cpu_model": _cpu_model()}


def batched_vs_plain(G, batch, n_records, seed, reps=3):
    """A seeded sample of n_records records of `batch`, scored (a) as one
    batched call and (b) as plain per-record calls, under two spans each:
    device-resident (every input uploaded before the clock starts; plain
    uploads each record as its own device batch) and host buffers
    (phmm_score_batch: staging, H2D, compute, D2H).  The best of `reps`
    passes of each side."""
    rng = np.random.default_rng(seed)
    recs = sorted(int(g) for g in rng.choice(batch.num_groups, size=n_records, replace=False))
    sub = batch.subset(recs)
    per = [sub.subset([g]) for g in range(sub.num_groups)]
    cells = sub.cells()
    out = {"records": n_records, "alignments": sub.num_alignments, "cells": cells,
           "sample": f"{n_records} records drawn by seed {seed} from the headline batch"}

    def best(fn):
        fn()  # warm-up (buffers grow, graphs are captured)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    with G.Engine(mode=G.MODE_BATCHED) as eng:
        db = eng.upload(sub)
        d = eng.alloc(4 * sub.num_alignments)

        def run_b():
            eng.score_device(db, d)
            eng.synchronize()
        t_b_dev = best(run_b)
        scores_b = eng.download(d, sub.num_alignments)
        eng.free(d)
        db.free()
        t_b_host = best(lambda: eng.score_batch(sub))
    with G.Engine(mode=G.MODE_PLAIN) as eng:
        dbs = [eng.upload(p) for p in per]
        ds = [eng.alloc(4 * max(p.num_alignments, 1)) for p in per]

        def run_p():
            for db_, d_ in zip(dbs, ds):
                eng.score_device(db_, d_)
            eng.synchronize()
        t_p_dev = best(run_p)
        scores_p = np.concatenate([eng.download(d_, p.num_alignments) for d_, p in zip(ds, per)])
        for db_, d_ in zip(dbs, ds):
            eng.free(d_)
            db_.free()

        def run_ph():
            for p in per:
                eng.score_batch(p)
        t_p_host = best(run_ph)
    out["device_resident"] = {
        "batched_gcups": round(cells / t_b_dev / 1e9, 1),
        "plain_per_record_gcups": round(cells / t_p_dev / 1e9, 1),
        "speedup": round(t_p_dev / t_b_dev, 2)}
    out["host_buffers"] = {
        "batched_gcups": round(cells / t_b_host / 1e9, 1),
        "plain_per_record_gcups": round(cells / t_p_host / 1e9, 1),
        "speedup": round(t_p_host / t_b_host, 2)}
    # per-record calls run on the one-template table, the batched call on the
    # tuned one: bit-identical except where the range-extended rescue accepts
    # an alignment on one table and hands it to FP64 on the other
    same = scores_b.view(np.uint32) == scores_p.view(np.uint32)
    out["plain_vs_batched"] = {"identical_fraction": round(float(np.mean(same)), 6),
                               "max_abs_dlog10": max_delta(scores_p, scores_b)}
    return out


def bench_config(G, cfg, seed, records, steps, warmup, n_parity, time_kernels=False,
                 mode=None):
    """One configuration, device-resident, batched engine (or `mode`): GCUPS,
    rescued fraction and parity on a seeded sample."""
    batch = G.HostBatch.synthetic(cfg, seed=seed, num_records=records)
    cells = batch.cells()
    mode = G.MODE_BATCHED if mode is None else mode
    with G.Engine(mode=mode, time_kernels=time_kernels) as eng:
        db = eng.upload(batch)
        d = eng.alloc(4 * batch.num_alignments)
        dt = timed_steps(eng, db, d, steps, warmup)
        st = eng.stats()
        scores = eng.download(d, batch.num_alignments)
        eng.free(d)
        db.free()
    res = {"workload": WORKLOADS[cfg], "seed": seed, "records": batch.num_groups,
           "alignments": batch.num_alignments, "cells": cells, "steps": steps,
           "gcups": round(cells * steps / dt / 1e9, 2),
           "ms_per_step": round(dt / steps * 1e3, 3),
           # below the FP32 threshold: finished by the FP32 pass's per-lane
           # scaling, or recomputed after it (of those, in FP64)
           "below_threshold_fraction":
               round((st.scaled_alignments + st.rescued) / max(steps * batch.num_alignments, 1), 5),
           "scaled_fraction": round(st.scaled_alignments / max(steps * batch.num_alignments, 1), 5),
           "rescued_fraction": round(st.rescued / max(steps * batch.num_alignments, 1), 5),
           "f64_fraction": round(st.f64_alignments / max(steps * batch.num_alignments, 1), 5)}
    if time_kernels and st.kernel_seconds > 0:
        res["fp32_stage_ms"] = round(st.kernel_seconds / steps * 1e3, 3)
        res["rescue_stage_ms"] = round(st.rescue_seconds / steps * 1e3, 3)
    if n_parity:
        res["parity"] = parity_of(scores, batch, n_parity, seed)
    return res, batch, st, dt


def bench_engine_time(G, batch, calls=3):
    """The reference's engine time (gpuworker.cuh:562, 966-969): one
    phmm_score_batch call from host buffers -- H2D, compute and D2H, the
    copies of a piece overlapping the kernels of the previous one -- per pass
    over the batch, from pageable arrays (staged through pinned memory) and
    from arrays in pinned memory (copied directly, as the reference's
    batches are pinned)."""
    out = {"calls": calls}
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        for name, b in (("pageable", batch), ("pinned", eng.pinned_copy(batch))):
            eng.score_batch(b)
            t0 = time.perf_counter()
            for _ in range(calls):
                eng.score_batch(b)
            dt = (time.perf_counter() - t0) / calls
            out[name] = {"gcups": round(batch.cells() / dt / 1e9, 2),
                         "ms_per_call": round(dt * 1e3, 3)}
            del b
    return out


def bench_config0(G, seed=1000, steps=20):
    """Configuration 0, the reference's CPU-runnable case: the batched engine
    from device-resident inputs, align_host on every host thread, and parity
    over every alignment."""
    from oracle import oracle as O
    res, batch, _, _ = bench_config(G, 0, seed, -1, steps, 3, n_parity=1 << 30)
    cores = host_cores()
    dc = _time_cpu(batch, O, cores)
    res["cpu_gcups"] = round(batch.cells() / dc / 1e9, 4)
    res["cpu_cores"] = cores
    res["cpu_kind"] = "reference" if O.ref_lib() is not None else "port"
    return res


def bench_serving(G, rates, seconds, pool_records, seed, parity_records=24):
    """Poisson arrivals of configuration-4 requests (one record each), and
    parity of the served scores of a few pool records vs the FP64 oracle.
    Each rate runs on a fresh server: 0.3 s of arrivals (not counted), then
    `seconds` of arrivals whose latencies are reported."""
    pool = [G.HostBatch.synthetic(4, seed=seed + i, num_records=1) for i in range(pool_records)]
    served = {}
    out = []
    with G.Engine(mode=G.MODE_BATCHED) as eng:
        # warm-up, as a server would before taking traffic: single requests,
        # then one burst of the whole pool, so the engine's buffers and graphs
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

            def arrivals(duration, take):
                """Poisson arrivals at `rate` for `duration` seconds, then
                drain; returns (submitted, refused, elapsed)."""
                submitted = refused = got = 0
                t0 = time.perf_counter()
                t_next = t0
                i = 0
                while True:
                    now = time.perf_counter()
                    if now - t0 >= duration:
                        break
                    while t_next <= now:
                        if srv.submit(pool[i % len(pool)], i):
                            submitted += 1
                        else:
                            refused += 1
                        i += 1
                        t_next += rng.exponential(1.0 / rate)
                    res = srv.poll(timeout_ms=0)
                    while res is not None:
                        take(res)
                        got += 1
                        res = srv.poll(timeout_ms=0)
                while got < submitted:
                    res = srv.poll(timeout_ms=10000)
                    if res is None:
                        break
                    take(res)
                    got += 1
                return submitted, refused, time.perf_counter() - t0

            # a short run at the same rate first (steady state: the staging
            # ring and arena pages touched, every launch size seen)
            arrivals(0.3, lambda res: None)
            st0 = srv.stats()
            lat, done, cells_done = [], 0, 0

            def take(res):
                nonlocal done, cells_done
                tag, sc, l = res
                lat.append(l)
                k = tag % len(pool)
                cells_done += pool[k].cells()
                if k < parity_records and k not in served:
                    served[k] = sc
                done += 1
            submitted, refused, elapsed = arrivals(seconds, take)
            st = srv.stats()
            st["launches"] -= st0["launches"]
            srv.close()
            lat_ms = np.array(lat) * 1e3
            out.append({"rate_req_s": rate, "submitted": submitted, "refused": refused,
                        "completed": done, "gcups": round(cells_done / elapsed / 1e9, 1),
                        "p50_ms": round(float(np.percentile(lat_ms, 50)), 2) if done else None,
                        "p99_ms": round(float(np.percentile(lat_ms, 99)), 2) if done else None,
                        "max_queue_depth": st["max_queue_depth"], "launches": st["launches"]})
            log(f"serving {rate} req/s: {out[-1]}")
    worst, n_al = 0.0, 0
    for k, sc in sorted(served.items()):
        p = parity_of(sc, pool[k], 400, seed + k)
        worst = max(worst, p["max_abs_dlog10"])
        n_al += p["sample_alignments"]
    return {"workload": WORKLOADS[4], "pool_records": pool_records, "seconds_per_rate": seconds,
            "rates": out,
            "parity": {"max_abs_dlog10": worst, "sample_alignments": n_al, "oracle": "FP64",
                       "sample": f"up to 400 alignments of each of {len(served)} served "
                                 "pool records"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--records", type=int, default=20000, help="headline (configuration-2) records")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="batched", choices=["batched", "plain"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (alignments "
              "would shard by record, each GPU scoring whole records with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    import gpupairhmm_amd as G
    mode = G.MODE_BATCHED if args.mode == "batched" else G.MODE_PLAIN
    log(f"configuration 2: {args.records} records ...")
    head, batch, st, dt = bench_config(G, 2, 1002, args.records, args.steps, args.warmup,
                                       n_parity=20000, time_kernels=True, mode=mode)
    cells = batch.cells()
    log(f"configuration 2: {head['gcups']} GCUPS, {head['ms_per_step']} ms per step")
    value = cells * args.steps / dt / 1e9
    kernel_avg = st.kernel_seconds / args.steps if st.kernel_seconds > 0 else None
    achieved = cells * FLOP_PER_CELL / kernel_avg / 1e12 if kernel_avg else None
    traffic = traffic_step = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            prof = json.load(f)
        if "dominant_kernel_fetch_bytes_per_launch" in prof:
            traffic = (prof["dominant_kernel_fetch_bytes_per_launch"] +
                       prof["dominant_kernel_write_bytes_per_launch"])
        traffic_step = prof.get("hbm_bytes_per_step")
    roofline = {"bound": "valu", "achieved": round(achieved, 2) if achieved else None,
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4) if achieved else None,
                "traffic": traffic,
                "traffic_kernel": "k_fwd_tile_set<float, 19, true> (one launch per step)",
                "traffic_per_step": traffic_step,
                "kernel": "k_fwd_tile_set<float> (the FP32 forward stage of one step: one "
                          "persistent launch per tile template, event-timed together)",
                "kernel_ms_per_step": round(kernel_avg * 1e3, 3) if kernel_avg else None,
                "flop_per_cell": FLOP_PER_CELL,
                "note": "compute-bound on FP32 VALU (a 3-term recurrence, no MFMA); peak is "
                        "MI355X's FP32 157.3 TFLOP/s; traffic is HBM bytes per launch of the "
                        "dominant kernel (and per step, all kernels) from rocprofv3 FETCH_SIZE + "
                        "WRITE_SIZE, calibrated with a copy kernel"}
    result = {
        "metric": METRIC,
        "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator of the mixed-length workload, seed 1002)",
        "config": {"workload": WORKLOADS[2] + "; FP32 with range-extended FP32 / FP64 "
                               "rescue, device-resident",
                   "records": batch.num_groups, "alignments": batch.num_alignments,
                   "cells": cells, "mode": args.mode},
        "roofline": roofline,
        "parity": head["parity"],
        "below_threshold_fraction": head["below_threshold_fraction"],
        "scaled_fraction": head["scaled_fraction"],
        "rescued_fraction": head["rescued_fraction"],
        "f64_fraction": head["f64_fraction"],
        "rescue_stage_ms": head.get("rescue_stage_ms"),
    }
    if not args.no_cpu:
        log("CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(batch)
    if not args.no_extras:
        log("configuration 2: batched vs plain per-record ...")
        result["batched_vs_plain"] = batched_vs_plain(G, batch, n_records=300, seed=2002)
        log(f"  {result['batched_vs_plain']}")
        del batch
        log("configuration 0 ...")
        result["config0"] = bench_config0(G)
        log("configuration 1 ...")
        c1, b1, _, _ = bench_config(G, 1, 1001, 10000, 10, 3, n_parity=10000)
        c1["plain"] = bench_config(G, 1, 1001, 10000, 10, 3, n_parity=0,
                                   mode=G.MODE_PLAIN)[0]["gcups"]
        c1["engine_time"] = bench_engine_time(G, b1)
        result["config1"] = c1
        del b1
        log("configuration 3 ...")
        result["config3"] = bench_config(G, 3, 1003, 50000, 3, 1, n_parity=1000)[0]
        log("configuration 4: serving ...")
        result["config4"] = bench_serving(G, rates=[1000, 5000, 10000, 20000], seconds=1.0,
                                          pool_records=400, seed=1004)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
