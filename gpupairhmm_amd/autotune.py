"""Autotuner for the gfx950 tile tables.

The reference ships hand-written per-architecture tables of (group size,
registers) layouts (tuningconfigs.hpp:33-239) and picks the smallest tile that
holds a read (gpuworker.cuh:243-273).  This tool measures instead, on the GPU:

1. FP32 tile rates.  For each read-length class bound B of a probe set it
   times every candidate (lanes, rows) whose tile holds B + 1 rows on
   synthetic records of reads of length B, for both arithmetic forms (FAST
   reads: constant indel qualities; SAFE reads: indel qualities that jump
   between rows) and several hap lengths, with the engine's FP32 tile kernels
   (the in-pass scaled instances, 12 to 20 rows per lane, as a default
   engine runs them; plain mode, the event-timed forward and rescue stages).
2. Class boundaries.  From those rates it fits a slot-rate model per rows
   per lane and form (cells per second of a full tile, scaled by the lanes a
   wave leaves idle and the hap's fill steps) and chooses up to 32 class
   bounds by dynamic programming, minimising the modelled time of a
   read-length distribution -- by default the mix of the benchmark
   workloads -- where every alignment pays for its class's whole tile.
3. The table.  Each chosen class is measured again at its bound, against a
   hap 100 bases longer than its reads and a long one, and takes the
   candidate with the least time per cell (FAST and SAFE weighted by the
   FAST share, the two haps alike).
4. FP64 tile rates.  For rows per lane 4, 6, 8, 10 and both forms it times
   the FP64 rescue stage with every alignment sent to it; the engine picks
   each bin's FP64 tile from these rates.

Outputs: csrc/tuning_gfx950.inc (compiled into the engine as its default)
and tuning_gfx950.json (every measurement).

    python -m gpupairhmm_amd.autotune [--quick] [--no-write]

Needs the GPU; rebuild the library afterwards (``make -C gpupairhmm_amd/csrc``).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# rows per lane of the FP32 tiles with an in-pass scaled instance: the default
# engine scales its FP32 tiles in the pass and runs the other row counts
# (4, 8, 24, 28, 32) unscaled, handing each alignment that falls below
# FLT_MIN to the FP64 rescue -- a cost the probe batches (no alignment below
# FLT_MIN) cannot charge, so those rows are not candidates
ROWS = (12, 13, 16, 19, 20)
ROWS64 = (4, 6, 8, 10)
# probe classes of the rate measurements; 103 and 151 hold reads of 101 and
# 150-151 bases (common sequencer lengths) in a tile with at most one spare row
BOUNDS = (15, 31, 47, 63, 79, 95, 103, 111, 127, 143, 151, 159, 175, 191, 207, 223, 255, 287, 319,
          383, 447, 511, 639, 767, 1023)
HAP_LENS = (150, 400, 1000)
MAX_CLASSES = 32     # phmm_engine_set_tiles accepts up to 32 rows
FAST_SHARE = 0.8     # weight of the FAST form when a class's tile is chosen


def candidates(bound: int, rows=ROWS):
    out = []
    for r in rows:
        lanes = -(-(bound + 1) // r)
        for g in (lanes, lanes + 1):
            if 1 <= g <= 64 and g * r >= bound + 1 and (g, r) not in out:
                out.append((g, r))
    return out


def probe_alignments(read_len: int) -> int:
    """Alignments a probe batch needs so that the persistent grid runs many
    rounds of tasks whatever the hap length: a batch that fills the device
    only once or twice times the fill, the drain and one task's latency
    rather than the tile's throughput.  Eight rounds of the device's waves
    (256 CUs x 4 SIMDs x 3 waves), each wave holding as many alignments as
    the fewest-lane 20-row candidate packs, times the 4 haps a read-stationary
    task sweeps."""
    per_wave = 64 // -(-(read_len + 1) // 20)
    return 8 * (256 * 4 * 3) * 4 * per_wave


def make_batch(G, read_len: int, hap_len: int, cells_target: float, seed: int, safe=False):
    """Records of 32 reads x 4 haps, all reads of one length, drawn from the
    record's first hap with 1% errors, at least probe_alignments(read_len)
    alignments.  safe: insertion qualities that jump between rows, so no
    read takes the rescaled (FAST) form."""
    rng = np.random.default_rng(seed)
    R, H = 32, 4
    per_rec = R * H * read_len * hap_len
    n_rec = max(4, int(cells_target // per_rec), -(-probe_alignments(read_len) // (R * H)))
    pad = lambda L: L + (4 - L % 4)
    rl = np.full(n_rec * R, read_len, dtype=np.int32)
    hl = np.full(n_rec * H, hap_len, dtype=np.int32)
    rbytes, hbytes = pad(read_len), pad(hap_len)
    bases = np.frombuffer(b"ACGT", dtype=np.uint8)
    haps = bases[rng.integers(0, 4, (n_rec * H, hbytes))]
    haps[:, hap_len:] = ord("N")
    # reads: windows of the record's first hap at random offsets
    span = max(1, hap_len - read_len + 1)
    offs = rng.integers(0, span, n_rec * R)
    src = np.repeat(np.arange(n_rec) * H, R)
    if hap_len >= read_len:
        idx = offs[:, None] + np.arange(read_len)[None, :]
        body = haps[src[:, None], idx]
    else:
        body = bases[rng.integers(0, 4, (n_rec * R, read_len))]
    err = rng.random(body.shape) < 0.01
    body = np.where(err, bases[rng.integers(0, 4, body.shape)], body)
    reads = np.full((n_rec * R, rbytes), ord("N"), dtype=np.uint8)
    reads[:, :read_len] = body
    bq = rng.integers(10, 41, (n_rec * R, rbytes)).astype(np.uint8)
    if safe:
        iq = rng.integers(10, 46, (n_rec * R, rbytes)).astype(np.uint8)
    else:
        iq = np.full((n_rec * R, rbytes), 40, dtype=np.uint8)
    iq[:, read_len - 1:] = 45
    gq = np.full((n_rec * R, rbytes), 10, dtype=np.uint8)
    return G.HostBatch.from_arrays(
        read_data=reads.ravel(), hap_data=haps.ravel(), base_quals=bq.ravel(),
        ins_quals=iq.ravel(), del_quals=iq.ravel().copy(), gcp_quals=gq.ravel(),
        read_lengths=rl, hap_lengths=hl,
        read_offsets=np.arange(n_rec * R + 1, dtype=np.int32) * rbytes,
        hap_offsets=np.arange(n_rec * H + 1, dtype=np.int32) * hbytes,
        num_reads_per_group=np.full(n_rec, R, dtype=np.int32),
        num_haps_per_group=np.full(n_rec, H, dtype=np.int32),
        reads_prefix=np.arange(n_rec + 1, dtype=np.int32) * R,
        haps_prefix=np.arange(n_rec + 1, dtype=np.int32) * H)


def time_config(G, eng, dbatch, d_scores, cfg, bound, reps, stage="fwd"):
    """Best event-timed duration over `reps` calls with the one-class table
    (lanes, rows, bound): stage "fwd" is the FP32 stage plus the rescue stage
    (a tile without a scaled instance hands what falls below the FP32
    threshold to FP64, which a forward-only time would not charge it);
    "rescue" the rescue stage alone."""
    eng.set_tiles([(cfg[0], cfg[1], bound)])
    eng.score_device(dbatch, d_scores)  # warm-up
    eng.synchronize()
    best = float("inf")
    for _ in range(reps):
        eng.reset_stats()
        eng.score_device(dbatch, d_scores)
        eng.synchronize()
        st = eng.stats()
        best = min(best, st.kernel_seconds + st.rescue_seconds if stage == "fwd"
                   else st.rescue_seconds)
    return best


def measure_classes(G, eng, bounds, hap_lens, cells, reps, forms=("fast", "safe"), log=print):
    """Every candidate of every bound, per form and hap length: GCUPS."""
    out = []
    for bound in bounds:
        for form in forms:
            for hl in hap_lens:
                batch = make_batch(G, bound, hl, cells, seed=bound * 7 + hl, safe=form == "safe")
                db = eng.upload(batch)
                d = eng.alloc(4 * batch.num_alignments)
                ncell = batch.cells()
                for cfg in candidates(bound):
                    t = time_config(G, eng, db, d, cfg, bound, reps)
                    out.append({"bound": bound, "form": form, "hap_len": hl, "lanes": cfg[0],
                                "rows": cfg[1], "seconds": t, "gcups": ncell / t / 1e9})
                eng.free(d)
                db.free()
            best = max((r for r in out if r["bound"] == bound and r["form"] == form),
                       key=lambda r: r["gcups"])
            log(f"len<={bound:5d} {form}: best {best['lanes']:2d}x{best['rows']:2d} "
                f"{best['gcups']:8.1f} GCUPS (hap {best['hap_len']})")
    return out


def final_haps(bound: int):
    """Hap lengths a chosen class is measured against: one 100 bases longer
    than its reads (the mixed-length workload's median extra) and one of
    2000 or more (long active regions)."""
    return (bound + 100, max(2000, 4 * bound))


def lane_eff(g):
    return (64 // g) * g / 64.0


def fit_rates(meas):
    """Slot rate per (rows, form): cells per second of a tile whose rows are
    all in use, with the wave's lanes used and no fill steps -- the median
    over the measurements of gcups * tile_rows / read_len / lane_eff /
    (hap / (hap + lanes))."""
    acc = {}
    for m in meas:
        g, r, b, hl = m["lanes"], m["rows"], m["bound"], m["hap_len"]
        sr = m["gcups"] * (g * r) / b / lane_eff(g) / (hl / (hl + g))
        acc.setdefault((r, m["form"]), []).append(sr)
    return {k: float(np.median(v)) for k, v in acc.items()}


def model_cost(g, r, form, rates, hap_len=300.0):
    """Modelled seconds per cell-row slot of a tile (g, r): the tile's rows
    over the slot rate, with its idle lanes and fill steps."""
    sr = rates.get((r, form))
    if not sr:
        return float("inf")
    return (g * r) / (sr * lane_eff(g) * hap_len / (hap_len + g))


def workload_lengths(max_len=1023):
    """Read-length distribution (by alignments) of the benchmark workloads:
    the mixed-length workload (log-normal, median 150, sigma 0.35, clipped to
    36-300) weighted most, the fixed 100-, 150- and 250-base workloads, and a
    floor over every length so no class is left without a tile."""
    p = np.zeros(max_len + 1)
    rng = np.random.default_rng(0)
    mixed = np.clip(np.round(150 * np.exp(0.35 * rng.standard_normal(400000))), 36, 300)
    p += 0.6 * np.bincount(mixed.astype(int), minlength=max_len + 1)[: max_len + 1] / len(mixed)
    p[150] += 0.15
    p[250] += 0.10
    p[100] += 0.05
    p[16:] += 0.10 / (max_len + 1 - 16)
    p[0] = 0.0
    return p / p.sum()


def class_floor(b: int) -> int:
    """Smallest lower bound a of a class (a, b]: above 31 rows a class holds
    no read shorter than two thirds of its bound, so a read pays for at most
    half as many rows again as it has, whatever the workload's weights."""
    return 0 if b <= 31 else (2 * b) // 3


def choose_bounds(rates, p, k=MAX_CLASSES, fast_share=FAST_SHARE, rows=ROWS):
    """Class bounds b_1 < ... < b_k = len(p)-1 minimising
    sum_L p(L) * cost(tile of L's class), where a class (a, b] takes its
    cheapest modelled tile holding b + 1 rows and a >= class_floor(b).
    Dynamic programming over the bounds; returns [(bound, lanes, rows)]."""
    n = len(p) - 1
    P = np.concatenate([[0.0], np.cumsum(p[1:])])  # P[b] = mass of lengths 1..b
    best_tile = [None] * (n + 1)
    cost = np.full(n + 1, np.inf)
    for b in range(1, n + 1):
        for g, r in candidates(b, rows):
            c = (fast_share * model_cost(g, r, "fast", rates) +
                 (1 - fast_share) * model_cost(g, r, "safe", rates))
            if c < cost[b]:
                cost[b], best_tile[b] = c, (g, r)
    INF = float("inf")
    dp = np.full((k + 1, n + 1), INF)
    arg = np.zeros((k + 1, n + 1), dtype=np.int64)
    dp[0][0] = 0.0
    for j in range(1, k + 1):
        prev = dp[j - 1]
        for b in range(1, n + 1):
            if not np.isfinite(cost[b]):
                continue
            a = np.arange(class_floor(b), b)
            tot = prev[a] + (P[b] - P[a]) * cost[b]
            i = int(np.argmin(tot))
            dp[j][b], arg[j][b] = tot[i], a[i]
    j = int(np.argmin(dp[:, n]))
    out, b = [], n
    while b > 0 and j > 0:
        out.append((b,) + best_tile[b])
        b, j = int(arg[j][b]), j - 1
    return sorted(out), float(dp[:, n].min())


def table_cost(table, rates, p, fast_share=FAST_SHARE):
    """Modelled time of distribution p under a (lanes, rows, bound) table."""
    bounds = [t[2] for t in table]
    total = 0.0
    for L in range(1, len(p)):
        if p[L] == 0:
            continue
        i = int(np.searchsorted(bounds, L))
        if i >= len(table):
            continue
        g, r, _ = table[i]
        total += p[L] * (fast_share * model_cost(g, r, "fast", rates) +
                         (1 - fast_share) * model_cost(g, r, "safe", rates))
    return total


def measure_f64(G, cells, reps, log=print):
    """FP64 tile rates: every alignment of a batch of 151-base reads against
    300-base haps rescued (threshold above every FP32 sum), the rescue stage
    event-timed for each rows-per-lane choice and form.  Returns slot rates
    per (rows, form) (cells per second of full tiles)."""
    out = {}
    for form in ("fast", "safe"):
        batch = make_batch(G, 151, 300, cells / 8, seed=64, safe=form == "safe")
        ncell = batch.cells()
        for nr in ROWS64:
            if form == "safe" and nr > 8:
                continue  # 10-row SAFE FP64 tiles need more than 256 VGPRs
            g = -(-152 // nr)
            # range_rescue=0: every alignment goes to FP64 (in-pass scaling
            # would finish them in the FP32 pass)
            with G.Engine(mode=G.MODE_PLAIN, time_kernels=True, rescue_min_sum=float("inf"),
                          range_rescue=0, f64_rows=nr) as e2:
                db = e2.upload(batch)
                d = e2.alloc(4 * batch.num_alignments)
                t = time_config(G, e2, db, d, (8, 19), 151, reps, stage="rescue")
                e2.free(d)
                db.free()
            sr = ncell / t * (g * nr) / 151 / lane_eff(g) / (300 / (300 + g))
            out[f"{nr},{form}"] = sr
            log(f"FP64 {form} {nr:2d} rows: {ncell / t / 1e9:7.1f} GCUPS, slot rate {sr / 1e9:7.1f}")
    return out


def write_inc(table, f64_rates, path):
    with open(path, "w") as f:
        f.write("// gfx950 tile tables.  Generated by gpupairhmm_amd/autotune.py from\n"
                "// measurements on an MI355X.  Included inside namespace phmm.\n"
                "//\n"
                "// FP32 tiles {lanes, rows, max_read_len}: class bounds chosen for the\n"
                "// benchmark workloads' read lengths, each class's tile measured.\n"
                "const phmm_tile_config kDefaultTiles[] = {\n")
        for i in range(0, len(table), 6):
            chunk = table[i:i + 6]
            f.write("    " + " ".join(f"{{{g}, {r}, {b}}}," for g, r, b in chunk) + "\n")
        f.write("};\nconst int kNumDefaultTiles = (int)(sizeof(kDefaultTiles) / "
                "sizeof(kDefaultTiles[0]));\n")
        f.write("// FP64 tile slot rates (Gcells/s of full tiles) by rows per lane 4, 6, 8,\n"
                "// 10: [SAFE, FAST]; 0: no such tile.\n"
                "const float kF64SlotRate[2][4] = {\n")
        for form in ("safe", "fast"):
            vals = [f64_rates.get(f"{nr},{form}", 0.0) / 1e9 for nr in ROWS64]
            f.write("    {" + ", ".join(f"{v:.1f}f" for v in vals) + "},\n")
        f.write("};\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--quick", action="store_true", help="fewer probe classes and cells")
    ap.add_argument("--no-write", action="store_true")
    ap.add_argument("--cells", type=float, default=2e9)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--classes", type=int, default=MAX_CLASSES)
    ap.add_argument("--out-dir", default=None,
                    help="write tuning_gfx950.{json,inc} here instead of into the package")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(PKG_DIR))
    import gpupairhmm_amd as G

    probe = BOUNDS[::3] if args.quick else BOUNDS
    hap_lens = HAP_LENS[:1] if args.quick else HAP_LENS
    cells = args.cells / 4 if args.quick else args.cells
    t0 = time.time()
    # the FP32 kernels that run by default: the scaled instances (in-pass
    # scaling needs the rescue on)
    with G.Engine(mode=G.MODE_PLAIN, time_kernels=True) as eng:
        meas = measure_classes(G, eng, probe, hap_lens, cells, args.reps)
        rates = fit_rates(meas)
        print("slot rates (Gcells/s):",
              {f"{r},{f}": round(v, 1) for (r, f), v in sorted(rates.items())}, flush=True)
        p = workload_lengths()
        bounds, model = choose_bounds(rates, p, k=args.classes)
        print("model bounds:", bounds, flush=True)
        # measure each chosen class at its bound, against a hap a little
        # longer than its reads and a long one, and keep the candidate with
        # the least time per cell over both forms and both haps
        final = []
        for b, _, _ in bounds:
            final += measure_classes(G, eng, [b], final_haps(b), cells, args.reps)
        table = []
        for b, _, _ in bounds:
            best, bt = None, float("inf")
            for g, r in candidates(b):
                per_cell = {f: [1 / m["gcups"] for m in final if m["bound"] == b and
                                m["lanes"] == g and m["rows"] == r and m["form"] == f]
                            for f in ("fast", "safe")}
                if not per_cell["fast"] or not per_cell["safe"]:
                    continue
                t = FAST_SHARE * sum(per_cell["fast"]) + (1 - FAST_SHARE) * sum(per_cell["safe"])
                if t < bt:
                    best, bt = (g, r), t
            table.append((best[0], best[1], b))
    f64 = measure_f64(G, cells, args.reps)
    out = {"device": "MI355X (gfx950)", "time": time.strftime("%Y-%m-%d %H:%M:%S"),
           "seconds": round(time.time() - t0, 1),
           "table": [{"lanes": g, "rows": r, "max_read_len": b} for g, r, b in table],
           "model": {"slot_rates": {f"{r},{f}": v for (r, f), v in rates.items()},
                     "modelled_cost": model, "bounds": bounds},
           "f64_slot_rates": f64,
           "measurements": meas + final}
    if not args.no_write:
        jdir = args.out_dir or PKG_DIR
        idir = args.out_dir or os.path.join(PKG_DIR, "csrc")
        os.makedirs(jdir, exist_ok=True)
        with open(os.path.join(jdir, "tuning_gfx950.json"), "w") as f:
            json.dump(out, f, indent=1)
        write_inc(table, f64, os.path.join(idir, "tuning_gfx950.inc"))
    print(json.dumps(out["table"]))
    return table


if __name__ == "__main__":
    main()
