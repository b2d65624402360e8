#!/usr/bin/env python3
"""Committed txn/s of the batched CC + execute path on one MI355X.

Workload (the plain-batch key-value configuration): a table of 10,000,000
rows x 10 fields of 100 bytes (key = row id), filled from a seed in HBM; one
batch of 2^20 transactions x 16 ops, keys Zipf(0.99) over the rows and
distinct within a transaction, 50% READ of one field / 50% UPDATE of one
field with a seeded 100-byte payload.  The batch is uploaded before the
timed region; a step is one execution of the whole batch (CC + execute) on
the device-resident table.  The headline mode is the ordered mode (every
txn commits, each cell sees its ops in TID order); the lowest-TID-wins mode
is reported beside it.

Beside the headline line's fields it reports, from the same run:
  small_batch          execute-to-completion latency of 16..512-txn ordered
                       batches: multi-kernel pipeline vs the one-launch kernel
  serving              served vs plain batches on the same table (ordered
                       mode, open-loop Poisson arrivals of the batch's txns)
  batching             a stream of 200 uneven batches (1k-1M txns, 1-64 ops,
                       key-value and order-entry) through the batching form
                       vs plain per-batch launches
  serving_large_schema served vs plain on W=1024 order entry + a 100M-row
                       key-value table, arrivals from the order-entry mix
  order_entry_plain    one W=64 order-entry batch per launch, both modes
  cpu_baseline         the threaded CPU restatement of the reference's
                       pipeline (all cores, 1 core) and the serial oracle,
                       on this batch and on the reference's CPU config

Prints one JSON line (see the field list at the end of main()).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "committed txn/s + p50/p99 commit latency, served vs plain batch, 1 MI355X"
HBM_PEAK_GBS = 8000.0
TABLE_SEED = 0x4C54504D47000001
BATCH_SEED = 0x4C54504D47000002
SEG_TILE = 2048  # sorted elements per tile of the ordered-mode apply kernel


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batch_counts(batch, width: int) -> dict:
    """Counts the algorithmic-byte formulas need (host side, numpy)."""
    kind = (batch.meta & 3).astype(np.int64)
    col = ((batch.meta >> 2) & 63).astype(np.int64)
    cell = batch.rows.astype(np.int64) * 64 + col
    n = len(kind)
    order = np.argsort(cell, kind="stable")
    sc, sk = cell[order], kind[order]
    head = np.ones(n, bool)
    head[1:] = sc[1:] != sc[:-1]
    seg = np.cumsum(head) - 1
    is_upd = (sk != 0).astype(np.int64)
    # inclusive running "an update happened in this segment" (segment ids increase)
    cm = np.maximum.accumulate((seg << 1) | is_upd) & 1
    reads = sk == 0
    snapshot_reads = int(np.count_nonzero(reads & (cm == 0)))
    written_cells = int(np.unique(sc[sk != 0]).size)
    # written cells whose segment crosses a 2048-element tile boundary of the
    # sorted list: their final write goes through the cross-tile list
    seg_has_write = np.zeros(int(seg[-1]) + 1 if n else 0, bool)
    seg_has_write[seg[sk != 0]] = True
    b = np.arange(SEG_TILE, n, SEG_TILE)
    crossing = np.unique(seg[b][sc[b] == sc[b - 1]]) if n else np.zeros(0, np.int64)
    cross = int(np.count_nonzero(seg_has_write[crossing]))
    return {"n_ops": n, "n_txns": batch.n_txns, "reads": int(reads.sum()),
            "writes": int((sk != 0).sum()), "snapshot_reads": snapshot_reads,
            "written_cells": written_cells, "cross_tile_writes": cross,
            "seg_tiles": (n + SEG_TILE - 1) // SEG_TILE, "width": width}


def algorithmic_bytes(kernel: str, c: dict, commit=None, batch=None) -> float | None:
    """Bytes a kernel must move per launch (not counting re-reads or padding)."""
    n, t, w = c["n_ops"], c["n_txns"], c["width"]
    per = {
        "mo_keys": 14 * n,                     # rows 4 + meta 2 in, key 4 + val 4 out
        "rs_hist_ops": 6 * n,                  # digit 0 from the ops: rows 4 + meta 2 in
        "rs_scatter_ops": 14 * n,              # rows + meta in, key + val out
        "rs_hist": 4 * n,                      # keys in
        "rs_scatter": 16 * n,                  # key+val in and out
        "mo_seg_scan": 8 * n,                  # sorted key+val in
        "mo_warm_seeds": 8 * n,                # payload seeds streamed through the cache
        # sorted key+val; per READ its 100-B result, plus the snapshot cell or
        # the 8-B payload seed of the in-batch UPDATE it sees; per cell whose
        # segment lies in one tile, payload seed 8 + the cell write; per cell
        # whose segment crosses tiles, its 16-B entry in the cross-tile list
        "mo_apply_reads": 8 * n + c["reads"] * w + c["snapshot_reads"] * w
                          + 8 * (c["reads"] - c["snapshot_reads"])
                          + (c["written_cells"] - c["cross_tile_writes"]) * (8 + w)
                          + 16 * c["cross_tile_writes"],
        # per cross-tile cell: its list entry 16 + payload seed 8 + the cell write
        "mo_final_writes": c["cross_tile_writes"] * (16 + 8 + w),
    }
    if commit is not None and batch is not None:
        kind = batch.meta & 3
        op_t = np.repeat(np.arange(batch.n_txns), np.diff(batch.txn_off.astype(np.int64)))
        cm = commit[op_t] != 0
        cr = int(np.count_nonzero(cm & (kind == 0)))
        cw = int(np.count_nonzero(cm & (kind != 0)))
        wr = int(np.count_nonzero(kind != 0))
        per.update({
            # meta 2 every op, row 4 + lock RMW 16 per write op, txn_off 4 per txn
            "cc_reserve": 2 * n + 20 * wr + 4 * t,
            # meta 2 + row 4 + lock 8 per op, txn_off 4 + commit flag 1 per txn,
            # committed READ: cell in + result out
            "cc_validate_read": 14 * n + 5 * t + cr * 2 * w,
            # meta 2 + row 4 per op, commit flag 1 + txn_off 4 per txn,
            # committed write: payload seed 8 + cell write
            "cc_write": 6 * n + 5 * t + cw * (8 + w),
        })
    return per.get(kernel)


def pmc_traffic(workload_key: str) -> dict:
    """HBM bytes per launch by kernel, from the committed counter summaries
    under profiles/ (lines "- `kernel` traffic per launch: N B [workload]",
    written from rocprofv3 FETCH_SIZE + WRITE_SIZE runs scaled by the copy
    kernel's calibration).  Empty when no summary matches this workload."""
    import glob
    import re
    out = {}
    pat = re.compile(r"^- `(\w+)` traffic per launch: (\d+) B \[(.+)\]")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*pmc*.md"))):
        with open(path) as f:
            for line in f:
                m = pat.match(line.strip())
                if m and m.group(3) == workload_key:
                    out[m.group(1)] = int(m.group(2))
    return out


def run_mode(e, batch, mode: int, steps: int, warmup: int) -> dict:
    for _ in range(warmup):
        e.execute(mode)
        e.wait()
    e.wait()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        e.execute(mode)
        e.wait()
        lat.append(time.perf_counter() - s0)
    t1 = time.perf_counter()
    st = e.stats()
    return {"wall_s": t1 - t0, "lat": lat, "committed": st["committed"], "aborted": st["aborted"],
            "exec_ms_last": st["exec_ms"], "upload_ms": st["upload_ms"]}


def profile_mode(e, mode: int) -> tuple[dict, dict]:
    """One extra step with per-kernel HIP events: total ms and launches per
    stage (the radix-sort stages repeat once per digit)."""
    e.execute(mode, profile=True)
    e.wait()
    st = e.stats()
    return st["kernels"], st["kernel_launches"]


def serve_compare(e, pool, args) -> dict:
    """Open-loop Poisson arrivals of single transactions drawn from the batch,
    served on the same device-resident table, ordered mode:
      served: continuous batching, cut at --serve-max-batch txns or when the
              oldest waited --serve-wait-us, 2 batches in flight, admitted by
              their device bytes against --serve-hbm-mb;
      plain:  batches of exactly --plain-batch txns, one at a time;
      plain at equal p99: plain batches of each size in --plain-sweep, and
              the one with the most committed txn/s whose p99 is at most the
              served p99 (the smallest p99 when none is);
      capacity at p99: per served p99, the highest offered rate each form
              carries within it (capacity_at_p99)."""
    from ltpmg_amd import MODE_O
    budget = int(args.serve_hbm_mb * (1 << 20))
    out = {"arrivals": args.serve_arrivals, "plain_batch": args.plain_batch,
           "served_max_batch": args.serve_max_batch, "served_max_wait_us": args.serve_wait_us,
           "served_hbm_inflight_bytes": budget, "plain_sweep": args.plain_sweep, "rates": []}
    keys = ("committed_per_s", "p50_ms", "p99_ms", "p999_ms", "mean_batch_txns", "batches",
            "host_pack_s", "device_ms", "wall_s")

    def arrivals(rate):  # about --serve-seconds of arrivals, at most --serve-arrivals
        return int(min(args.serve_arrivals, max(1 << 16, rate * args.serve_seconds)))

    def plain(rate, b):
        return e.serve(pool, arrival_rate=rate, n_arrivals=arrivals(rate), max_batch=b, min_batch=b,
                       n_inflight=1, mode=MODE_O, seed=7)

    # one short untimed run first: the first DMA from the freshly pinned pool
    # and the first slot resizes are one-time costs
    e.serve(pool, arrival_rate=0, n_arrivals=1 << 18, max_batch=args.serve_max_batch, n_inflight=2,
            mode=MODE_O, seed=3)
    for rate in args.serve_rates:
        row = {"arrival_rate": rate, "arrivals": arrivals(rate)}
        # an eighth of the stream at the same rate first, untimed: the first
        # run at a new load after a lighter one pays one-time costs in the
        # runtime's copy path (the bandwidth-bound uploads of large batches)
        e.serve(pool, arrival_rate=rate, n_arrivals=max(1 << 14, arrivals(rate) // 8), max_batch=args.serve_max_batch,
                max_wait_us=args.serve_wait_us, n_inflight=2, mode=MODE_O, seed=5, hbm_inflight_bytes=budget)
        r = e.serve(pool, arrival_rate=rate, n_arrivals=arrivals(rate), max_batch=args.serve_max_batch,
                    max_wait_us=args.serve_wait_us, n_inflight=2, mode=MODE_O, seed=7,
                    hbm_inflight_bytes=budget)
        row["served"] = {k: r[k] for k in keys + ("peak_inflight_bytes", "budget_holds")}
        log(f"serve served @ {rate:.3g}/s: {r['committed_per_s']:.3g} txn/s, p50 {r['p50_ms']:.3f} ms, "
            f"p99 {r['p99_ms']:.3f} ms, {r['batches']} batches, pack {r['host_pack_s'] * 1e3:.1f} ms, "
            f"peak in flight {r['peak_inflight_bytes'] / 2**20:.0f} MiB")
        r = plain(rate, args.plain_batch)
        row["plain"] = {k: r[k] for k in keys}
        log(f"serve plain @ {rate:.3g}/s: {r['committed_per_s']:.3g} txn/s, p50 {r['p50_ms']:.3f} ms, "
            f"p99 {r['p99_ms']:.3f} ms, {r['batches']} batches")
        sweep = []
        for b in args.plain_sweep:
            r = plain(rate, b)
            sweep.append({"batch": b, **{k: r[k] for k in ("committed_per_s", "p50_ms", "p99_ms")}})
        p99 = row["served"]["p99_ms"]
        ok = [x for x in sweep if x["p99_ms"] <= p99]
        best = max(ok, key=lambda x: x["committed_per_s"]) if ok else min(sweep, key=lambda x: x["p99_ms"])
        row["plain_sweep"] = sweep
        row["plain_at_equal_p99"] = {**best, "p99_within_served": bool(ok)}
        row["served_over_plain_at_equal_p99"] = row["served"]["committed_per_s"] / best["committed_per_s"]
        log(f"serve @ {rate:.3g}/s: plain at p99 <= {p99:.2f} ms: batch {best['batch']}, "
            f"{best['committed_per_s']:.3g} txn/s (p99 {best['p99_ms']:.2f} ms); served "
            f"{row['served_over_plain_at_equal_p99']:.2f}x")
        out["rates"].append(row)
    out["capacity_at_p99"] = capacity_at_p99(out["rates"])
    for c in out["capacity_at_p99"]:
        log(f"serve capacity at p99 <= {c['p99_ms']:.2f} ms: served {c['served_rate']:.3g}/s, "
            f"plain {c['plain_rate']:.3g}/s (batch {c['plain_batch']})")
    return out


def small_batch_latency(e, rows: int, n_cols: int, width: int, reps: int = 200) -> dict:
    """Execute-to-completion latency of small ordered batches of the C2
    shape (16 ops per txn, zipf 0.99, 50/50 READ/UPDATE) on the C2 table,
    through the multi-kernel pipeline and through the one-launch kernel
    (batches of at most LTPMG_SMALL_MAX_OPS ops): host time of the execute
    call, wall time from execute to completion, device time."""
    from ltpmg_amd import MODE_O, OP_UPDATE, gen_kv
    out = []
    for txns in (16, 64, 256, 512):
        b = gen_kv(txns, rows, n_cols, ops=16, zipf_theta=0.99, read_pct=50, write_kind=OP_UPDATE,
                   seed=BATCH_SEED + 7)
        e.submit(b)
        row = {"txns": txns, "ops": b.n_ops}
        for name, pipe in (("pipeline", True), ("one_launch", False)):
            for _ in range(10):
                e.execute(MODE_O, pipeline=pipe)
                e.wait()
            host, wall, dev = [], [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                e.execute(MODE_O, pipeline=pipe)
                t1 = time.perf_counter()
                e.wait()
                wall.append(time.perf_counter() - t0)
                host.append(t1 - t0)
                dev.append(e.stats()["exec_ms"])
            row[name] = {"host_us": 1e6 * pct(host, 50), "p50_us": 1e6 * pct(wall, 50),
                         "p99_us": 1e6 * pct(wall, 99), "device_us": 1e3 * pct(dev, 50)}
        log(f"small batch {txns} txns ({b.n_ops} ops): pipeline {row['pipeline']['p50_us']:.0f} us "
            f"(device {row['pipeline']['device_us']:.0f}), one launch {row['one_launch']['p50_us']:.0f} us "
            f"(device {row['one_launch']['device_us']:.0f}), execute+wait p50")
        out.append(row)
    return {"reps": reps, "sizes": out}


def capacity_at_p99(rates: list) -> list:
    """For each served p99 P measured above: the highest offered rate of the
    sweep that each form carries (>= 95 % of the offered txn/s, the rest being the drain after the last arrival) at a p99 of
    at most P -- plain at its best batch size for that rate.  0 when no rate
    qualifies."""
    out = []
    for ref in rates:
        p99 = ref["served"]["p99_ms"]
        served = [r["arrival_rate"] for r in rates
                  if r["served"]["p99_ms"] <= p99 and r["served"]["committed_per_s"] >= 0.95 * r["arrival_rate"]]
        plain, batch = 0.0, None
        for r in rates:
            ok = [x for x in r["plain_sweep"]
                  if x["p99_ms"] <= p99 and x["committed_per_s"] >= 0.95 * r["arrival_rate"]]
            if ok and r["arrival_rate"] > plain:
                plain, batch = r["arrival_rate"], min(ok, key=lambda x: x["p99_ms"])["batch"]
        out.append({"p99_ms": p99, "served_rate": max(served, default=0.0), "plain_rate": plain,
                    "plain_batch": batch})
    return out


def batching_compare(args) -> dict:
    """The batching form against the plain per-batch path on one stream of
    uneven batches, all resident in HBM before timing:
      sizes    floor(exp(U(ln 1e3, ln 1e6))) txns per batch (fixed seed)
      batches  alternately the key-value table (ops per txn U{1..64}, zipf
               0.5 and 0.99 in turn, 50% READ / 50% UPDATE) and the W=64
               order-entry schema (45% NewOrder, 43% Payment, 12% read-only)
      plain    every batch alone through the three-kernel / ordered
               pipeline, host wait after each batch
      batching the plan ltpmg_stream_tune picks (merged ordered groups,
               persistent fused lowest-TID-wins kernel, one captured graph)
    Both modes are run over the whole stream.  Value: committed txns of one
    pass over the stream / device time of that pass."""
    from ltpmg_amd import MODE_A, MODE_O, Engine, Stream, StreamPlan
    from ltpmg_amd.engine import OeStream, create_oe_tables
    rng = np.random.default_rng(BATCH_SEED + 4)
    nb = args.batching_batches
    sizes = np.floor(np.exp(rng.uniform(np.log(args.batching_min_txns),
                                        np.log(args.batching_max_txns), nb))).astype(np.int64)
    # the order-entry batches append rows; every run of the stream rewinds
    # the append tables to their preloaded rows, so one pass needs room for
    # the inserts of the order-entry half of the stream
    W, cap, kv_tab, n_cols, width = 64, int(sizes[1::2].sum()), 9, 10, 100
    e = Engine()
    create_oe_tables(e, W, cap, TABLE_SEED)
    e.create_table(kv_tab, args.rows, [width] * n_cols)
    e.fill_table(kv_tab, TABLE_SEED)
    from ltpmg_amd import OP_UPDATE, gen_kv
    oe = OeStream(W, seed=BATCH_SEED + 5)
    batches, kinds = [], []
    t0 = time.perf_counter()
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            theta = 0.5 if i % 4 == 0 else 0.99
            batches.append(gen_kv(int(n), args.rows, n_cols, ops=(1, 64), zipf_theta=theta,
                                  read_pct=50, write_kind=OP_UPDATE, table_id=kv_tab,
                                  seed=BATCH_SEED + 100 + i))
            kinds.append(f"kv{theta}")
        else:
            batches.append(oe.next(int(n)))
            kinds.append("oe")
        if (i + 1) % 25 == 0:
            log(f"batching stream: {i + 1}/{nb} batches generated")
    txns = int(sizes.sum())
    ops = int(sum(b.n_ops for b in batches))
    log(f"batching stream: {nb} batches, {txns} txns, {ops} ops, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    out = {"batches": nb, "txns": txns, "ops": ops,
           "batch_txns_min": int(sizes.min()), "batch_txns_max": int(sizes.max()),
           "order_entry_partitions": W, "kv_rows": args.rows, "modes": {}}
    for name, mode in (("ordered", MODE_O), ("lowest_tid_wins", MODE_A)):
        with Stream(e, batches, mode, rewind=True) as st:
            plain_ms = [st.run(StreamPlan.plain()) for _ in range(2)]
            committed = int(st.committed().sum())
            plan, trials = st.tune()
            tuned = []
            for _ in range(args.batching_runs):
                w0 = time.perf_counter()
                ms = st.run()
                tuned.append((ms, time.perf_counter() - w0))
            assert int(st.committed().sum()) == committed
            n_groups = len(st.groups())
        p_ms = min(plain_ms)
        b_ms = statistics.median(t[0] for t in tuned)
        row = {"committed": committed,
               "plain": {"device_ms": p_ms, "committed_txn_per_s": committed / (p_ms * 1e-3)},
               "batching": {"device_ms": b_ms, "committed_txn_per_s": committed / (b_ms * 1e-3),
                            "wall_ms": 1e3 * statistics.median(t[1] for t in tuned),
                            "launch_groups": n_groups, "tuning_runs": trials,
                            "plan": plan.__dict__},
               "speedup": p_ms / b_ms}
        out["modes"][name] = row
        log(f"batching {name}: plain {p_ms:.1f} ms, tuned {b_ms:.1f} ms "
            f"({row['speedup']:.2f}x, {n_groups} groups, plan {plan})")
    e.close()
    return out


def serve_c5(args) -> dict:
    """Serving on the large schema: open-loop Poisson arrivals of single
    transactions drawn from the order-entry mix (45% NewOrder 5-15 lines,
    43% Payment, 12% read-only; 15% of Payments by last name) over W=1024
    partitions, with a 100M-row x 10 x 100 B key-value table also resident
    (about 185 GB of tables with 128-byte slots).  NewOrder and Payment
    append rows; the append tables hold 100 orders per district at the start
    (the ordered mode numbers at most 2^32 cells) and room for 2^22
    transactions' inserts, and are truncated back to the preloaded rows
    before each run.  Ordered mode.  The engine's HBM budget bounds the
    slots, which are sized before the clock starts.  Forms: plain (fixed
    batches, one in flight) at two batch sizes, served (continuous batching,
    two in flight)."""
    from ltpmg_amd import MODE_O, Engine
    from ltpmg_amd.engine import OeStream, create_oe_tables
    W, cap, pre, kv_rows = 1024, 1 << 22, 100, args.c5_kv_rows
    budget = int(args.c5_hbm_budget_gb * (1 << 30))
    e = Engine(n_slots=2, hbm_budget=budget)
    t0 = time.perf_counter()
    _, append = create_oe_tables(e, W, cap, TABLE_SEED, preload=pre)
    e.create_table(9, kv_rows, [100] * 10)
    e.fill_table(9, TABLE_SEED)
    tables_gb = e.stats_hbm() / (1 << 30)
    log(f"serving (large schema): {tables_gb:.1f} GiB of tables filled in {time.perf_counter() - t0:.1f} s")
    pool = OeStream(W, seed=BATCH_SEED + 9, byname_pct=15, table_seed=TABLE_SEED, preload=pre).next(args.c5_pool)
    out = {"partitions": W, "kv_rows": kv_rows, "tables_gib": tables_gb, "hbm_budget_gib": args.c5_hbm_budget_gb,
           "pool_txns": pool.n_txns, "pool_ops": pool.n_ops, "mode": "ordered", "rates": []}
    for rate in args.c5_rates:
        n = int(max(1 << 17, min(1 << 22, rate * args.c5_seconds)))
        row = {"arrival_rate": rate, "arrivals": n}
        forms = [(f"plain_{b}", dict(max_batch=b, min_batch=b, n_inflight=1)) for b in args.c5_plain_batches]
        forms.append(("served", dict(max_batch=args.c5_served_max, max_wait_us=args.c5_wait_us, n_inflight=2)))
        for name, kw in forms:
            for t, k in append.items():
                e.truncate_table(t, k)
            r = e.serve(pool, arrival_rate=rate, n_arrivals=n, mode=MODE_O, seed=11, **kw)
            row[name] = {k: r[k] for k in ("committed_per_s", "p50_ms", "p99_ms", "p999_ms",
                                           "mean_batch_txns", "batches", "host_pack_s", "device_ms")}
            log(f"serve[large] {name} @ {rate:.3g}/s: {r['committed_per_s']:.3g} txn/s, "
                f"p50 {r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, p999 {r['p999_ms']:.2f} ms")
        out["rates"].append(row)
    out["hbm_allocated_gib"] = e.stats_hbm() / (1 << 30)
    e.close()
    return out


def pct(xs, q):
    xs = sorted(xs)
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def host_cpu() -> dict:
    """The host's CPU model, core counts and this process's CPU quota."""
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            quota = None if q == "max" else int(q) / int(per)
    except (OSError, ValueError):
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "cgroup_cpus": quota}


def _time_cpu(fn, runs: int) -> tuple[float, float]:
    times = [fn() for _ in range(runs)]
    return statistics.median(times), sum(times)


def _sample(batch, s: int):
    from ltpmg_amd.engine import Batch
    n_ops = int(batch.txn_off[s])
    return Batch(batch.txn_off[: s + 1].copy(), batch.rows[:n_ops].copy(), batch.meta[:n_ops].copy(),
                 batch.args[:n_ops].copy())


def _cpu_legs(orc, im, batch, mode: int, stride: int, threads: int, one_core_txns: int,
              serial_txns: int, runs: int) -> dict:
    """The threaded restatement on `threads` threads (whole batch) and on 1
    thread, and the oracle's serial replay, each on its own sample."""
    out = {}
    for name, s, th in (("all_cores", batch.n_txns, threads), ("one_core", min(one_core_txns, batch.n_txns), 1)):
        sb = _sample(batch, s)
        med, tot = _time_cpu(lambda: orc.cpu_pipeline(im, sb, mode, stride,
                                                      n_threads=th)[2], runs)
        out[name] = {"value": s / med, "unit": "txn/s", "threads": th, "txns": s, "ops": sb.n_ops,
                     "median_s": med, "total_s": tot, "runs": runs}
    s = min(serial_txns, batch.n_txns)
    sb = _sample(batch, s)

    def serial():
        t0 = time.perf_counter()
        orc.replay(im, sb, mode, stride)
        return time.perf_counter() - t0
    med, tot = _time_cpu(serial, runs)
    out["serial_oracle"] = {"value": s / med, "unit": "txn/s", "threads": 1, "txns": s, "ops": sb.n_ops,
                            "median_s": med, "total_s": tot, "runs": runs,
                            "note": "orc_replay: commit rule then serial replay in TID order "
                                    "(wall time of the ctypes call)"}
    return out


def cpu_baseline(batch, rows: int, width: int, n_cols: int, runs: int) -> dict:
    """CPU baselines on the GPU box's host (rank 0, N=1), bounded samples:
      C2  the headline key-value batch, ordered mode: the threaded C++
          restatement of the reference pipeline on every core this process
          may use (whole batch) and on 1 core (first 2^16 txns), and the
          oracle's serial replay (first 2^16 txns);
      C1  the reference's CPU configuration (100k rows x 10 int32, 8192 txns
          x 10 ops, uniform keys, 50% READ of the row / 50% RMW +1, lowest
          TID wins): the same three legs on the whole batch.
    The reference's own CPU path cannot run here (its CPU executor is CUDA
    host code, cuda/src/gpulauncher.cu:708-753); "port" is the restatement."""
    import oracle_ffi as orc
    from ltpmg_amd import MODE_A, OP_RMW_ADD, gen_kv
    cpu = host_cpu()
    threads = cpu["affinity"]  # every core this process may run on, within its CPU quota
    if cpu["cgroup_cpus"]:
        threads = max(1, min(threads, int(np.ceil(cpu["cgroup_cpus"]))))
    # C2: only the cells the batch touches are filled (untouched pages stay unmapped)
    im = orc.Image({0: (rows, [width] * n_cols)})
    col = (batch.meta >> 2) & 63
    for c in range(n_cols):
        im.fill_cells(TABLE_SEED, 0, c, np.unique(batch.rows[col == c]))
    c2 = _cpu_legs(orc, im, batch, orc.MODE_O, 128, threads, 1 << 16, 1 << 16, runs)
    del im
    c1_batch = gen_kv(8192, 100_000, 10, ops=10, read_pct=50, write_kind=OP_RMW_ADD, read_all_cols=True,
                      rmw_delta=1, seed=BATCH_SEED + 11)
    im1 = orc.Image({0: (100_000, [4] * 10)}).fill(TABLE_SEED)
    c1 = _cpu_legs(orc, im1, c1_batch, MODE_A, 4, threads, 8192, 8192, max(runs, 9))
    head = c2["all_cores"]
    return {"value": head["value"], "unit": "txn/s", "cores": threads, "kind": "port",
            "sample": f"C2: the whole headline batch ({head['txns']} txns, {head['ops']} ops), ordered "
                      f"mode, same table seed; C++ restatement of the reference pipeline (histogram, "
                      f"scan, scatter, per-row TID sort, per-row apply) on {threads} threads, timed "
                      f"around CC + execute only, median of {runs} runs; one-core and serial-oracle "
                      f"legs and the C1 configuration beside it",
            **cpu, "c2": c2, "c1": c1}


def order_entry_c3(args) -> dict:
    """Plain batch on the order-entry schema: W = 64 partitions, one batch of
    --c3-txns transactions of the 45% NewOrder (5-15 lines, 1% remote) / 43%
    Payment / 12% read-only mix (15% of customer lookups by last name), both
    modes.  NewOrder and Payment append rows, so before each step the append
    tables are truncated back to their preloaded rows (outside the timed
    execute + wait)."""
    from ltpmg_amd import MODE_A, MODE_O, Engine
    from ltpmg_amd.engine import OeStream, create_oe_tables
    W, n = 64, args.c3_txns
    e = Engine()
    _, append = create_oe_tables(e, W, n, TABLE_SEED)
    batch = OeStream(W, seed=BATCH_SEED + 3, byname_pct=15, table_seed=TABLE_SEED).next(n)
    e.submit(batch)
    out = {"partitions": W, "txns": batch.n_txns, "ops": batch.n_ops, "modes": {}}
    for name, mode in (("ordered", MODE_O), ("lowest_tid_wins", MODE_A)):
        lat = []
        for k in range(args.warmup + args.c3_steps):
            for t, rows in append.items():
                e.truncate_table(t, rows)
            t0 = time.perf_counter()
            e.execute(mode)
            e.wait()
            if k >= args.warmup:
                lat.append(time.perf_counter() - t0)
        st = e.stats()
        med = statistics.median(lat)
        out["modes"][name] = {"committed": st["committed"], "inserted_rows": st["inserted"],
                              "committed_txn_per_s": st["committed"] / med, "step_ms_median": 1e3 * med,
                              "device_ms": st["exec_ms"], "steps": len(lat)}
        log(f"C3 {name}: {st['committed']}/{batch.n_txns} committed, {st['inserted']} rows appended, "
            f"{1e3 * med:.3f} ms per step, {st['committed'] / med:.4g} committed txn/s")
    e.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=10_000_000)
    ap.add_argument("--txns", type=int, default=1 << 20)
    ap.add_argument("--ops", type=int, default=16)
    ap.add_argument("--theta", type=float, default=0.99)
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[1e6, 5e6, 2e7, 5e7, 1e8, 2e8],
                    help="arrival rates (txn/s) for the served-vs-plain comparison; none to skip")
    ap.add_argument("--serve-arrivals", type=int, default=1 << 22)
    ap.add_argument("--serve-seconds", type=float, default=0.25,
                    help="arrival window per rate (capped at --serve-arrivals arrivals)")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the small-batch latency comparison")
    ap.add_argument("--plain-batch", type=int, default=1 << 18)
    ap.add_argument("--serve-max-batch", type=int, default=1 << 18)
    ap.add_argument("--serve-wait-us", type=float, default=0.0,
                    help="served form: cut when the oldest waited this long (0: as soon as a slot is free)")
    ap.add_argument("--serve-hbm-mb", type=float, default=2048.0,
                    help="served form: device bytes the batches in flight may hold")
    ap.add_argument("--plain-sweep", type=int, nargs="*",
                    default=[1 << 12, 1 << 13, 1 << 14, 1 << 15, 1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20])
    ap.add_argument("--no-c5", action="store_true",
                    help="skip serving on the large order-entry + key-value schema")
    ap.add_argument("--c5-kv-rows", type=int, default=100_000_000)
    ap.add_argument("--c5-hbm-budget-gb", type=float, default=240.0)
    ap.add_argument("--c5-pool", type=int, default=1 << 20)
    ap.add_argument("--c5-rates", type=float, nargs="*", default=[1e6, 5e6, 2e7, 5e7])
    ap.add_argument("--c5-seconds", type=float, default=0.5, help="arrival window per rate")
    ap.add_argument("--c5-plain-batches", type=int, nargs="*", default=[1 << 15, 1 << 17])
    ap.add_argument("--c5-served-max", type=int, default=1 << 17)
    ap.add_argument("--c5-wait-us", type=float, default=0.0)
    ap.add_argument("--no-c3", action="store_true", help="skip the order-entry plain batch")
    ap.add_argument("--c3-txns", type=int, default=1 << 20)
    ap.add_argument("--c3-steps", type=int, default=5)
    ap.add_argument("--no-batching", action="store_true",
                    help="skip the batching-form stream comparison")
    ap.add_argument("--batching-batches", type=int, default=200)
    ap.add_argument("--batching-min-txns", type=int, default=1000)
    ap.add_argument("--batching-max-txns", type=int, default=1_000_000)
    ap.add_argument("--batching-runs", type=int, default=3)
    ap.add_argument("--copy-selftest", type=int, default=0, metavar="BYTES",
                    help="only time the plain streaming-copy kernel over BYTES and exit "
                         "(the counter calibration run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the batched transaction path on a "
              "single MI355X (multi-GPU sharding is design-only); use --gpus 1.", file=sys.stderr)
        return 2

    from ltpmg_amd import MODE_A, MODE_O, OP_UPDATE, Engine, gen_kv

    if args.copy_selftest:
        e = Engine(n_slots=1)
        gbs = e.copy_bandwidth(args.copy_selftest, 5)
        print(json.dumps({"copy_bytes": args.copy_selftest, "iters": 6, "copy_gbs": gbs}), flush=True)
        e.close()
        return 0

    n_cols, width = 10, 100
    log(f"generating batch: {args.txns} txns x {args.ops} ops, zipf {args.theta}, {args.rows} rows")
    batch = gen_kv(args.txns, args.rows, n_cols, ops=args.ops, zipf_theta=args.theta, read_pct=50,
                   write_kind=OP_UPDATE, seed=BATCH_SEED)
    counts = batch_counts(batch, width)
    log(f"batch: {batch.n_ops} ops, {counts['written_cells']} written cells")
    e = Engine(n_slots=2)
    e.create_table(0, args.rows, [width] * n_cols)
    e.fill_table(0, TABLE_SEED)
    log("table filled")
    # the batch is uploaded before the timed region; the first upload pays
    # one-time costs, so it is repeated and the last one is the warm figure
    uploads = []
    for _ in range(3):
        t0 = time.perf_counter()
        e.submit(batch)
        e.wait()
        wall = time.perf_counter() - t0
        e.execute(MODE_O)  # the slot's stats need an executed batch
        e.wait()
        uploads.append((wall, e.stats()["upload_ms"]))
    upload_wall, upload_ms = uploads[-1]
    log(f"upload: {upload_ms:.2f} ms DMA warm ({uploads[0][1]:.2f} ms first), {upload_wall * 1e3:.1f} ms submit wall")

    res = {}
    for name, mode in (("ordered", MODE_O), ("lowest_tid_wins", MODE_A)):
        steps = args.steps if mode == MODE_O else max(1, args.steps // 2)
        warm = args.warmup if mode == MODE_O else 1
        r = run_mode(e, batch, mode, steps, warm)
        log(f"{name}: {steps} steps in {r['wall_s']:.3f} s, committed {r['committed']}")
        r["kernels_ms"], r["launches"] = profile_mode(e, mode)
        commit, _ = e.fetch(results=False)
        r["commit"] = commit
        res[name] = r

    head = res["ordered"]
    value = head["committed"] * args.steps / head["wall_s"]
    ms_per_step = 1000.0 * head["wall_s"] / args.steps

    # roofline of the dominant kernel of the headline mode
    ker, launches = head["kernels_ms"], head["launches"]
    avg = {k: ker[k] / launches[k] for k in ker}
    dom = max(avg, key=lambda k: avg[k])
    alg = algorithmic_bytes(dom, counts)
    achieved = alg / (avg[dom] * 1e-3) / 1e9 if alg else None
    copy_gbs = e.copy_bandwidth(1 << 30, 10)
    wkey = f"kv {args.rows}x{n_cols}x{width}B, {args.txns}x{args.ops} ops, zipf {args.theta}"
    traffic = pmc_traffic(wkey)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                "traffic": traffic.get(dom), "avg_launch_ms": avg[dom], "algorithmic_bytes": alg,
                "measured_copy_gbs": copy_gbs,
                "frac_of_measured_copy": achieved / copy_gbs if achieved and copy_gbs else None}
    log(f"dominant kernel {dom}: {avg[dom]:.3f} ms, {achieved:.0f} GB/s; copy kernel {copy_gbs:.0f} GB/s")
    per_kernel = {}
    for name, r in res.items():
        for k, ms in r["kernels_ms"].items():
            a = algorithmic_bytes(k, counts, r["commit"], batch)
            l = r["launches"][k]
            per_kernel[f"{name}/{k}"] = {
                "ms": ms, "launches": l,
                "gbs": (a * l / (ms * 1e-3) / 1e9) if a and ms > 0 else None,
                "traffic_per_launch": traffic.get(k)}

    small = None
    if not args.no_small:
        small = small_batch_latency(e, args.rows, n_cols, width)

    serving = None
    if args.serve_rates:
        serving = serve_compare(e, batch, args)

    e.close()
    order_entry = None
    if not args.no_c3:
        order_entry = order_entry_c3(args)
    batching = None
    if not args.no_batching:
        batching = batching_compare(args)
    serving_large = None
    if not args.no_c5:
        serving_large = serve_c5(args)

    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {host_cpu()}")
        cpu = cpu_baseline(batch, args.rows, width, n_cols, args.cpu_runs)
        log(f"cpu baseline: C2 {cpu['c2']['all_cores']['value']:.4g} txn/s on {cpu['cores']} threads, "
            f"{cpu['c2']['one_core']['value']:.4g} on 1, serial oracle {cpu['c2']['serial_oracle']['value']:.4g}; "
            f"C1 {cpu['c1']['all_cores']['value']:.4g} / {cpu['c1']['one_core']['value']:.4g} / "
            f"{cpu['c1']['serial_oracle']['value']:.4g}")

    modes = {}
    for name, r in res.items():
        steps = len(r["lat"])
        modes[name] = {
            "committed_txn_per_s": r["committed"] * steps / r["wall_s"],
            "committed": r["committed"], "aborted": r["aborted"],
            "device_ms_per_batch": r["exec_ms_last"],
            "p50_commit_latency_ms": 1000 * pct(r["lat"], 50),
            "p99_commit_latency_ms": 1000 * pct(r["lat"], 99),
        }
    out = {
        "metric": METRIC, "value": value, "unit": "txn/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded table fill and Zipf batch; no dataset)",
        "config": {"workload": f"plain batch, ordered mode: KV table {args.rows} rows x {n_cols} "
                               f"x {width} B fields, {args.txns} txns x {args.ops} ops, "
                               f"zipf {args.theta}, 50% READ / 50% UPDATE",
                   "rows": args.rows, "fields": n_cols, "field_bytes": width,
                   "batch_txns": args.txns, "ops_per_txn": args.ops, "zipf_theta": args.theta,
                   "mode": "ordered", "parallelism": "single GPU"},
        "p50_commit_latency_ms": modes["ordered"]["p50_commit_latency_ms"],
        "p99_commit_latency_ms": modes["ordered"]["p99_commit_latency_ms"],
        "modes": modes,
        "upload_ms": upload_ms, "upload_ms_first": uploads[0][1], "upload_wall_s": upload_wall,
        "upload_bytes": int(batch.n_ops * 14 + (batch.n_txns + 1) * 4),
        "roofline": roofline, "kernels": per_kernel, "cpu_baseline": cpu, "serving": serving,
        "small_batch": small,
        "order_entry_plain": order_entry, "batching": batching, "serving_large_schema": serving_large,
        "batch_counts": counts,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
