#!/usr/bin/env python3
"""Ingest throughput of the per-timestep co-occurrence pipeline on one MI355X.

A step is one timestep of the workload: every object's position goes through
the QTree build and partition, the border lookup, pair refinement fused with
the meeting-table update, meeting identification and, whenever the memtable
exceeds its restriction, the sort-based flush into a run of CTFs
(process_with_gpu in the reference, src/cuda/geometry.cu:1573-2218).

The default workload is the 10,000,000-object street-grid walk with 20
Gaussian hotspots (r = 20 m, minimum meeting 10 s, runs of 8 Mi records in
256 CTFs) named by the metric "point-updates/s ingested at 10M objects". The
trace is generated on the device before timing, so the timed region holds
only pipeline steps on points already resident in HBM.

Output: one JSON line: metric and value; the roofline of the dominant
kernel from HIP events on the library's stream, with every stage's
algorithmic GB/s; the reference's CPU path timed on the host's cores on a
bounded sample of the same workload; and a serving phase (5M objects at 1 Hz
plus 2,000 queries/s with every hit materialised) reporting p50/p99.
"""
import argparse
import collections
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemcpy.restype = ctypes.c_int
    return lib


def mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def gen_rand(seed, oid, step, salt):
    """stc_rand of stcdb_amd/csrc/stc_gen.h, vectorised over oid"""
    with np.errstate(over="ignore"):
        h = mix64(np.uint64(seed) ^ (np.uint64(salt) * np.uint64(0xD1B54A32D192ED03)))
        h = mix64(h ^ oid)
        return mix64(h ^ (np.uint64(step) * np.uint64(0x8CB92BA72F3D8DD7)))


def cpu_sample_ids(gen):
    """A bounded sample of the workload that keeps its density: every object
    of hotspot 0 and every 20th of the uniformly started objects."""
    n = gen["num_objects"]
    oid = np.arange(n, dtype=np.uint64)
    if gen["n_hotspots"] == 0:
        return np.nonzero(oid % np.uint64(20) == 0)[0]
    u = (gen_rand(gen["seed"], oid, 0, 3) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    hot = u < gen["hotspot_frac"]
    hid = gen_rand(gen["seed"], oid, 0, 4) % np.uint64(gen["n_hotspots"])
    return np.nonzero((hot & (hid == 0)) | (~hot & (oid % np.uint64(20) == 0)))[0]


def cpu_threads():
    """the host threads a CPU baseline may use: the process's CPUs, capped by
    OMP_NUM_THREADS (the share of the host this process is given)"""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(cap))) if cap and cap.isdigit() else n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(w, frames_host, steps, schema):
    """The reference's CPU path (tracer::process without -g: filter,
    reachability, update_meetings, update_schema on pthreads; restated in
    oracle/stc_cpu_path.c) on a bounded sample of the same workload. The
    first step builds the tree (once, as build_schema does before the loop)
    and is not timed; the rate is sample objects x timed steps / the sum of
    the four stages' wall times."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ids = cpu_sample_ids(w["gen"])  # frames_host holds these rows, in this order
    cfg = dict(w["cfg"])
    cfg["num_objects"] = int(len(ids))
    cfg["schema_mode"] = 2 if schema == "dynamic" else 0
    threads = cpu_threads()
    buckets = 1 << 27  # ~0.37 of the sample's ~49M pairs per step: no probe chain near a full table
    cp = O.CpuPath(cfg, w["box"], threads=threads, buckets=buckets, keep_pairs=False)
    cp.step(frames_host[0], 0)
    stage = {k: 0.0 for k in O.CpuPath.STAGES}
    pairs = 0
    for t in range(1, steps + 1):
        r = cp.step(frames_host[t], t)
        pairs += r["pairs"]
        for k, v in r["ms"].items():
            stage[k] += v
    cp.close()
    total_s = sum(stage.values()) / 1e3
    return dict(value=round(len(ids) * steps / total_s, 1), unit="point-updates/s", cores=threads, kind="port",
                cpu_model=cpu_model(), nproc=os.cpu_count(),
                ms_per_step={k: round(v / steps, 2) for k, v in stage.items()},
                sample=f"reference CPU path (filter, reachability, update_meetings, update_schema; "
                       f"{threads} threads, {buckets} meeting buckets) on {len(ids)} objects of the same workload "
                       f"(all of hotspot 0 and every 20th uniformly started object, density preserved), "
                       f"{steps} timesteps after the first; {pairs} pairs found; {total_s:.1f} s of stages")


def cpu_flush_baseline(w, frames_host, steps):
    """A CPU flush baseline, timed apart from the CPU path (which has no
    flush): the oracle's single-thread flush (STR sort, record placement, CTF
    widths, bitmaps, key packing) on every 10th object of the CPU sample, with
    a flush at every step (kv_restriction 1) and records from the third step
    on (minimum meeting 1 s). Reports records flushed per second of flush."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    sub = [f[::10].copy() for f in frames_host[:steps + 1]]
    cfg = dict(w["cfg"], num_objects=int(len(sub[0])), kv_restriction=1, min_meet_time=1,
               schema_mode=1 if w["cfg"].get("dynamic_schema") else 0)
    orc = O.Oracle(cfg, w["box"])
    ms, recs, flushes = 0.0, 0, 0
    for t, xy in enumerate(sub):
        r = orc.step(xy, t)
        if r["flushed"]:
            m, k = orc.last_flush()
            ms += m
            recs += k
            flushes += 1
    orc.close()
    if not flushes or ms <= 0:
        return None
    return dict(value=round(recs / (ms * 1e-3), 1), unit="records/s", cores=1, kind="port",
                ms_per_flush=round(ms / flushes, 2), records_per_flush=int(recs / flushes),
                sample=f"the oracle's flush on {len(sub[0])} objects of the same workload (every 10th of the CPU "
                       f"path's sample), {flushes} flushes of {steps + 1} steps; {ms / 1e3:.1f} s of flush")


def measured_traffic(kernel, workload, n):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter
    summary for this workload (FETCH_SIZE and WRITE_SIZE, calibrated with a copy
    of known size); None when no summary covers it."""
    path = os.path.join(PROFILES, f"pipeline_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and line.startswith("- `"):
            k, v = line[3:].split("`", 1)[0], line.split(":", 1)[1].strip()
            vals[k] = v
    try:
        if int(vals["objects"]) != n:
            return None
        return int(vals[f"{kernel}.traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def make_query_stream(rng, n_objects, box, t_now, count):
    """C5's query mix: 50 % object history, 40 % square range, 10 % pair
    lookup; time windows log-uniform in [10 s, 1 h] ending at t_now (a scalar or
    one step per query); range sides
    log-uniform in [100 m, 5 km] centred uniformly on the map"""
    from stcdb_amd.engine import QF_COUNT, QUERY_DTYPE, Q_OID, Q_PAIR, Q_RANGE
    from stcdb_amd.workloads import DPM_LAT, dpm_lon
    q = np.zeros(count, dtype=QUERY_DTYPE)
    u = rng.random(count)
    q["kind"] = np.where(u < 0.5, Q_OID, np.where(u < 0.9, Q_RANGE, Q_PAIR))
    # range co-occurrence queries report exact counts, as the reference's box search does
    # (search_count, workbench.cpp:478-572); object and pair queries return every hit
    q["flags"] = np.where(q["kind"] == Q_RANGE, QF_COUNT, 0)
    q["oid"] = rng.integers(0, n_objects, count)
    q["oid2"] = rng.integers(0, n_objects, count)
    window = np.exp(rng.uniform(np.log(10.0), np.log(3600.0), count)).astype(np.int64)
    q["t1"] = t_now
    q["t0"] = np.maximum(t_now - window, 0)
    (lo0, lo1), (hi0, hi1) = box
    side = np.exp(rng.uniform(np.log(100.0), np.log(5000.0), count))
    cx, cy = rng.uniform(lo0, hi0, count), rng.uniform(lo1, hi1, count)
    hx, hy = side / 2 * dpm_lon(lo1), side / 2 * DPM_LAT
    q["box"] = np.stack([cx - hx, cy - hy, cx + hx, cy + hy], axis=1)
    return q


def answer(pipe, q, buf):
    """a query batch over the last published state: every hit of its object and
    pair queries in host memory (the buffer grows when short), range queries
    counted; a batch that admission by free device memory refuses is split"""
    import stcdb_amd._lib as L
    try:
        hits, counts, st = pipe.query_at(q, buf[0])
    except L.StcError as e:
        if e.code != L.STC_E_CAPACITY or len(q) == 1:
            raise
        a = answer(pipe, q[:len(q) // 2], buf)
        b = answer(pipe, q[len(q) // 2:], buf)
        return a[0] + b[0], a[1] + b[1], min(a[2], b[2])
    stored = int(counts[q["flags"] == 0].sum())
    if stored > len(buf[0]):
        buf[0] = np.zeros(stored + stored // 2, dtype=HIT_DTYPE_BENCH())
        hits, counts, st = pipe.query_at(q, buf[0])
    return stored, int(counts.sum()), st


def HIT_DTYPE_BENCH():
    from stcdb_amd.engine import HIT_DTYPE
    return HIT_DTYPE


def serve(args, objects=0, schema="dynamic"):
    """Continuous batching: an ingest thread steps the 5M-object walk at 1 Hz
    while this thread answers a Poisson stream of queries; every query that
    has arrived when the previous batch returns joins the next batch, which
    runs on the library's query stream over the state of the last completed
    step, overlapping a step in flight. Latency = answer time - arrival time on
    the wall clock."""
    import threading
    from stcdb_amd import DeviceTrace, Pipeline, workload
    over = {"num_objects": objects} if objects else {}
    w = workload("c5", dynamic_schema=1 if schema == "dynamic" else 0, **over)
    cfg, gen = w["cfg"], w["gen"]
    n = cfg["num_objects"]
    pipe = Pipeline(cfg, w["box"])
    trace = DeviceTrace(gen)
    rng = np.random.default_rng(5)
    for t in range(args.warmup):
        r = pipe.step_device(trace.next(t), t)
        print(f"[serve] warmup step {t}: {r['ms_total']:.1f} ms, {r['pairs']} pairs, runs {pipe.num_runs()}",
              file=sys.stderr, flush=True)
    T = args.serve_seconds
    k = int(rng.poisson(args.qps * T))
    arrivals = np.sort(rng.uniform(0.0, T, k))
    # each query asks about the window ending at the step current when it arrives
    queries = make_query_stream(rng, n, w["box"], args.warmup + np.floor(arrivals).astype(np.int64), k)
    step_spans, step_ms, errors = [], [], []
    t_start = time.perf_counter()

    def ingest():
        try:
            for s in range(T):
                t = args.warmup + s
                wait = s - (time.perf_counter() - t_start)
                if wait > 0:
                    time.sleep(wait)
                a = time.perf_counter() - t_start
                r = pipe.step_device(trace.next(t), t)
                step_spans.append((a, time.perf_counter() - t_start))
                step_ms.append(r["ms_total"])
                print(f"[serve] second {s}: step {r['ms_total']:.1f} ms, runs {pipe.num_runs()}", file=sys.stderr,
                      flush=True)
        except Exception as e:
            errors.append(e)

    th = threading.Thread(target=ingest)
    th.start()
    lat, batches, spans, hits_total, counted_total, done = [], [], [], 0, 0, 0
    buf = [np.zeros(1 << 21, dtype=HIT_DTYPE_BENCH())]
    while done < k:
        now = time.perf_counter() - t_start
        ready = int(np.searchsorted(arrivals, now, side="right"))
        if ready <= done:
            time.sleep(min(max(arrivals[done] - now, 0.0), 0.002))
            continue
        a = time.perf_counter() - t_start
        stored, counted, _ = answer(pipe, queries[done:ready], buf)
        fin = time.perf_counter() - t_start
        spans.append((a, fin))
        lat.extend((fin - arrivals[done:ready]) * 1000.0)
        batches.append(ready - done)
        hits_total += stored
        counted_total += counted
        done = ready
    th.join()
    if errors:
        raise errors[0]
    overlapping = sum(any(x < d and c < y for c, d in step_spans) for x, y in spans)
    lat = np.array(lat)
    out = {
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "queries": int(len(lat)),
        "qps": args.qps,
        "hits_materialised": hits_total,
        "hits_counted": counted_total,
        "batches": len(batches),
        "batches_overlapping_a_step": overlapping,
        "mean_batch": round(float(np.mean(batches)), 1),
        "ingest_ms_per_step": round(float(np.mean(step_ms)), 3),
        "objects": n,
        "seconds": T,
        "warmup_steps": args.warmup,
        "runs_in_hbm": pipe.num_runs(),
        "schema": schema,
    }
    pipe.close()
    trace.close()
    return out


def batched(args):
    """The batched-streams workload (c4): 64 independent streams (log-normal
    sizes, median 50k objects, ~6M in total; each its own box, reach and leaf
    capacity), one library context each. Each stream's leaf capacity is first
    chosen by the tuner on the stream's own first frames; then the streams
    step together through stc_step_batch (host threads, each context on its
    own streams) and, for comparison, one after another as 64 plain sequential
    runs would. Frames are generated on the device before timing."""
    import torch
    from stcdb_amd import DeviceTrace, Pipeline, batched_streams, step_batch
    from stcdb_amd.tuner import CANDIDATES, tune_leaf_capacity
    streams = batched_streams()
    total = sum(w["cfg"]["num_objects"] for w in streams)
    W, K = args.warmup, args.steps
    nframes = W + 2 * K
    h = hip()
    frames = []
    for w in streams:
        n = w["cfg"]["num_objects"]
        f = torch.empty((nframes, n, 2), dtype=torch.float64, device="cuda")
        tr = DeviceTrace(w["gen"])
        for t in range(nframes):
            assert h.hipMemcpy(f[t].data_ptr(), tr.next(t).value, n * 16, 3) == 0
        tr.close()
        frames.append(f)
    torch.cuda.synchronize()
    print(f"[c4] {len(streams)} streams, {total} objects, {nframes} frames each resident", file=sys.stderr, flush=True)
    ptr = lambda i, t: ctypes.c_void_p(frames[i][t].data_ptr())
    t0 = time.perf_counter()
    tuned = []
    for i, w in enumerate(streams):
        best, _ = tune_leaf_capacity(w["cfg"], w["box"], [ptr(i, t) for t in range(args.tune_steps)],
                                     on_device=True)
        tuned.append(best)
    tune_s = time.perf_counter() - t0
    print(f"[c4] tuned in {tune_s:.1f} s: capacities {dict(collections.Counter(tuned))}", file=sys.stderr, flush=True)
    pipes = [Pipeline(dict(w["cfg"], grid_capacity=c), w["box"]) for w, c in zip(streams, tuned)]
    for t in range(W):
        step_batch(pipes, [ptr(i, t) for i in range(len(pipes))], t, threads=args.threads, on_device=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(W, W + K):  # one stream after another
        for i, p in enumerate(pipes):
            p.step_device(ptr(i, t), t)
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for t in range(W + K, W + 2 * K):  # all streams at once
        step_batch(pipes, [ptr(i, t) for i in range(len(pipes))], t, threads=args.threads, on_device=True)
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t0
    for p in pipes:
        p.close()
    line = {
        "metric": "point-updates/s over 64 batched streams (c4)",
        "value": round(total * K / batch_s, 1),
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(batch_s * 1000 / K, 3),
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (64 seeded street-grid walks generated on the device)",
        "sequential_value": round(total * K / seq_s, 1),
        "sequential_ms_per_step": round(seq_s * 1000 / K, 3),
        "threads": args.threads,
        "tuner": {"candidates": list(CANDIDATES), "steps": args.tune_steps, "seconds": round(tune_s, 1),
                  "chosen": dict(collections.Counter(tuned))},
        "config": {"workload": "c4", "streams": len(streams), "objects": total,
                   "objects_min": min(w["cfg"]["num_objects"] for w in streams),
                   "objects_max": max(w["cfg"]["num_objects"] for w in streams)},
    }
    print(json.dumps(line), flush=True)


def compact(args):
    """compaction of runs kept in HBM: the workload's steps run until --compact
    runs exist, then one stc_compact_runs merges them; reports records merged
    per second"""
    import torch
    from stcdb_amd import DeviceTrace, Pipeline, workload
    over = {"num_objects": args.objects} if args.objects else {}
    w = workload(args.workload, **over)
    cfg = dict(w["cfg"])
    cfg["device_runs"] = 1
    n = cfg["num_objects"]
    pipe = Pipeline(cfg, w["box"])
    trace = DeviceTrace(w["gen"])
    t = 0
    while pipe.num_runs() < args.compact:
        r = pipe.step_device(trace.next(t), t)
        print(f"[compact] step {t}: {r['ms_total']:.1f} ms, runs {pipe.num_runs()}", file=sys.stderr, flush=True)
        t += 1
    trace.close()
    k = args.compact
    infos = [pipe.run_info(i) for i in range(k)]
    records = sum(i["kv_count"] + i["oversize_count"] for i in infos)
    in_bytes = sum(i["key_bytes"] for i in infos)
    pipe.sync_runs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pipe.compact_runs(0, k)
    pipe.sync_runs()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    out = pipe.run_info(0)
    assert out["kv_count"] + out["oversize_count"] == records and pipe.num_runs() == 1
    line = {
        "metric": f"records/s merged by compaction of {k} runs at {n // 1000000}M objects",
        "value": round(records / elapsed, 1),
        "unit": "records/s",
        "ms": round(elapsed * 1000, 3),
        "records": records,
        "input_key_bytes": in_bytes,
        "output_key_bytes": out["key_bytes"],
        "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic (runs flushed from the seeded street-grid walk)",
        "config": {"workload": args.workload, "objects": n, "runs": k, "ctf_count": cfg["ctf_count"]},
    }
    print(json.dumps(line), flush=True)
    pipe.close()


def stage_roofline(outs, n):
    """per stage of a step: mean ms (HIP events on the library's stream) and
    the stage's algorithmic HBM bytes (points in, pairs and meetings out; the
    flush over the steps that flushed), as GB/s and a fraction of the peak"""
    def mean(k, rows=None):
        rows = outs if rows is None else rows
        return float(np.mean([r[k] for r in rows])) if rows else 0.0
    pairs = mean("pairs")
    slots = mean("table_buckets")
    flushed = [r for r in outs if r["flushed"]]
    recs = mean("new_records")
    stages = {
        # QTree descent and leaf placement: 16 B point in, 4 B leaf-ordered id out
        "partition": (mean("ms_partition"), 20.0 * n),
        # border lookup: the border points' 16 B positions
        "filter": (mean("ms_filter"), 16.0 * mean("border_points")),
        # pair refinement (reverse lists, collect, place, segment sorts): points in, 8 B pair keys out
        "refine": (mean("ms_refine"), 20.0 * n + 8.0 * pairs),
        # meetings: a 32 B bucket read and written per pair, the previous array scanned,
        # 2 x (16 B key + 16 B box) per recorded meeting
        "meetings": (mean("ms_meetings"), 64.0 * pairs + 32.0 * slots + 32.0 * recs),
    }
    if flushed:
        R = mean("kv_count", flushed)
        # STR 2 x 12 B per object, main sort 2 x 32 B, bitmap 16 B and encode 32 B + ~13 B packed per record
        # the flush's own device time: with async_flush a step reports the previous step's flush
        done = [r for r in outs if r["ms_flush_done"] > 0]
        stages["flush"] = (mean("ms_flush_done", done), 24.0 * n + (64.0 + 16.0 + 32.0 + 13.0) * R)
    out = {}
    for k, (ms, b) in stages.items():
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[k] = dict(ms=round(ms, 3), alg_bytes=b, gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--objects", type=int, default=0, help="override the object count")
    ap.add_argument("--schema", choices=["dynamic", "rebuild"], default="dynamic",
                    help="QTree mode: kept and updated across steps (the reference's default) or rebuilt every step "
                         "(--disable_dynamic_schema)")
    ap.add_argument("--flush", choices=["async", "sync"], default="sync",
                    help="a step's flush overlaps the next step up to its merge (stc_config.async_flush), or "
                         "completes inside the step (process_with_gpu's order)")
    ap.add_argument("--cpu-steps", type=int, default=40,
                    help="timed steps of the CPU baseline (about 0.25 s each on 16 threads: ~10 s of CPU work)")
    ap.add_argument("--host-steps", type=int, default=10,
                    help="extra steps fed from pinned host memory (PCIe-inclusive rate, reported beside value)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving phase of the default run")
    ap.add_argument("--serve", action="store_true",
                    help="serving only: 5M objects ingested at 1 Hz plus 2,000 queries/s; prints p50/p99 latency")
    ap.add_argument("--qps", type=float, default=2000.0)
    ap.add_argument("--serve-seconds", type=int, default=8)
    ap.add_argument("--serve-warmup", type=int, default=12)
    ap.add_argument("--threads", type=int, default=16, help="host threads of stc_step_batch (workload c4)")
    ap.add_argument("--tune-steps", type=int, default=4, help="frames each tuner candidate steps (workload c4)")
    ap.add_argument("--compact", type=int, default=0,
                    help="compaction mode: flush this many runs into HBM, then time merging them into one")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="copy a buffer of this size once before the steps (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.tune_steps < 2:
        ap.error("--tune-steps must be at least 2: the tuner times every frame after a candidate's first")
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (the pipeline is single-GPU; "
              f"see the sharding design for several GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.serve:
        sv = serve(argparse.Namespace(**dict(vars(args), warmup=args.serve_warmup)), args.objects, args.schema)
        print(json.dumps(dict({"metric": "p50/p99 query ms at 2k QPS with ingest at 1 Hz (5M objects)",
                               "value": sv["p99_ms"], "unit": "ms (p99)", "n_gpus": 1, "higher_is_better": False,
                               "data": "synthetic (seeded street-grid walk on the device; seeded Poisson queries)"},
                              **sv)), flush=True)
        return
    if args.compact:
        compact(args)
        return
    if args.workload == "c4":
        batched(args)
        return

    import torch
    from stcdb_amd import DeviceTrace, Pipeline, workload

    over = {"dynamic_schema": 1 if args.schema == "dynamic" else 0, "async_flush": 1 if args.flush == "async" else 0}
    if args.objects:
        over["num_objects"] = args.objects
    w = workload(args.workload, **over)
    cfg, gen = w["cfg"], w["gen"]
    n = cfg["num_objects"]
    total = args.warmup + args.steps
    nframes = total + args.host_steps

    # inputs for every step, generated on the device before timing
    frames = torch.empty((nframes, n, 2), dtype=torch.float64, device="cuda")
    trace = DeviceTrace(gen)
    h = hip()
    for t in range(nframes):
        p = trace.next(t)
        assert h.hipMemcpy(frames[t].data_ptr(), p.value, n * 16, 3) == 0
    trace.close()
    torch.cuda.synchronize()
    print(f"[bench] workload {args.workload}: {n} objects, {nframes} frames resident, schema {args.schema}",
          file=sys.stderr, flush=True)

    if args.copy_probe_mib:
        src = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {args.copy_probe_mib} MiB read, {args.copy_probe_mib} MiB written", file=sys.stderr,
              flush=True)
        del src, dst

    pipe = Pipeline(cfg, w["box"])
    for t in range(args.warmup):
        r = pipe.step_device(ctypes.c_void_p(frames[t].data_ptr()), t)
        pipe.drop_runs()
        print(f"[bench] warmup step {t}: {r['ms_total']:.1f} ms, {r['pairs']} pairs, {r['leaves']} leaves",
              file=sys.stderr, flush=True)

    outs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.warmup, total):
        outs.append(pipe.step_device(ctypes.c_void_p(frames[t].data_ptr()), t))
        # a flushed run's copy to pinned host memory (the reference's D2H of h_keys, sids,
        # bitmaps and oversize buffer, geometry.cu:2019-2035) is queued with the next step's
        # pair count and always completes; dropping the run only recycles its host block
        pipe.drop_runs()
    pipe.sync_runs()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for i, r in enumerate(outs):
        print(f"[bench] step {args.warmup + i}: {r['ms_total']:.1f} ms (collect {r['ms_count_kernel']:.1f}, "
              f"place {r['ms_refine_kernel']:.1f}, merge {r['ms_identify_kernel']:.1f}, flush {r['ms_flush_done']:.1f}, "
              f"schema {r['ms_schema']:.2f}) pairs {r['pairs']} records {r['new_records']} leaves {r['leaves']} "
              f"flushed {r['flushed']}", file=sys.stderr, flush=True)

    value = n * args.steps / elapsed

    # the same pipeline fed from pinned host memory: stc_step includes the H2D copy of the points
    host_value = None
    if args.host_steps:
        pinned = torch.empty((args.host_steps, n, 2), dtype=torch.float64, pin_memory=True)
        pinned.copy_(frames[total:total + args.host_steps])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.host_steps):
            pipe.step_host_ptr(ctypes.c_void_p(pinned[i].data_ptr()), total + i)
            pipe.drop_runs()
        pipe.sync_runs()
        torch.cuda.synchronize()
        host_value = n * args.host_steps / (time.perf_counter() - t0)
        del pinned
    del frames
    pipe.close()
    torch.cuda.empty_cache()

    # roofline of the dominant kernel (algorithmic bytes per launch / its HIP-event time)
    def mean(k):
        return float(np.mean([r[k] for r in outs]))

    pairs = mean("pairs")
    slots = mean("table_buckets")  # pair slots of a step = buckets of its meeting array
    kernels = {
        # collect pass: the N points in (16 B position + 4 B id), a 4 B count per owner out, one
        # 4 B source (the other point's index) per slot out
        "k_pair_collect": ("pair collect pass: each combination decided once, sources per owner", mean("ms_count_kernel"),
                           24.0 * n + 4.0 * slots),
        # place pass: each owner's 8 B run descriptor and 4 B sources in, one 12 B slot (the
        # partner and its float position) per pair out
        "k_pair_place": ("pair place pass: owner segments sorted by partner", mean("ms_refine_kernel"),
                         8.0 * n + 16.0 * slots),
        # merge: the step's slots (12 B) and the previous meetings (28 B) in, the owners' float
        # positions (8 B each) and segment offsets (4 B each) once, the new meetings (28 B) and
        # 2 records (32 B each) per recorded meeting out
        "k_meet_merge": ("merge with the previous step's meetings, expiry, records", mean("ms_identify_kernel"),
                         12.0 * slots + 28.0 * slots + 28.0 * slots + 12.0 * n + 32.0 * mean("new_records")),
    }
    kname = max(kernels, key=lambda k: kernels[k][1])
    desc, ms, alg = kernels[kname]
    achieved = alg / (ms * 1e-3) / 1e9
    roof = dict(bound="hbm", kernel=f"{kname} ({desc})", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(kname, args.workload, n),
                alg_bytes_per_launch=alg, avg_ms=round(ms, 3),
                kernels={k: dict(ms=round(v[1], 3), alg_bytes=v[2], gbs=round(v[2] / (v[1] * 1e-3) / 1e9, 1),
                                 traffic=measured_traffic(k, args.workload, n))
                         for k, v in kernels.items()},
                stages=stage_roofline(outs, n))

    cpu = cpu_flush = None
    flushed = [r for r in outs if r["flushed"]]
    gpu_flush = None
    if flushed:
        fr = float(np.mean([r["kv_count"] for r in flushed]))
        fm = float(np.mean([r["ms_flush_done"] for r in outs if r["ms_flush_done"] > 0]))
        gpu_flush = dict(records_per_s=round(fr / (fm * 1e-3), 1), ms_per_flush=round(fm, 3), records_per_flush=int(fr))
    if not args.no_cpu:
        trace = DeviceTrace(gen)
        frames_host = []
        ids = cpu_sample_ids(gen)
        buf = np.empty((n, 2), dtype=np.float64)
        for t in range(args.cpu_steps + 1):
            p = trace.next(t)
            assert h.hipMemcpy(buf.ctypes.data, p.value, n * 16, 2) == 0
            frames_host.append(np.ascontiguousarray(buf[ids]))  # the sample's rows only
        trace.close()
        cpu = cpu_baseline(w, frames_host, args.cpu_steps, args.schema)
        cpu_flush = cpu_flush_baseline(w, frames_host, 12)
        del frames_host
        print(f"[bench] cpu baseline: {cpu['value']:.0f} point-updates/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    serving = None
    if not args.no_serve:
        sargs = argparse.Namespace(**dict(vars(args), warmup=args.serve_warmup))
        serving = serve(sargs, 0, args.schema)

    line = {
        "metric": "point-updates/s ingested at 10M objects",
        "value": round(value, 1),
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1000 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded street-grid walk generated on the device)",
        "config": {"workload": args.workload, "objects": n, "schema": args.schema, "flush": args.flush,
                   "reach_m": cfg["reach_distance"],
                   "min_meet_s": cfg["min_meet_time"], "leaf_capacity": cfg["grid_capacity"],
                   "kv_restriction": cfg["kv_restriction"], "ctf_count": cfg["ctf_count"],
                   "pairs_per_step": int(pairs),
                   "flushes": int(sum(r["flushed"] for r in outs))},
        "host_input_value": round(host_value, 1) if host_value else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "flush": {"gpu": gpu_flush, "cpu_baseline": cpu_flush},
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
