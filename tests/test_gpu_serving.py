"""Serving with continuous batching: query batches issued from one host thread
while another thread steps the same context. Batches run on the context's
query stream against the state the last completed step published, so they
overlap the running step on the device.

Each batch reports the timestep of the state it answered
(stc_query_batch_at). Afterwards a second context replays the same steps
alone and answers every query of every batch on its own, at the state that
batch saw: counts and hit lists must be identical, bit for bit. The first
test uses a small dense workload with a flush every few steps (so the
memtable epochs and the run list change under the queries); the second is
the c5 serving workload at its size (5M objects) with its query mix."""
import threading
import time

import numpy as np
import pytest

from stcdb_amd import HostTrace, Pipeline, workload
from stcdb_amd.engine import HIT_DTYPE, QF_COUNT, QUERY_DTYPE

pytestmark = pytest.mark.gpu


def random_queries(rng, n_objects, box, t_now, count, side_m=(20.0, 600.0), window=(1, 30), count_ranges=False):
    lon_m = 1.0 / 82900.0  # about a metre of longitude at 41.6 N, in degrees
    lat_m = 1.0 / 111320.0
    q = np.zeros(count, dtype=QUERY_DTYPE)
    u = rng.random(count)
    q["kind"] = np.where(u < 0.5, 1, np.where(u < 0.9, 2, 3))
    q["oid"] = rng.integers(0, n_objects, count)
    q["oid2"] = rng.integers(0, n_objects, count)
    w = rng.integers(window[0], window[1] + 1, count)
    q["t1"] = t_now
    q["t0"] = np.maximum(t_now - w, 0)
    (lo0, lo1), (hi0, hi1) = box
    side = np.exp(rng.uniform(np.log(side_m[0]), np.log(side_m[1]), count))
    cx, cy = rng.uniform(lo0, hi0, count), rng.uniform(lo1, hi1, count)
    q["box"] = np.stack([cx - side / 2 * lon_m, cy - side / 2 * lat_m, cx + side / 2 * lon_m,
                         cy + side / 2 * lat_m], 1)
    if count_ranges:
        q["flags"] = np.where(q["kind"] == 2, QF_COUNT, 0)
    return q


def serve_and_replay(cfg, box, frames, warm, rng, count_ranges, max_batch, hits_cap):
    gpu = Pipeline(cfg, box)
    for t in range(warm):
        gpu.step(frames[t], t)
    n = cfg["num_objects"]
    step_spans, done = [], threading.Event()
    errors = []

    def ingest():
        try:
            for t in range(warm, len(frames)):
                t0 = time.perf_counter()
                gpu.step(frames[t], t)
                step_spans.append((t0, time.perf_counter()))
        except Exception as e:  # surfaced by the main thread
            errors.append(e)
        finally:
            done.set()

    th = threading.Thread(target=ingest)
    buf = np.zeros(hits_cap, dtype=HIT_DTYPE)
    recorded, spans = [], []
    th.start()
    t_seen = warm - 1
    while not done.is_set() or not recorded:
        q = random_queries(rng, n, box, t_seen + 1, int(rng.integers(1, max_batch + 1)), count_ranges=count_ranges)
        t0 = time.perf_counter()
        hits, counts, st = gpu.query_at(q, buf)
        spans.append((t0, time.perf_counter()))
        assert int(counts[q["flags"] == 0].sum()) <= len(buf)
        recorded.append((st, q, hits.copy(), counts.copy()))
        t_seen = st
    th.join()
    assert not errors, errors
    gpu.close()
    # the batches ran while steps ran
    overlap = sum(any(a < d and c < b for c, d in step_spans) for a, b in spans)
    print(f"[serving] {len(recorded)} batches, {overlap} of them overlapping a step, "
          f"states {sorted(set(r[0] for r in recorded))[:3]}..{max(r[0] for r in recorded)}", flush=True)
    assert overlap > 0
    # replay alone: every query on its own, at the state its batch saw
    by_t = {}
    for r in recorded:
        by_t.setdefault(r[0], []).append(r)
    ref = Pipeline(cfg, box)
    checked = hits_total = 0
    for t in range(len(frames)):
        ref.step(frames[t], t)
        for _, q, hits, counts in by_t.get(t, []):
            for i in range(len(q)):
                h1, c1 = ref.query(q[i:i + 1])
                assert int(c1[0]) == int(counts[i]), (t, i)
                mine = hits[hits["query"] == i].copy()
                mine["query"] = 0
                assert np.array_equal(h1.view(np.uint8), mine.view(np.uint8)), (t, i)
                checked += 1
                hits_total += len(mine)
    ref.close()
    print(f"[serving] {checked} queries equal alone and in the mix ({hits_total} hits)", flush=True)
    return checked, hits_total


def test_alone_equals_mix_with_flushes():
    w = workload("c1", num_objects=30000, min_meet_time=3, kv_restriction=4000, reach_distance=20.0,
                 dynamic_schema=1)
    w["gen"].update(n_hotspots=5, hotspot_frac=0.7, hotspot_sigma=250.0)
    cfg = dict(w["cfg"], device_runs=1)
    frames = [xy.copy() for _, xy in HostTrace(w["gen"]).steps(45)]
    checked, hits = serve_and_replay(cfg, w["box"], frames, 6, np.random.default_rng(17), False, 48, 1 << 21)
    assert checked > 200 and hits > 1000


def test_alone_equals_mix_c5_size():
    w = workload("c5", dynamic_schema=1)
    frames = [xy.copy() for _, xy in HostTrace(w["gen"]).steps(14)]
    checked, hits = serve_and_replay(w["cfg"], w["box"], frames, 10, np.random.default_rng(5), True, 64, 1 << 22)
    assert checked > 100
