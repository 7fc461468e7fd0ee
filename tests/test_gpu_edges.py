"""Edge cases of the pipeline, GPU against the oracle: a single object,
coincident objects (one leaf overflowing its buffer), points on QTree split
lines and on the map border, pairs within a few ulps of the reach distance
(decided on the host), ragged object counts, an input outside the map;
and, at the full 10M-object size, invariants that tie each step's counters
together."""
import math

import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import Pipeline, workload
from stcdb_amd._lib import STC_E_ARG, STC_E_INPUT, StcError
from stcdb_amd.engine import DeviceTrace

pytestmark = pytest.mark.gpu


def config(n, **over):
    over.setdefault("grid_capacity", 16)
    w = workload("c1", num_objects=n, min_meet_time=2, keep_pairs=1, kv_restriction=60, **over)
    return w["cfg"], w["box"]


def run_frames(cfg, box, frames, tag, min_pairs=0):
    gpu = Pipeline(cfg, box)
    orc = O.Oracle(cfg, box)
    print(f"[edges {tag}] n={cfg['num_objects']} steps={len(frames)}", flush=True)
    band = 0
    for t, xy in enumerate(frames):
        g = gpu.step(xy, t)
        o = orc.step(xy, t)
        band += g["ambiguous_pairs"]
        assert g["pairs"] >= min_pairs, (tag, t, g["pairs"])
        for k, ko in [("leaves", "leaves"), ("nodes", "nodes"), ("overflow_points", "overflow"),
                      ("active_meetings", "active"), ("expired", "expired"), ("kv_count", "kv_count"),
                      ("flushed", "flushed")]:
            assert g[k] == o[ko], (tag, t, k, g[k], o[ko])
        assert np.array_equal(gpu.pairs(), orc.pairs()), (tag, t)
        gb, gc = gpu.leaves()
        ob, oc = orc.leaves()
        assert np.array_equal(gb, ob) and np.array_equal(gc, oc), (tag, t)
        gk, gs, ge, gm = gpu.active()
        ok, os_, oe, om = orc.active()
        assert np.array_equal(gk, ok) and np.array_equal(gs, os_) and np.array_equal(ge, oe), (tag, t)
        assert np.array_equal(gm.view(np.uint32), om.view(np.uint32)), (tag, t)
    assert gpu.num_runs() == orc.num_runs()
    for r in range(gpu.num_runs()):
        a, b = gpu.run(r), orc.run(r)
        assert np.array_equal(a["key_bytes"], b["key_bytes"]) and np.array_equal(a["sids"], b["sids"]), (tag, r)
    gpu.close()
    orc.close()
    return band


def walk(box, n, steps, seed, step_m=3.0, spread=0.3):
    """n objects in a small square, moving a few metres per step"""
    rng = np.random.default_rng(seed)
    (lo0, lo1), (hi0, hi1) = box
    c = np.array([(lo0 + hi0) / 2, (lo1 + hi1) / 2])
    xy = c + rng.normal(0, spread * 1e-3, (n, 2))
    frames = []
    for _ in range(steps):
        frames.append(xy.copy())
        xy = xy + rng.normal(0, step_m * 1e-5, (n, 2))
    return frames


def test_two_objects_meet_and_part():
    # capacity 1: the tree splits until the two objects sit in different leaves or the finest cells
    cfg, box = config(2, grid_capacity=1)
    (lo0, lo1), (hi0, hi1) = box
    c = np.array([(lo0 + hi0) / 2 + 1e-5, (lo1 + hi1) / 2 - 1e-5])
    frames = []
    for t in range(14):
        gap = 5e-5 if 3 <= t < 10 else 5e-3  # within reach for steps 3..9
        frames.append(np.array([c, c + [gap, gap / 3]]))
    run_frames(cfg, box, frames, "two")


def test_too_few_objects_for_the_capacity_is_rejected():
    # the reference's rebuild never writes the root when it does not split (geometry.cu:1346-1350): rejected
    cfg, box = config(50, grid_capacity=200)
    xy = walk(box, 50, 1, 2)[0]
    gpu = Pipeline(cfg, box)
    with pytest.raises(StcError) as e:
        gpu.step(xy, 0)
    assert e.value.code == STC_E_ARG
    orc = O.Oracle(cfg, box)
    with pytest.raises(RuntimeError):
        orc.step(xy, 0)
    gpu.close()
    orc.close()


def test_coincident_objects_overflow_one_leaf():
    # 500 objects on one point: the leaf keeps its 400 lowest ids (the reference drops the rest)
    cfg, box = config(500)
    (lo0, lo1), (hi0, hi1) = box
    p = np.array([lo0 + 0.3 * (hi0 - lo0), lo1 + 0.6 * (hi1 - lo1)])
    frames = [np.tile(p + np.array([1e-6, -2e-6]) * (t % 3), (500, 1)) for t in range(8)]
    run_frames(cfg, box, frames, "coincident")


def test_points_on_split_lines_and_map_border():
    cfg, box = config(700)
    (lo0, lo1), (hi0, hi1) = box
    rng = np.random.default_rng(4)
    frames = []
    for t in range(10):
        # QTree split lines of the first levels, exact in double, and the low map border
        k = rng.integers(1, 16, (700, 2))
        xs = lo0 + (hi0 - lo0) * k[:, 0] / 16.0
        ys = lo1 + (hi1 - lo1) * k[:, 1] / 16.0
        xy = np.stack([xs, ys], axis=1)
        xy[::7, 0] = lo0
        xy[::11, 1] = lo1
        xy[::5] += rng.normal(0, 3e-5, (len(xy[::5]), 2))
        xy = np.clip(xy, [lo0, lo1], [np.nextafter(hi0, lo0), np.nextafter(hi1, lo1)])
        frames.append(xy)
    run_frames(cfg, box, frames, "split-lines")


def dpm_lon(lat):
    """degrees of longitude per metre at latitude `lat`, in the reference's
    operation order (util.h:38,57) with the C library's sin"""
    return 360.0 / (math.sin((90 - abs(lat)) * 3.14159265 / 180) * 40076.0 * 1000.0)


def test_pairs_at_the_reach_distance():
    """Pairs placed a few ulps from the reach distance: near the origin a
    coordinate ulp is far below the 16-ulp guard band around r, so many pairs
    fall inside it and are decided on the host; the pairs, meetings and runs
    must still equal the oracle's, which evaluates every pair exactly."""
    cfg, _ = config(600, grid_capacity=10)
    box = [[0.001, 0.001], [0.011, 0.011]]
    r = cfg["reach_distance"]
    rng = np.random.default_rng(12)
    frames = []
    for t in range(8):
        p = rng.uniform(0.002, 0.010, (300, 2))
        th = rng.uniform(0, 2 * np.pi, 300)
        q = np.empty_like(p)
        for i in range(300):
            dx = r * dpm_lon(p[i, 1]) * math.cos(th[i])
            dy = r * (360.0 / (40076.0 * 1000.0)) * math.sin(th[i])
            q[i] = p[i] + np.array([dx, dy])
            for d in range(2):  # a few ulps either way
                for _ in range(int(rng.integers(0, 6))):
                    q[i, d] = np.nextafter(q[i, d], np.inf if rng.random() < 0.5 else -np.inf)
        frames.append(np.concatenate([p, q]))
    band = run_frames(cfg, box, frames, "reach-distance")
    assert band > 0, "no pair fell inside the guard band"


@pytest.mark.parametrize("n", [63, 65, 4097])
def test_ragged_object_counts(n):
    cfg, box = config(n, grid_capacity=8)
    run_frames(cfg, box, walk(box, n, 12, n, spread=0.8), f"ragged-{n}")


def test_point_outside_the_map_is_an_input_error():
    cfg, box = config(100)
    gpu = Pipeline(cfg, box)
    xy = walk(box, 100, 1, 3)[0]
    gpu.step(xy, 0)
    xy[17, 0] = box[1][0] + 1e-3
    with pytest.raises(StcError) as e:
        gpu.step(xy, 1)
    assert e.value.code == STC_E_INPUT
    gpu.close()


def test_full_size_step_invariants():
    """10M objects: each step's counters agree with each other and with the
    previous step (every pair opens or continues a meeting; expired meetings
    are the previous step's live ones that did not continue; records grow by
    two per emitted meeting)"""
    w = workload("c3")
    pipe = Pipeline(w["cfg"], w["box"])
    trace = DeviceTrace(w["gen"])
    prev = None
    for t in range(14):
        r = pipe.step_device(trace.next(t), t)
        pipe.drop_runs()
        print(f"[edges full] t={t} pairs={r['pairs']} active={r['active_meetings']} kv={r['kv_count']}", flush=True)
        assert r["ambiguous_nodes"] == 0
        # every live meeting was opened or continued by a pair of this step; a pair found
        # twice in one step (an ulp sliver between leaves) may count twice as continuing
        dup = r["new_keys"] + r["continuing"] - r["active_meetings"]
        assert 0 <= dup <= r["pairs"] + r["ambiguous_pairs"] - r["active_meetings"]
        assert r["table_occupied"] >= r["active_meetings"]
        if prev is not None:
            assert 0 <= r["expired"] - (prev["active_meetings"] - r["continuing"]) <= dup
            if not prev["flushed"]:
                assert r["kv_count"] == prev["kv_count"] + r["new_records"]
            assert r["new_records"] % 2 == 0 and r["new_records"] <= 2 * r["expired"]
        prev = r
    assert prev["flushed"] and prev["new_records"] > 0
    pipe.close()
    trace.close()


def test_one_dense_cluster_every_segment_length():
    """5,000 objects within a few metres: every pair is within reach, so the
    owners' segments take every length from 0 to 4,999 slots (the place pass,
    the wave sort of 513-1024, the block sort of 1025-4096 and the sort over
    global scratch beyond), and the first step's pair sources overflow their
    initial buffer (grown and collected again)"""
    cfg, box = config(5000, grid_capacity=200)
    frames = walk(box, 5000, 4, seed=9, step_m=0.2, spread=0.01)
    run_frames(cfg, box, frames, "dense-cluster", min_pairs=10_000_000)  # ~95 % of the 12.5M possible pairs
