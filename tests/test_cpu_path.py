"""The restatement of the reference's CPU path (oracle/stc_cpu_path.c: the
threaded filter, reachability, update_meetings and update_schema of
src/tracing/process.cpp:452-459, 608-612) checked against the oracle run with
the same tree rules (schema_mode 2: build_schema's tree, update_schema_unit's
width rule):

* every step's leaves (boxes and counts) and pair set are the oracle's;
* the meetings it identifies are the oracle's records, direction pid < target
  (the CPU path keeps no MBR and applies no map filter; all points here are
  inside the map), for every pair whose home slot (Cantor key mod the table
  size, refinement.cpp:38) lies more than 16 slots from any other pair's:
  where probe chains form, the reset of expired buckets (meeting.cpp:73) lets
  a live pair be inserted again with a new start, which splits its meeting,
  and which pair takes a shared slot first depends on the threads' order;
* results do not depend on the thread count.
"""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import HostTrace, workload


def records_as_meetings(orc):
    keys, _ = orc.records()
    lo, hi = keys[:, 0], keys[:, 1]
    pid = (hi >> np.uint64(26)) & np.uint64((1 << 26) - 1)
    tgt = hi & np.uint64((1 << 26) - 1)
    end = (lo & np.uint64(0xFFFFFFFF)) + np.uint64(orc.end_time_min())
    dur = lo >> np.uint64(32)
    sel = pid < tgt
    out = np.stack([pid[sel], tgt[sel], (end - dur)[sel], end[sel]], 1).astype(np.int64)
    return set(map(tuple, out.tolist()))


@pytest.mark.parametrize("name,over,steps", [
    ("c1", dict(num_objects=6000, reach_distance=20.0, min_meet_time=3), 40),
    ("c1", dict(num_objects=12000, grid_capacity=12, grid_amplify=4.0, reach_distance=15.0, min_meet_time=3), 25),
])
def test_cpu_path_equals_oracle(name, over, steps):
    w = workload(name, dynamic_schema=1, **over)
    w["gen"].update(n_hotspots=4, hotspot_frac=0.6, hotspot_sigma=400.0)
    cfg = dict(w["cfg"])
    cfg["schema_mode"] = 2
    orc = O.Oracle(cfg, w["box"])
    B = 1 << 25
    cp = O.CpuPath(cfg, w["box"], threads=8, buckets=B)
    meet = set()
    leaf_sets = set()
    keys = set()
    for t, xy in HostTrace(w["gen"]).steps(steps):
        o = orc.step(xy, t)
        c = cp.step(xy, t)
        assert o["overflow"] == 0
        ob, oc = orc.leaves()
        cb, cc = cp.leaves()
        assert np.array_equal(ob, cb) and np.array_equal(oc, cc), t
        assert np.array_equal(orc.pairs(), cp.pairs()), t
        assert c["units"] == o["units_inside"] + o["units_cross"], t
        meet |= set(map(tuple, cp.meetings().astype(np.int64).tolist()))
        keys |= set(orc.pairs().tolist())
        leaf_sets.add(len(ob))
    # pairs whose home slot is near another pair's may be split by the table
    home = {}
    for k in keys:
        a, b = k >> 32, k & 0xFFFFFFFF
        home[(a, b)] = ((a + b) * (a + b + 1) // 2 + b) % B
    slots = np.sort(np.array(list(home.values()), dtype=np.int64))
    def lonely(ab):
        h = home[ab]
        i = np.searchsorted(slots, h)
        near = slots[max(i - 2, 0):i + 3]
        return np.count_nonzero(np.abs(near - h) <= 16) == 1
    expect = records_as_meetings(orc)
    keep = lambda ms: {m for m in ms if lonely((m[0], m[1]))}
    assert keep(meet) == keep(expect)
    assert len(keep(meet)) > 0.8 * len(expect) and len(expect) > 50 and len(leaf_sets) > 1


def test_cpu_path_thread_count_independent():
    w = workload("c1", num_objects=8000, reach_distance=20.0, min_meet_time=2, dynamic_schema=1)
    runs = []
    for threads in (1, 5, 16):
        cp = O.CpuPath(w["cfg"], w["box"], threads=threads, buckets=1 << 22)
        got = []
        for t, xy in HostTrace(w["gen"]).steps(15):
            cp.step(xy, t)
            m = cp.meetings()
            got.append((cp.pairs(), m[np.lexsort(m.T[::-1])]))
        runs.append(got)
        cp.close()
    for other in runs[1:]:
        for (pa, ma), (pb, mb) in zip(runs[0], other):
            assert np.array_equal(pa, pb) and np.array_equal(ma, mb)
