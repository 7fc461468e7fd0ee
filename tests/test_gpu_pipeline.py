"""GPU parity of the per-step pipeline against the CPU oracle, through the C ABI.

Every step compares, bit for bit: the QTree leaves (boxes and member counts),
the set of pairs within reach, the active meeting table (key, start, end,
float MBR), and at the end the memtable records (128-bit keys and boxes) as
sorted sets.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import DeviceTrace, HostTrace, Pipeline, workload

pytestmark = pytest.mark.gpu


def sorted_records(keys, boxes):
    order = np.lexsort((keys[:, 0], keys[:, 1]))
    return keys[order], boxes[order]


def schema_set(nodes_boxes, types, children):
    """a QTree as {box: (type, child boxes)} over its valid slots"""
    out = {}
    for i in np.nonzero(types != 0)[0]:
        b = tuple(nodes_boxes[i].tolist())
        out[b] = (int(types[i]), tuple(tuple(nodes_boxes[c].tolist()) for c in children[i]) if types[i] == 1 else None)
    return out


def gpu_schema(gpu):
    q = gpu.schema()
    return schema_set(q["mbr"], q["type"].astype(np.int64), q["children"])


def oracle_schema(orc):
    boxes, mids, typ, ch = orc.nodes()
    return schema_set(boxes, typ.astype(np.int64), ch)


def run_parity(w, steps, check_every=1, tag="", frames=None):
    """the library and the oracle on the same steps (frames: an iterable of
    (t, xy); by default the workload's host trace)"""
    cfg, box = dict(w["cfg"]), w["box"]
    cfg["keep_pairs"] = 1
    print(f"[parity {tag}] n={cfg['num_objects']} cap={cfg['grid_capacity']} r={cfg['reach_distance']} "
          f"steps={steps} dynamic={cfg.get('dynamic_schema', 0)}", flush=True)
    gpu = Pipeline(cfg, box)
    orc = O.Oracle(cfg, box)
    dyn = cfg.get("dynamic_schema", 0)
    totals = dict(pairs=0, records=0, ambiguous=0, leaves=set())
    for t, xy in (frames if frames is not None else HostTrace(w["gen"]).steps(steps)):
        g = gpu.step(xy, t)
        o = orc.step(xy, t)
        assert g["leaves"] == o["leaves"], t
        assert g["nodes"] == o["nodes"], t
        assert g["border_points"] == o["border"], t
        assert g["cross_units"] == o["units_cross"], t
        assert g["candidates"] == o["candidates"], t
        assert g["pairs"] >= o["pairs"]  # the device list may hold an ulp-sliver duplicate
        assert g["overflow_points"] == o["overflow"], t
        assert g["active_meetings"] == o["active"], t
        assert g["expired"] == o["expired"], t
        assert g["kv_count"] == o["kv_count"], t
        assert g["ambiguous_nodes"] == 0, t
        totals["ambiguous"] += g["ambiguous_pairs"]
        if t % check_every == 0 or t == steps - 1:
            gp, op = gpu.pairs(), orc.pairs()
            assert np.array_equal(gp, op), (t, len(gp), len(op))
            gb, gc = gpu.leaves()
            ob, oc = orc.leaves()
            assert np.array_equal(gb, ob) and np.array_equal(gc, oc), t
            gk, gs, ge, gm = gpu.active()
            ok, os_, oe, om = orc.active()
            assert np.array_equal(gk, ok), t
            assert np.array_equal(gs, os_) and np.array_equal(ge, oe), t
            assert np.array_equal(gm.view(np.uint32), om.view(np.uint32)), t
        if dyn:
            assert gpu_schema(gpu) == oracle_schema(orc), t
        totals["leaves"].add(o["leaves"])
        totals["pairs"] += o["pairs"]
    gk, gb = sorted_records(*gpu.records())
    ok, ob = sorted_records(*orc.records())
    assert np.array_equal(gk, ok)
    assert np.array_equal(gb.view(np.uint32), ob.view(np.uint32))
    totals["records"] = len(ok)
    print(f"[parity {tag}] ok: {totals}", flush=True)
    gpu.close()
    orc.close()
    return totals


def test_c1_full_run_parity():
    w = workload("c1")
    tot = run_parity(w, w["steps"], tag="c1")
    assert tot["records"] > 1000 and tot["pairs"] > 10000


def test_c2_full_size_parity():
    # the 1M-object hotspot workload at its full size: ~11M pairs per step, every step bit-exact
    # (the oracle takes ~3 s per step); twelve steps let meetings of 10 s expire into records
    w = workload("c2")
    tot = run_parity(w, 12, check_every=4, tag="c2")
    assert tot["pairs"] > 100_000_000 and tot["records"] > 0


def test_deep_tree_parity():
    # a small leaf capacity forces deep trees, finest-level splits and many border lookups
    w = workload("c1", grid_capacity=6, num_objects=6000)
    tot = run_parity(w, 40, tag="deep")
    assert tot["pairs"] > 0


def test_dense_hotspot_parity():
    w = workload("c1", num_objects=30000, grid_capacity=24, reach_distance=15.0, min_meet_time=3)
    w["gen"].update(n_hotspots=6, hotspot_frac=0.8, hotspot_sigma=250.0)
    tot = run_parity(w, 30, check_every=5, tag="hotspot")
    assert tot["records"] > 0


def test_device_trace_matches_host_trace():
    w = workload("c1", num_objects=5000)
    dt = DeviceTrace(w["gen"])
    ht = HostTrace(w["gen"])
    import ctypes
    for t, xy in ht.steps(12):
        p = dt.next(t)
        buf = np.empty_like(xy)
        import stcdb_amd._lib as L  # noqa: F401
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert hip.hipMemcpy(buf.ctypes.data, p, buf.nbytes, 2) == 0
        assert np.array_equal(buf.view(np.uint64), xy.view(np.uint64)), t
    dt.close()


def test_step_device_equals_step_host():
    w = workload("c1", num_objects=4000, keep_pairs=1)
    a = Pipeline(w["cfg"], w["box"])
    b = Pipeline(w["cfg"], w["box"])
    dt = DeviceTrace(w["gen"])
    ht = HostTrace(w["gen"])
    for t, xy in ht.steps(15):
        ra = a.step(xy, t)
        rb = b.step_device(dt.next(t), t)
        for k in ["pairs", "active_meetings", "kv_count", "leaves"]:
            assert ra[k] == rb[k], (t, k)
    assert np.array_equal(a.pairs(), b.pairs())
    ka, _ = sorted_records(*a.records())
    kb, _ = sorted_records(*b.records())
    assert np.array_equal(ka, kb)
    a.close(), b.close(), dt.close()
