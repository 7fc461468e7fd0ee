"""GPU parity of the flush: every run (CTF metadata, packed key bytes, bitmaps,
object sids, oversize buffer, sorted records) equals the oracle's byte for byte."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import HostTrace, Pipeline, workload

pytestmark = pytest.mark.gpu


def compare_runs(w, steps, tag):
    cfg = dict(w["cfg"])
    cfg["keep_pairs"] = 3
    print(f"[flush {tag}] n={cfg['num_objects']} kv_restriction={cfg['kv_restriction']} ctf={cfg['ctf_count']} "
          f"compat={cfg['compat']} steps={steps}", flush=True)
    gpu = Pipeline(cfg, w["box"])
    orc = O.Oracle(cfg, w["box"])
    tr = HostTrace(w["gen"])
    for t, xy in tr.steps(steps):
        g = gpu.step(xy, t)
        o = orc.step(xy, t)
        assert g["flushed"] == o["flushed"], t
        assert g["kv_count"] == o["kv_count"], t
    assert gpu.num_runs() == orc.num_runs() and orc.num_runs() >= 1
    for r in range(orc.num_runs()):
        gr, orr = gpu.run(r), orc.run(r)
        assert gr["info"] == orr["info"], r
        for f in orr["ctfs"].dtype.names:
            assert np.array_equal(gr["ctfs"][f], orr["ctfs"][f]), (r, f)
        for f in ["key_bytes", "bitmaps", "sids", "o_keys", "o_bitmap", "sorted_keys"]:
            assert np.array_equal(gr[f], orr[f]), (r, f)
        for f in ["o_boxes", "sorted_boxes"]:
            assert np.array_equal(gr[f].view(np.uint32), orr[f].view(np.uint32)), (r, f)
    print(f"[flush {tag}] ok: {orc.num_runs()} runs", flush=True)
    return orc.num_runs()


@pytest.mark.parametrize("compat", [0, 1])
def test_flush_parity_c1(compat):
    w = workload("c1", kv_restriction=400, ctf_count=16, compat=compat)
    assert compare_runs(w, 80, f"c1-compat{compat}") >= 3


def test_flush_parity_oversize_tail():
    w = workload("c1", kv_restriction=400, ctf_count=16, area_threshold=1e-7)
    compare_runs(w, 60, "oversize")


def test_flush_parity_many_ctfs_dense():
    w = workload("c1", num_objects=30000, grid_capacity=24, reach_distance=15.0, min_meet_time=3,
                 kv_restriction=5000, ctf_count=100)
    w["gen"].update(n_hotspots=6, hotspot_frac=0.8, hotspot_sigma=250.0)
    compare_runs(w, 25, "dense")


def test_async_flush_runs_equal_oracle():
    """async_flush: each step's flush runs on its own stream and thread and
    overlaps the next step; the runs are still the oracle's, byte for byte"""
    w = workload("c1", num_objects=30000, grid_capacity=24, reach_distance=15.0, min_meet_time=3,
                 kv_restriction=3000, ctf_count=100, dynamic_schema=1, async_flush=1)
    w["gen"].update(n_hotspots=6, hotspot_frac=0.8, hotspot_sigma=250.0)
    assert compare_runs(w, 25, "async") >= 5


def test_async_flush_state_equals_synchronous():
    """with a flush at almost every step, a context with async_flush and one
    without hold the same meetings and memtable after every step, report the
    same counters, and end with the same runs"""
    w = workload("c1", num_objects=20000, min_meet_time=2, kv_restriction=200, ctf_count=16, dynamic_schema=1)
    w["gen"].update(n_hotspots=4, hotspot_frac=0.7, hotspot_sigma=300.0)
    a = Pipeline(dict(w["cfg"], async_flush=1, keep_pairs=1), w["box"])
    b = Pipeline(dict(w["cfg"], keep_pairs=1), w["box"])
    flushed = done = 0
    for t, xy in HostTrace(w["gen"]).steps(30):
        ga, gb = a.step(xy, t), b.step(xy, t)
        for k in ("pairs", "active_meetings", "expired", "kv_count", "flushed", "leaves"):
            assert ga[k] == gb[k], (t, k)
        flushed += gb["flushed"]
        done += ga["ms_flush_done"] > 0
        assert np.array_equal(a.pairs(), b.pairs()), t
        ka, sa, ea, ma = a.active()
        kb, sb, eb, mb = b.active()
        assert np.array_equal(ka, kb) and np.array_equal(sa, sb) and np.array_equal(ea, eb), t
        assert np.array_equal(ma.view(np.uint32), mb.view(np.uint32)), t
        if t % 3 == 2:  # reading the memtable waits for a pending flush; other steps join it in the merge
            ra, rb = a.records(), b.records()
            assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1].view(np.uint32), rb[1].view(np.uint32)), t
    assert flushed >= 20 and done >= flushed - 1
    assert a.num_runs() == b.num_runs() == flushed
    for r in range(b.num_runs()):
        x, y = a.run(r), b.run(r)
        assert x["info"] == y["info"], r
        for f in ["key_bytes", "bitmaps", "sids", "o_keys", "o_bitmap"]:
            assert np.array_equal(x[f], y[f]), (r, f)
    a.close()
    b.close()
