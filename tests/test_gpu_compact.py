"""GPU parity of compaction (stc_compact_runs): the merged run equals the
oracle's byte for byte (CTF metadata, packed keys, bitmaps, sids, oversize
buffer, sorted records), the runs around it are untouched, and queries over
the merged state answer as the oracle's."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import HostTrace, Pipeline, workload
from stcdb_amd._lib import StcError
from test_gpu_query import compare
from test_oracle_query import make_queries

pytestmark = pytest.mark.gpu


def same_runs(gpu, orc, tag):
    assert gpu.num_runs() == orc.num_runs(), tag
    for r in range(orc.num_runs()):
        gr, orr = gpu.run(r), orc.run(r)
        assert gr["info"] == orr["info"], (tag, r)
        for f in orr["ctfs"].dtype.names:
            assert np.array_equal(gr["ctfs"][f], orr["ctfs"][f]), (tag, r, f)
        for f in ["key_bytes", "bitmaps", "sids", "o_keys", "o_bitmap", "sorted_keys"]:
            assert np.array_equal(gr[f], orr[f]), (tag, r, f)
        for f in ["o_boxes", "sorted_boxes"]:
            assert np.array_equal(gr[f].view(np.uint32), orr[f].view(np.uint32)), (tag, r, f)


def lockstep(w, steps, tag):
    cfg = dict(w["cfg"])
    cfg["keep_pairs"] = 3
    cfg["device_runs"] = 1
    print(f"[compact {tag}] n={cfg['num_objects']} kv_restriction={cfg['kv_restriction']} ctf={cfg['ctf_count']} "
          f"compat={cfg['compat']} steps={steps}", flush=True)
    gpu = Pipeline(cfg, w["box"])
    orc = O.Oracle(cfg, w["box"])
    for t, xy in HostTrace(w["gen"]).steps(steps):
        g = gpu.step(xy, t)
        o = orc.step(xy, t)
        assert g["flushed"] == o["flushed"] and g["kv_count"] == o["kv_count"], t
    return gpu, orc


@pytest.mark.parametrize("compat", [0, 1])
def test_compaction_parity(compat):
    w = workload("c1", kv_restriction=400, ctf_count=16, compat=compat)
    gpu, orc = lockstep(w, 90, f"compat{compat}")
    assert orc.num_runs() >= 4
    rng = np.random.default_rng(21 + compat)
    for first, count in [(1, 2), (0, 2), (0, 2)]:
        print(f"[compact compat{compat}] merge {first}+{count} of {orc.num_runs()}", flush=True)
        gpu.compact_runs(first, count)
        orc.compact_runs(first, count)
        same_runs(gpu, orc, f"merge {first}+{count}")
        compare(gpu, orc, make_queries(orc, w, rng, 150), f"after merge {first}+{count}")
    gpu.close()
    orc.close()


def test_compaction_parity_oversize():
    w = workload("c1", kv_restriction=400, ctf_count=16, area_threshold=1e-7)
    gpu, orc = lockstep(w, 60, "oversize")
    n = orc.num_runs()
    assert n >= 3
    gpu.compact_runs(0, n)
    orc.compact_runs(0, n)
    same_runs(gpu, orc, "oversize")
    assert orc.run(0)["info"]["oversize_count"] > 100
    compare(gpu, orc, make_queries(orc, w, np.random.default_rng(4), 150), "oversize")
    gpu.close()
    orc.close()


def test_compaction_parity_dense_then_steps_continue():
    w = workload("c1", num_objects=30000, grid_capacity=24, reach_distance=15.0, min_meet_time=3,
                 kv_restriction=5000, ctf_count=100)
    w["gen"].update(n_hotspots=6, hotspot_frac=0.8, hotspot_sigma=250.0)
    gpu, orc = lockstep(w, 20, "dense")
    n = orc.num_runs()
    assert n >= 2
    gpu.compact_runs(0, n)
    orc.compact_runs(0, n)
    same_runs(gpu, orc, "dense")
    # the memtable is untouched: the next steps flush as before
    tr = HostTrace(w["gen"])
    for t, xy in tr.steps(26):
        if t < 20:
            continue
        g = gpu.step(xy, t)
        o = orc.step(xy, t)
        assert g["flushed"] == o["flushed"] and g["kv_count"] == o["kv_count"], t
    same_runs(gpu, orc, "dense after steps")
    gpu.close()
    orc.close()


def test_compaction_needs_runs_in_hbm():
    w = workload("c1", kv_restriction=400, ctf_count=16)
    gpu = Pipeline(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(60):
        gpu.step(xy, t)
    n = gpu.num_runs()
    assert n >= 2
    with pytest.raises(StcError):
        gpu.compact_runs(0, 2)
    with pytest.raises(StcError):
        gpu.compact_runs(n - 1, 2)
    assert gpu.num_runs() == n
    gpu.close()
