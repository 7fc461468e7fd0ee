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
