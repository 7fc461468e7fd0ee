"""GPU parity of the dynamic schema (the reference's default mode,
src/util/config.h:35) against the CPU oracle, through the C ABI.

* the initial QTree built on the device (stc_build_schema; the reference's
  partitioner::build_schema, src/tracing/filter.cpp:27-93) equals the
  oracle's, node for node;
* whole runs with the tree kept across steps and updated after each one
  (cuda_update_schema_*, src/cuda/geometry.cu:1114-1234): every step's leaves
  and counts, pairs, active meetings and the updated tree equal the oracle's,
  and the memtable records at the end; on the street-grid workloads and on a
  map at 60 degrees north whose leaves are twice as wide as high, where the
  height rule of the update differs from a width rule;
* a tree installed with stc_set_schema (the oracle's) gives the same run, and
  a malformed tree is rejected.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import stcdb_amd._lib as L
from stcdb_amd import HostTrace, Pipeline, workload
from stcdb_amd.engine import QTNODE_DTYPE
from test_gpu_pipeline import gpu_schema, oracle_schema, run_parity
from test_oracle_schema import MAP, cluster_walk

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,over", [("c1", {}), ("c1", dict(grid_capacity=8, num_objects=8000)),
                                       ("c2", dict(num_objects=200_000, grid_capacity=50))])
def test_initial_schema_equals_oracle(name, over):
    w = workload(name, dynamic_schema=1, **over)
    print(f"[schema build] {name} {over}", flush=True)
    xy = next(iter(HostTrace(w["gen"]).steps(1)))[1]
    gpu = Pipeline(w["cfg"], w["box"])
    orc = O.Oracle(w["cfg"], w["box"])
    gpu.build_schema(xy)
    orc.build_schema(xy)
    g, o = gpu_schema(gpu), oracle_schema(orc)
    assert g == o
    assert sum(v[0] == 2 for v in o.values()) > 40
    gpu.close()
    orc.close()


def test_dynamic_c1_full_run_parity():
    w = workload("c1", dynamic_schema=1)
    tot = run_parity(w, w["steps"], check_every=4, tag="dyn c1")
    assert tot["records"] > 1000


def test_dynamic_hotspot_parity():
    # deep and changing trees: 30k objects around 6 hotspots, leaf capacity 12
    w = workload("c1", num_objects=30000, grid_capacity=12, reach_distance=15.0, min_meet_time=3, dynamic_schema=1)
    w["gen"].update(n_hotspots=6, hotspot_frac=0.8, hotspot_sigma=250.0)
    tot = run_parity(w, 30, check_every=3, tag="dyn hotspot")
    assert tot["records"] > 0 and len(tot["leaves"]) > 5  # the tree changes from step to step


@pytest.mark.parametrize("delay", [1, 2])
def test_dynamic_height_rule_parity(delay):
    # leaves twice as wide as high: the update splits by height (GPU rule)
    n = 6000
    frames = cluster_walk(n, 16, seed=21, step_m=9.0)
    from stcdb_amd.workloads import pipeline_config
    cfg = pipeline_config(n, 20.0, 3, grid_capacity=8, dynamic_schema=1, schema_update_delay=delay)
    w = dict(cfg=cfg, box=MAP, gen=None)
    tot = run_parity(w, len(frames), tag=f"dyn height delay {delay}", frames=enumerate(frames))
    assert len(tot["leaves"]) > 5


def test_set_schema_runs_like_the_oracle():
    w = workload("c1", num_objects=8000, grid_capacity=10, dynamic_schema=1, keep_pairs=1)
    frames = list(HostTrace(w["gen"]).steps(12))
    orc = O.Oracle(w["cfg"], w["box"])
    orc.build_schema(frames[0][1])
    boxes, mids, typ, ch = orc.nodes()
    q = np.zeros(len(typ), dtype=QTNODE_DTYPE)
    q["type"] = typ
    q["mbr"] = boxes
    q["mid_x"], q["mid_y"] = mids[:, 0], mids[:, 1]
    q["children"] = np.where(typ[:, None] == 1, ch, 0)
    gpu = Pipeline(w["cfg"], w["box"])
    gpu.set_schema(q)
    assert gpu_schema(gpu) == oracle_schema(orc)
    for t, xy in frames:
        g = gpu.step(xy, t)
        o = orc.step(xy, t)
        assert g["leaves"] == o["leaves"] and g["pairs"] >= o["pairs"], t
        assert np.array_equal(gpu.pairs(), orc.pairs()), t
        assert gpu_schema(gpu) == oracle_schema(orc), t
    # malformed trees are rejected
    bad = q.copy()
    bad["children"][0][1] = bad["children"][0][0]  # a node reached twice
    with pytest.raises(L.StcError):
        gpu.set_schema(bad)
    bad = q.copy()
    bad["type"][0] = 2                               # the root must be a branch
    with pytest.raises(L.StcError):
        gpu.set_schema(bad)
    gpu.close()
    orc.close()


def test_schema_entry_points_need_dynamic_mode():
    w = workload("c1", num_objects=4000)
    gpu = Pipeline(w["cfg"], w["box"])
    n = ctypes.c_uint64()
    assert gpu.lib.stc_get_schema(gpu.ctx, None, 0, ctypes.byref(n)) == L.STC_E_STATE
    gpu.close()
