"""stc_live_search_noalloc (the allocation-free entry point) against
stc_live_search through the C ABI: the same hits, bit for bit, on two calls
in a row (the second with fewer objects, so it must not see the first call's
leftovers in the shared scratch array), and on a call whose list repeats
objects; the size limit is reported, not overrun."""
import ctypes

import numpy as np
import pytest

import stcdb_amd._lib as L
from stcdb_amd import HostTrace, Pipeline, workload
from stcdb_amd.engine import HIT_DTYPE

pytestmark = pytest.mark.gpu


def noalloc(gpu, pids, cap):
    pids = np.ascontiguousarray(pids, dtype=np.uint32)
    hits = np.zeros(cap, dtype=HIT_DTYPE)
    n = ctypes.c_uint64()
    rc = gpu.lib.stc_live_search_noalloc(gpu.ctx, pids.ctypes.data_as(ctypes.c_void_p), len(pids),
                                         hits.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n))
    return rc, hits[:min(n.value, cap)], n.value


def test_noalloc_live_search_equals_live_search():
    w = workload("c1", num_objects=20000, min_meet_time=3, reach_distance=25.0)
    w["gen"].update(n_hotspots=4, hotspot_frac=0.7, hotspot_sigma=200.0)
    gpu = Pipeline(w["cfg"], w["box"])
    print("[noalloc] 20000 objects, 25 steps", flush=True)
    for t, xy in HostTrace(w["gen"]).steps(25):
        gpu.step(xy, t)
    k, s, e, _ = gpu.active()
    rng = np.random.default_rng(3)
    owners = np.unique((k >> np.uint64(32)).astype(np.uint32))
    big = np.concatenate([rng.choice(owners, 3000, replace=False), rng.integers(0, 20000, 1000)]).astype(np.uint32)
    small = big[:200]
    dup = np.concatenate([small[:50], small[:50], small[10:30]])
    for label, pids in [("big", big), ("small", small), ("repeats", dup)]:
        ref = gpu.live_search(pids)
        rc, got, n = noalloc(gpu, pids, len(ref) + 16)
        print(f"[noalloc] {label}: {len(pids)} objects, {n} hits", flush=True)
        assert rc == 0 and n == len(ref)
        assert np.array_equal(got, ref), label
        assert len(ref) > 0
    # limits: more objects than the scratch holds is an error, nothing is written
    rc, _, _ = noalloc(gpu, np.arange(16385, dtype=np.uint32), 16)
    assert rc == L.STC_E_CAPACITY
    gpu.close()
