"""GPU parity of the query path (stc_query_batch) against the oracle: object
history, square range and pair queries over live meetings, the memtable and
runs kept in HBM, compared hit for hit (ids, times and the double MBR bits)."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import Pipeline, workload
from stcdb_amd.engine import HIT_DTYPE, HostTrace
from test_oracle_query import make_queries

pytestmark = pytest.mark.gpu


def dense(**over):
    w = workload("c1", num_objects=3000, kv_restriction=300, min_meet_time=3, **over)
    w["gen"].update(n_hotspots=3, hotspot_frac=0.7, hotspot_sigma=150.0)
    w["cfg"]["device_runs"] = 1
    return w


def compare(gpu, orc, q, tag):
    gh, gc = gpu.query(q)
    oh, oc = orc.query(q)
    print(f"[query {tag}] {len(q)} queries, {len(oh)} hits", flush=True)
    assert gh.dtype == HIT_DTYPE and len(gh) == len(oh), (tag, len(gh), len(oh))
    assert np.array_equal(gc, oc), tag
    assert np.array_equal(gh.view(np.uint8), oh.view(np.uint8)), tag
    return len(oh)


@pytest.mark.parametrize("compat", [0, 1])
def test_query_parity_over_live_memtable_and_runs(compat):
    w = dense(compat=compat)
    gpu = Pipeline(w["cfg"], w["box"])
    orc = O.Oracle(w["cfg"], w["box"])
    rng = np.random.default_rng(5 + compat)
    checked = 0
    for t, xy in HostTrace(w["gen"]).steps(48):
        g = gpu.step(xy, t)
        o = orc.step(xy, t)
        assert g["kv_count"] == o["kv_count"] and g["flushed"] == o["flushed"], t
        if t in (12, 30, 47) or (t > 35 and g["flushed"]):
            checked += compare(gpu, orc, make_queries(orc, w, rng, 150), f"t={t} flushed={g['flushed']}")
    assert gpu.num_runs() >= 2 and checked > 500
    gpu.close()
    orc.close()


def test_query_counts_survive_a_small_cap():
    w = dense()
    gpu = Pipeline(w["cfg"], w["box"])
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(30):
        gpu.step(xy, t)
        orc.step(xy, t)
    q = make_queries(orc, w, np.random.default_rng(3), 60)
    _, oc = orc.query(q)
    assert oc.sum() > 20
    hits, gc = gpu.query(q, cap=7)
    assert len(hits) == 7 and np.array_equal(gc, oc)
    gpu.close()
    orc.close()


def test_count_only_queries_match_oracle():
    # counts without hit lists take the planner's shortcut for CTFs wholly inside a range query
    w = dense()
    gpu = Pipeline(w["cfg"], w["box"])
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(44):
        gpu.step(xy, t)
        orc.step(xy, t)
    rng = np.random.default_rng(8)
    q = make_queries(orc, w, rng, 120)
    (lo0, lo1), (hi0, hi1) = w["box"]
    big = q[:30].copy()
    big["kind"] = 2
    big["t0"], big["t1"] = 0, 1000
    for i in range(30):  # nested boxes from a quarter of the map to all of it
        f = 0.25 + 0.75 * i / 29
        big[i]["box"] = (lo0, lo1, lo0 + f * (hi0 - lo0), lo1 + f * (hi1 - lo1))
    q = np.concatenate([q, big])
    _, oc = orc.query(q)
    hits, gc = gpu.query(q, cap=0)
    print(f"[query counts] {len(q)} queries, {int(oc.sum())} hits", flush=True)
    assert len(hits) == 0 and oc[-1] > 100
    assert np.array_equal(gc, oc)
    gpu.close()
    orc.close()


def test_count_flag_mixed_batch_matches_oracle():
    """a batch whose range queries are flagged STC_QF_COUNT (as the serving
    loop sends them) and whose object and pair queries store every hit"""
    w = dense()
    gpu = Pipeline(w["cfg"], w["box"])
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(44):
        gpu.step(xy, t)
        orc.step(xy, t)
    q = make_queries(orc, w, np.random.default_rng(9), 150)
    (lo0, lo1), (hi0, hi1) = w["box"]
    big = q[:10].copy()
    big["kind"], big["t0"], big["t1"] = 2, 0, 1000
    for i in range(10):
        big[i]["box"] = (lo0, lo1, lo0 + (0.3 + 0.07 * i) * (hi0 - lo0), lo1 + (0.3 + 0.07 * i) * (hi1 - lo1))
    q = np.concatenate([q, big])
    q["flags"] = np.where(q["kind"] == 2, 1, 0)
    n = compare(gpu, orc, q, "count flag")
    assert n > 0
    gpu.close()
    orc.close()
