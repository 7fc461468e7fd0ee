"""Live contact search (stc_live_search, the reference's search_single /
search_multi, src/cuda/geometry.cu:605-662) against the oracle's live
meetings, with the trace fed from .tr chunk files."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import Pipeline, traces, workload
from stcdb_amd.engine import HIT_DTYPE, HostTrace

pytestmark = pytest.mark.gpu


def expected(orc, pids, t, m):
    keys, start, end, mbr = orc.active()
    rows = []
    for qi, p in enumerate(pids):
        for k, s, e, b in zip(keys, start, end, mbr):
            a, c = int(k) >> 32, int(k) & 0xFFFFFFFF
            if e == t and t - s >= m + 1 and p in (a, c):
                rows.append((qi, p, c if p == a else a, int(s), int(e), 0, tuple(float(v) for v in b)))
    h = np.zeros(len(rows), dtype=HIT_DTYPE)
    for i, r in enumerate(rows):
        h[i] = r
    return O.sort_hits(h)


def test_live_search_two_hops(tmp_path):
    w = workload("c1", num_objects=3000, min_meet_time=3)
    w["gen"].update(n_hotspots=3, hotspot_frac=0.7, hotspot_sigma=150.0)
    prefix = str(tmp_path / "c1_")
    traces.write_trace(prefix, HostTrace(w["gen"]), 30)
    gpu = Pipeline(w["cfg"], w["box"])
    orc = O.Oracle(w["cfg"], w["box"])
    m = w["cfg"]["min_meet_time"]
    hop1 = None
    for t, xy in traces.iter_trace(prefix, 3000, 30):
        gpu.step(xy, t)
        orc.step(xy, t)
        if t == 20:
            keys, start, _, _ = orc.active()
            long_lived = [int(k) >> 32 for k, s in zip(keys, start) if t - s >= m + 1]
            pids = np.array(long_lived[:5] + [2999], dtype=np.uint32)
            hop1 = gpu.live_search(pids)
            assert len(hop1) > 0 and np.array_equal(hop1, expected(orc, pids, t, m))
        if t == 21:  # the reference's second hop: the contacts found, one step later
            targets = np.unique(hop1["target"])
            hop2 = gpu.live_search(targets)
            assert np.array_equal(hop2, expected(orc, targets, t, m))
            print(f"[live] hop1 {len(hop1)} contacts, hop2 {len(hop2)}", flush=True)
    gpu.close()
    orc.close()
