"""GPU parity of per-second meeting persistence: the units a GPU step exports
(stc_get_step_meetings) equal the oracle's byte for byte, and a GPU context
replaying units through files (stc_step_meetings) reaches the oracle's
position-driven state step for step: counts, live meetings, memtable records
and runs."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import HostTrace, Pipeline, workload
from stcdb_amd._lib import StcError
from stcdb_amd.engine import MEETING_UNIT_DTYPE
from stcdb_amd.traces import read_meetings, write_meetings
from test_gpu_compact import same_runs

pytestmark = pytest.mark.gpu


def cantor(a, b):
    return (a + b) * (a + b + 1) // 2 + b


def test_step_units_match_oracle():
    w = workload("c1")
    cfg = dict(w["cfg"])
    cfg["keep_pairs"] = 1
    gpu = Pipeline(cfg, w["box"])
    orc = O.Oracle(cfg, w["box"])
    print(f"[meetings export] n={cfg['num_objects']}", flush=True)
    for t, xy in HostTrace(w["gen"]).steps(15):
        gpu.step(xy, t)
        orc.step(xy, t)
        g, o = gpu.step_meeting_units(), orc.step_meeting_units()
        assert g.dtype == MEETING_UNIT_DTYPE and len(g) == len(o) > 0, t
        assert np.array_equal(g.view(np.uint8), o.view(np.uint8)), t
    gpu.close()
    orc.close()


@pytest.mark.parametrize("compat", [0, 1])
def test_replay_matches_oracle_position_run(tmp_path, compat):
    w = workload("c1", kv_restriction=400, ctf_count=16, compat=compat)
    cfg = dict(w["cfg"])
    cfg["keep_pairs"] = 3
    n = cfg["num_objects"]
    orc = O.Oracle(cfg, w["box"])
    steps, per_step = [], []
    for t, xy in HostTrace(w["gen"]).steps(130):
        per_step.append(orc.step(xy, t))
        steps.append(orc.step_meeting_units())
    prefix = str(tmp_path) + "/"
    for st in range(0, 130, 100):
        write_meetings(prefix, st, n, steps[st:st + 100])
    print(f"[meetings replay compat={compat}] {sum(len(u) for u in steps)} units over 130 steps", flush=True)
    gpu = Pipeline(cfg, w["box"])
    rng = np.random.default_rng(17 + compat)
    for st in range(0, 130, 100):
        blocks = read_meetings(prefix, st, n)
        for t in range(st, min(st + 100, 130)):
            u = blocks[t % 100]
            u = u[rng.permutation(len(u))]  # arrival order does not matter
            g = gpu.step_meetings(u, t)
            want = per_step[t]
            assert g["pairs"] == want["pairs"], t
            assert g["kv_count"] == want["kv_count"] and g["flushed"] == want["flushed"], t
            assert g["active_meetings"] == want["active"], t
            # the replayed step's pairs export again as the same units
            assert np.array_equal(gpu.step_meeting_units().view(np.uint8), steps[t].view(np.uint8)), t
    ka, sa, ea, ma = orc.active()
    kg, sg, eg, mg = gpu.active()
    assert np.array_equal(ka, kg) and np.array_equal(sa, sg) and np.array_equal(ea, eg)
    assert np.array_equal(ma.view(np.uint32), mg.view(np.uint32))
    same_runs(gpu, orc, f"replay compat={compat}")
    assert gpu.num_runs() >= 3
    gpu.close()
    orc.close()


def test_replay_rejects_bad_units():
    w = workload("c1")
    n = w["cfg"]["num_objects"]
    gpu = Pipeline(w["cfg"], w["box"])
    good = np.zeros(2, dtype=MEETING_UNIT_DTYPE)
    good["key"] = [cantor(3, 7), cantor(0, n - 1)]
    gpu.step_meetings(good, 0)
    for a, b in [(7, 3), (5, 5), (1, n)]:
        bad = np.zeros(1, dtype=MEETING_UNIT_DTYPE)
        bad["key"] = cantor(a, b)
        with pytest.raises(StcError):
            gpu.step_meetings(bad, 1)
    dup = np.zeros(2, dtype=MEETING_UNIT_DTYPE)
    dup["key"] = cantor(3, 7)
    with pytest.raises(StcError):
        gpu.step_meetings(dup, 1)
    # the rejected calls changed nothing: step 1 still follows step 0
    g = gpu.step_meetings(good[:1], 1)
    assert g["active_meetings"] == 1
    gpu.close()
