"""Per-second meeting persistence in the oracle: the pairs a step finds,
saved as the reference's meeting units (key = Cantor(min pid, max pid),
geometry.cu:453-466) and replayed in place of the positions
(load_meetings_pers, geometry.cu:1609-1622), give the same meetings, records
and runs as the positions did, step for step; the units pass through files in
the layout of workbench::dump_meetings / load_meetings
(src/tracing/workbench.cpp:446-470)."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import workload
from stcdb_amd.engine import MEETING_UNIT_DTYPE, HostTrace
from stcdb_amd.traces import meetings_path, read_meetings, write_meetings


def cantor(a, b):
    return (a + b) * (a + b + 1) // 2 + b


def same_state(a, b, tag):
    assert a.counts()["kv_count"] == b.counts()["kv_count"], tag
    ka, sa, ea, ma = a.active()
    kb, sb, eb, mb = b.active()
    assert np.array_equal(ka, kb) and np.array_equal(sa, sb) and np.array_equal(ea, eb), tag
    assert np.array_equal(ma.view(np.uint32), mb.view(np.uint32)), tag
    ra, rb = a.records(), b.records()
    ia, ib = np.lexsort(ra[0].T[::-1]), np.lexsort(rb[0].T[::-1])
    assert np.array_equal(ra[0][ia], rb[0][ib]) and np.array_equal(ra[1][ia], rb[1][ib]), tag
    assert a.num_runs() == b.num_runs(), tag
    for r in range(a.num_runs()):
        x, y = a.run(r), b.run(r)
        assert x["info"] == y["info"], (tag, r)
        for f in ["key_bytes", "bitmaps", "sids", "o_keys", "o_bitmap"]:
            assert np.array_equal(x[f], y[f]), (tag, r, f)


def test_units_are_the_steps_pairs():
    w = workload("c1")
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(12):
        orc.step(xy, t)
        u = orc.step_meeting_units()
        keys = orc.pairs()
        assert u.dtype == MEETING_UNIT_DTYPE and len(u) == len(keys) > 0
        a, b = (keys >> np.uint64(32)).astype(object), (keys & np.uint64(0xFFFFFFFF)).astype(object)
        assert np.all(a < b)
        assert [int(k) for k in u["key"]] == [cantor(x, y) for x, y in zip(a, b)]
        assert np.all(u["start"] == t) and np.all(u["end"] == t)
        # the MBR is the float bbox of the two positions
        ai, bi = a.astype(np.int64), b.astype(np.int64)
        pa, pb = xy[ai].astype(np.float32), xy[bi].astype(np.float32)
        want = np.stack([np.minimum(pa[:, 0], pb[:, 0]), np.minimum(pa[:, 1], pb[:, 1]),
                         np.maximum(pa[:, 0], pb[:, 0]), np.maximum(pa[:, 1], pb[:, 1])], axis=1)
        assert np.array_equal(u["mbr"], want)
    orc.close()


@pytest.mark.parametrize("shuffle", [False, True])
def test_replay_through_files_equals_the_position_run(tmp_path, shuffle):
    w = workload("c1", kv_restriction=400, ctf_count=16)
    n = w["cfg"]["num_objects"]
    pos = O.Oracle(w["cfg"], w["box"])
    steps, per_step = [], []
    for t, xy in HostTrace(w["gen"]).steps(130):
        per_step.append(pos.step(xy, t))
        steps.append(pos.step_meeting_units())
    prefix = str(tmp_path) + "/"
    for st in range(0, 130, 100):
        write_meetings(prefix, st, n, steps[st:st + 100])
    raw = np.fromfile(meetings_path(prefix, 0, n), dtype=np.uint8)
    assert raw.size == 400 + 32 * sum(len(u) for u in steps[:100])
    rep = O.Oracle(w["cfg"], w["box"])
    rng = np.random.default_rng(9)
    for st in range(0, 130, 100):
        blocks = read_meetings(prefix, st, n)
        assert sum(len(b) for b in blocks) == sum(len(u) for u in steps[st:st + 100])
        for t in range(st, min(st + 100, 130)):
            u = blocks[t % 100]
            assert np.array_equal(u.view(np.uint8), steps[t].view(np.uint8))
            if shuffle:  # the reference's units arrive in the order its threads found them
                u = u[rng.permutation(len(u))]
            c = rep.step_meetings(u, t)
            want = per_step[t]
            assert c["pairs"] == want["pairs"] == len(u), t
            assert c["kv_count"] == want["kv_count"] and c["flushed"] == want["flushed"], t
    same_state(pos, rep, "end")
    assert rep.num_runs() >= 3
    pos.close()
    rep.close()


def test_replay_rejects_bad_units():
    w = workload("c1")
    n = w["cfg"]["num_objects"]
    orc = O.Oracle(w["cfg"], w["box"])
    good = np.zeros(2, dtype=MEETING_UNIT_DTYPE)
    good["key"] = [cantor(3, 7), cantor(0, n - 1)]
    orc.step_meetings(good, 0)
    for a, b in [(7, 3), (5, 5), (1, n)]:
        bad = np.zeros(1, dtype=MEETING_UNIT_DTYPE)
        bad["key"] = cantor(a, b)
        with pytest.raises(RuntimeError):
            orc.step_meetings(bad, 1)
    dup = np.zeros(2, dtype=MEETING_UNIT_DTYPE)
    dup["key"] = cantor(3, 7)
    with pytest.raises(RuntimeError):
        orc.step_meetings(dup, 1)
    orc.close()
