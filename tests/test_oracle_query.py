"""Pins the oracle's queries over the resident state (live meetings, memtable,
runs) against a plain-Python restatement of the same set semantics over the
oracle's exported state. Run keys are decoded as the reference decodes them
(CTF::ctf_get_key_oid, parse_key, new_parse_mbr: src/tracing/step_merge.cpp:
325-379), with the CTF box extents subtracted in float32 as the reference's
f_box arithmetic does."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import workload
from stcdb_amd.engine import HostTrace

MASK26 = (1 << 26) - 1


def decode_run(run):
    """(pid, target, start, end, box) of every CTF record and oversize record"""
    out = []
    off = 0
    for c in run["ctfs"]:
        kb = int(c["key_bit"]) // 8
        mb, eb, db, ib = (int(c[f]) for f in ("mbr_bit", "end_bit", "duration_bit", "id_bit"))
        e, ly, lx = int(c["edge_bit"]), int(c["low_y_bit"]), int(c["low_x_bit"])
        m = [np.float32(v) for v in c["mbr"]]
        w0, w1 = np.float32(m[2] - m[0]), np.float32(m[3] - m[1])
        longer = float(max(w0, w1))
        for i in range(int(c["kv_capacity"])):
            k = int.from_bytes(bytes(run["key_bytes"][off + i * kb:off + (i + 1) * kb]), "little")
            vm = k & ((1 << mb) - 1)
            k >>= mb
            end = (k & ((1 << eb) - 1)) + int(c["end_time_min"])
            k >>= eb
            dur = k & ((1 << db) - 1)
            k >>= db
            target = k & ((1 << ib) - 1)
            k >>= ib
            pid = k & ((1 << ib) - 1)
            y = vm & ((1 << e) - 1)
            vm >>= e
            x = vm & ((1 << e) - 1)
            vm >>= e
            low1 = vm & ((1 << ly) - 1)
            vm >>= ly
            low0 = vm & ((1 << lx) - 1)
            b0 = float(m[0]) + low0 / ((1 << lx) - 1) * float(w0)
            b1 = float(m[1]) + low1 / ((1 << ly) - 1) * float(w1)
            box = (b0, b1, b0 + x / ((1 << e) - 1) * longer, b1 + y / ((1 << e) - 1) * longer)
            out.append((pid, target, end - dur, end, box))
        off += int(c["kv_capacity"]) * kb
    base = run["info"]["end_time_min"]
    for (lo, hi), b in zip(run["o_keys"], run["o_boxes"]):
        lo, hi = int(lo), int(hi)
        end = (lo & 0xFFFFFFFF) + base
        out.append(((hi >> 26) & MASK26, hi & MASK26, end - (lo >> 32), end, tuple(float(v) for v in b)))
    return out


def brute_force(orc, queries):
    items = []  # (pid, target, start, end, box, source)
    keys, start, end, mbr = orc.active()
    for k, s, e, m in zip(keys, start, end, mbr):
        a, b = int(k) >> 32, int(k) & 0xFFFFFFFF
        box = tuple(float(v) for v in m)
        items += [(a, b, int(s), int(e), box, 0), (b, a, int(s), int(e), box, 0)]
    rk, rb = orc.records()
    base = orc.end_time_min()
    for (lo, hi), b in zip(rk, rb):
        lo, hi = int(lo), int(hi)
        end_t = (lo & 0xFFFFFFFF) + base
        items.append(((hi >> 26) & MASK26, hi & MASK26, end_t - (lo >> 32), end_t, tuple(float(v) for v in b), 1))
    for r in range(orc.num_runs()):
        items += [it + (2 + r,) for it in decode_run(orc.run(r))]
    hits = []
    for qi, q in enumerate(queries):
        for pid, target, s, e, box, src in items:
            if not (s <= q["t1"] and e >= q["t0"]):
                continue
            kind = q["kind"]
            if kind == 1 and pid == q["oid"]:
                pass
            elif kind == 3 and pid == q["oid"] and target == q["oid2"]:
                pass
            elif kind == 2 and pid < target and box[0] <= q["box"][2] and box[2] >= q["box"][0] and \
                    box[1] <= q["box"][3] and box[3] >= q["box"][1]:
                pass
            else:
                continue
            hits.append((qi, pid, target, s, e, src) + box)
    h = np.zeros(len(hits), dtype=O.HIT_DTYPE)
    for i, t in enumerate(hits):
        h[i] = (t[0], t[1], t[2], t[3], t[4], t[5], t[6:10])
    return O.sort_hits(h)


def make_queries(orc, w, rng, nq):
    keys, _ = orc.records()
    pids = [int(hi) >> 26 & MASK26 for _, hi in keys] or [0]
    tgts = [int(hi) & MASK26 for _, hi in keys] or [1]
    (lo0, lo1), (hi0, hi1) = w["box"]
    q = np.zeros(nq, dtype=O.QUERY_DTYPE)
    for i in range(nq):
        kind = [1, 2, 3][i % 3]
        t0 = int(rng.integers(0, 40))
        q[i]["kind"], q[i]["t0"], q[i]["t1"] = kind, t0, t0 + int(rng.integers(0, 30))
        j = int(rng.integers(0, len(pids)))
        q[i]["oid"] = pids[j] if rng.random() < 0.8 else int(rng.integers(0, w["cfg"]["num_objects"]))
        q[i]["oid2"] = tgts[j] if rng.random() < 0.8 else int(rng.integers(0, w["cfg"]["num_objects"]))
        cx, cy = rng.uniform(lo0, hi0), rng.uniform(lo1, hi1)
        side = 10 ** rng.uniform(-3.5, -1.5)
        q[i]["box"] = (cx - side, cy - side, cx + side, cy + side)
    return q


def test_oracle_queries_match_brute_force():
    w = workload("c1", num_objects=3000, kv_restriction=300, min_meet_time=3)
    w["gen"].update(n_hotspots=3, hotspot_frac=0.7, hotspot_sigma=150.0)
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(60):
        c = orc.step(xy, t)
        if t >= 40 and not c["flushed"] and c["kv_count"]:
            break  # end with records in the memtable
    assert orc.num_runs() >= 2 and len(orc.records()[0]) > 0
    rng = np.random.default_rng(11)
    q = make_queries(orc, w, rng, 90)
    got, counts = orc.query(q)
    want = brute_force(orc, q)
    assert len(got) == len(want) and len(got) > 100
    assert np.array_equal(got, want)
    assert counts.sum() == len(got)
    # every source contributes
    assert set(np.unique(got["source"])) >= {0, 1, 2}


def test_oracle_count_flag_stores_no_hits():
    """STC_QF_COUNT: the query's count is unchanged and none of its hits is stored"""
    w = workload("c1", num_objects=3000, kv_restriction=300, min_meet_time=3)
    w["gen"].update(n_hotspots=3, hotspot_frac=0.7, hotspot_sigma=150.0)
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(45):
        orc.step(xy, t)
    q = make_queries(orc, w, np.random.default_rng(12), 90)
    all_hits, all_counts = orc.query(q)
    q["flags"][::2] = 1
    hits, counts = orc.query(q)
    assert np.array_equal(counts, all_counts)
    keep = all_hits[q["flags"][all_hits["query"]] == 0]
    assert np.array_equal(hits, keep) and 0 < len(hits) < len(all_hits)
