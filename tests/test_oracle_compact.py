"""Pins the oracle's compaction (orc_compact_runs) by what a merge of runs must
keep, checked on the runs it produces:

* every record of the inputs is in the merged run once, with the same ids and
  start/end times, and nothing else is;
* an oversize record of the merged run holds, bit for bit, the float box the
  reference's compaction reads for it: the stored f_box of an input oversize
  record, or CTF::new_parse_mbr_f_box (src/tracing/step_merge.cpp:381-401) of
  an input CTF record, restated here in numpy;
* a CTF record's box is the input box to within the quantisation steps of
  the input and merged CTFs (the packing truncates: step_merge.cpp:325-379);
* an object oversize in any input is oversize in the merged run
  (tools/compaction.cpp:259-268), and every CTF record sits in the CTF its
  object's sid names;
* the run list shrinks by count - 1, the merged run spans the inputs' time
  bases, and object and pair queries answer the same counts before and after.

The runs are flushed with compat 0: with compat 1 a CTF's bounds skip half of
its records (geometry.cu:865-888), so packed boxes can overflow into the time
fields of their keys, and a merge carries such keys on as decoded.
"""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import workload
from stcdb_amd.engine import HostTrace
from test_oracle_query import MASK26

f32 = np.float32


def ctf_records(run):
    """per CTF record: (pid, target, start, end, ctf index, f_box as float32[4])"""
    out = []
    off = 0
    for ci, c in enumerate(run["ctfs"]):
        kb = int(c["key_bit"]) // 8
        mb, eb, db, ib = (int(c[f]) for f in ("mbr_bit", "end_bit", "duration_bit", "id_bit"))
        e, ly, lx = int(c["edge_bit"]), int(c["low_y_bit"]), int(c["low_x_bit"])
        m = [f32(v) for v in c["mbr"]]
        w0, w1 = f32(m[2] - m[0]), f32(m[3] - m[1])
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
            pid = (k >> ib) & ((1 << ib) - 1)
            y = vm & ((1 << e) - 1)
            x = (vm >> e) & ((1 << e) - 1)
            low1 = (vm >> (2 * e)) & ((1 << ly) - 1)
            low0 = (vm >> (2 * e + ly)) & ((1 << lx) - 1)
            b0 = f32(float(m[0]) + low0 / ((1 << lx) - 1) * float(w0))
            b1 = f32(float(m[1]) + low1 / ((1 << ly) - 1) * float(w1))
            b2 = f32(float(b0) + x / ((1 << e) - 1) * longer)
            b3 = f32(float(b1) + y / ((1 << e) - 1) * longer)
            out.append((pid, target, end - dur, end, ci, np.array([b0, b1, b2, b3], dtype=f32)))
        off += int(c["kv_capacity"]) * kb
    return out


def tail_records(run):
    base = run["info"]["end_time_min"]
    out = []
    for (lo, hi), b in zip(run["o_keys"], run["o_boxes"]):
        lo, hi = int(lo), int(hi)
        end = (lo & 0xFFFFFFFF) + base
        out.append(((hi >> 26) & MASK26, hi & MASK26, end - (lo >> 32), end, -1, np.asarray(b, dtype=f32)))
    return out


def step_of(ctfs):
    """largest decode error of a CTF box corner: one low step plus one edge step"""
    s = 0.0
    for c in ctfs:
        w0, w1 = float(c["mbr"][2] - c["mbr"][0]), float(c["mbr"][3] - c["mbr"][1])
        lo = max(w0 / ((1 << int(c["low_x_bit"])) - 1), w1 / ((1 << int(c["low_y_bit"])) - 1))
        s = max(s, lo + max(w0, w1) / ((1 << int(c["edge_bit"])) - 1))
    return s


def build(area_threshold=None, steps=90, ctf_count=16):
    over = dict(kv_restriction=400, ctf_count=ctf_count, compat=0)
    if area_threshold is not None:
        over["area_threshold"] = area_threshold
    w = workload("c1", **over)
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(steps):
        orc.step(xy, t)
    return w, orc


def id_queries(recs, rng, nq):
    """object and pair queries (they read no boxes) on ids drawn from recs"""
    q = np.zeros(nq, dtype=O.QUERY_DTYPE)
    for i in range(nq):
        pid, target, start, end = recs[int(rng.integers(0, len(recs)))][:4]
        t0 = max(0, start - int(rng.integers(0, 10)))
        q[i]["kind"], q[i]["oid"], q[i]["oid2"] = 1 + 2 * (i % 2), pid, target
        q[i]["t0"], q[i]["t1"] = t0, t0 + int(rng.integers(0, 20))
    return q


def check_merge(orc, w, first, count, rng):
    runs = [orc.run(r) for r in range(orc.num_runs())]
    inputs = runs[first:first + count]
    q = id_queries([r for R in inputs for r in ctf_records(R) + tail_records(R)], rng, 200)
    _, before = orc.query(q)
    orc.compact_runs(first, count)
    assert orc.num_runs() == len(runs) - count + 1
    for r in range(first):
        assert np.array_equal(orc.run(r)["key_bytes"], runs[r]["key_bytes"])
    for r in range(first + count, len(runs)):
        assert np.array_equal(orc.run(r - count + 1)["key_bytes"], runs[r]["key_bytes"])
    M = orc.run(first)
    info = M["info"]
    assert info["end_time_min"] == min(R["info"]["end_time_min"] for R in inputs)
    assert info["cur_time"] == max(R["info"]["cur_time"] for R in inputs)
    total = sum(R["info"]["kv_count"] + R["info"]["oversize_count"] for R in inputs)
    assert info["kv_count"] + info["oversize_count"] == total

    src = {}
    for R in inputs:
        for rec in ctf_records(R) + tail_records(R):
            src.setdefault(rec[:4], []).append(rec)
    got_ctf, got_tail = ctf_records(M), tail_records(M)
    got = sorted(r[:4] for r in got_ctf + got_tail)
    assert got == sorted(k for k, v in src.items() for _ in v)

    # oversize records: bit-exact float boxes as the reference's compaction reads them
    from_ctf = 0
    for rec in got_tail:
        cands = src[rec[:4]]
        hit = [c for c in cands if np.array_equal(c[5].view(np.uint32), rec[5].view(np.uint32))]
        assert hit, rec[:4]
        from_ctf += hit[0][4] >= 0
    # CTF records: within the quantisation of the input and merged CTFs
    tol = step_of(M["ctfs"]) + max(step_of(R["ctfs"]) for R in inputs) + 1e-5
    for rec in got_ctf:
        d = min(np.abs(c[5].astype(np.float64) - rec[5].astype(np.float64)).max() for c in src[rec[:4]])
        assert d <= tol, (rec[:4], d, tol)

    sids = M["sids"]
    was_over = np.zeros(len(sids), dtype=bool)
    for R in inputs:
        was_over |= R["sids"] == 1
    has_rec = np.zeros(len(sids), dtype=bool)
    for rec in got_ctf + got_tail:
        has_rec[rec[0]] = True
    assert np.all(sids[was_over & has_rec] == 1)
    for rec in got_ctf:
        assert sids[rec[0]] == rec[4] + 2
    for rec in got_tail:
        assert sids[rec[0]] == 1

    _, after = orc.query(q)
    assert np.array_equal(before, after) and before.sum() > 0
    return len(got_ctf), len(got_tail), from_ctf


@pytest.mark.parametrize("ctf_count", [16, 36])
def test_compaction_keeps_every_record(ctf_count):
    w, orc = build(ctf_count=ctf_count)
    assert orc.num_runs() >= 4
    rng = np.random.default_rng(ctf_count)
    nc, nt, _ = check_merge(orc, w, 1, 2, rng)
    assert nc > 500
    # merge the merged run with everything after it
    check_merge(orc, w, 1, orc.num_runs() - 1, rng)
    assert orc.num_runs() == 2
    orc.close()


def test_compaction_oversize_objects():
    w, orc = build(area_threshold=1e-7, steps=60)
    rng = np.random.default_rng(8)
    n = orc.num_runs()
    assert n >= 3
    nc, nt, from_ctf = check_merge(orc, w, 0, n, rng)
    # objects whose merged box crosses the threshold bring CTF records to the tail
    assert nt > 100 and from_ctf > 0 and nc > 0
    orc.close()


def test_compaction_arguments():
    w, orc = build(steps=60)
    n = orc.num_runs()
    assert n >= 2
    before = [orc.run(r)["key_bytes"].copy() for r in range(n)]
    orc.compact_runs(0, 1)  # one run: nothing to merge
    assert orc.num_runs() == n and np.array_equal(orc.run(0)["key_bytes"], before[0])
    for first, count in [(0, 0), (n - 1, 2), (n, 1)]:
        with pytest.raises(RuntimeError):
            orc.compact_runs(first, count)
    assert orc.num_runs() == n
    orc.close()
