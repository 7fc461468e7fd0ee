"""Pins the oracle's flush (runs of CTFs) with properties that do not depend on
re-typing its arithmetic: sortedness, capacities, round trips of the packed
keys through a restatement of the reference's own decoder
(CTF::parse_key / new_parse_mbr, src/tracing/step_merge.cpp:335-379) and
bitmap coverage."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import workload
from stcdb_amd.engine import HostTrace

OID_BIT = 26


def key128(k):
    return (int(k[1]) << 64) | int(k[0])


def run_oracle(w, steps):
    orc = O.Oracle(w["cfg"], w["box"])
    tr = HostTrace(w["gen"])
    for t, xy in tr.steps(steps):
        orc.step(xy, t)
    return orc


def decode_ctf_keys(ctf, raw):
    """CTF::parse_key + new_parse_mbr over the packed little-endian keys"""
    kb = int(ctf["key_bit"]) // 8
    cap = int(ctf["kv_capacity"])
    out = []
    mbr_bit, end_bit, dur_bit, id_bit = (int(ctf[f]) for f in ("mbr_bit", "end_bit", "duration_bit", "id_bit"))
    e, ly, lx = int(ctf["edge_bit"]), int(ctf["low_y_bit"]), int(ctf["low_x_bit"])
    m = [float(v) for v in ctf["mbr"]]
    longer = np.float32(max(np.float32(m[2]) - np.float32(m[0]), np.float32(m[3]) - np.float32(m[1])))
    for i in range(cap):
        k = int.from_bytes(bytes(raw[i * kb:(i + 1) * kb]), "little")
        value_mbr = k & ((1 << mbr_bit) - 1)
        k >>= mbr_bit
        end = k & ((1 << end_bit) - 1)
        k >>= end_bit
        dur = k & ((1 << dur_bit) - 1)
        k >>= dur_bit
        target = k & ((1 << id_bit) - 1)
        k >>= id_bit
        pid = k
        y = value_mbr & ((1 << e) - 1)
        value_mbr >>= e
        x = value_mbr & ((1 << e) - 1)
        value_mbr >>= e
        low1 = value_mbr & ((1 << ly) - 1)
        value_mbr >>= ly
        low0 = value_mbr & ((1 << lx) - 1)
        b0 = m[0] + low0 / ((1 << lx) - 1) * (m[2] - m[0])
        b1 = m[1] + low1 / ((1 << ly) - 1) * (m[3] - m[1])
        out.append((pid, target, dur, end, b0, b1, b0 + x / ((1 << e) - 1) * float(longer),
                    b1 + y / ((1 << e) - 1) * float(longer)))
    return out


@pytest.mark.parametrize("compat", [0, 1])
def test_oracle_runs_are_consistent(compat):
    w = workload("c1", kv_restriction=400, compat=compat, ctf_count=16)
    orc = run_oracle(w, 60)
    assert orc.num_runs() >= 2
    for r in range(orc.num_runs()):
        run = orc.run(r)
        info, ctfs = run["info"], run["ctfs"]
        sk = [key128(k) for k in run["sorted_keys"]]
        assert all(a < b for a, b in zip(sk, sk[1:])), "records must be strictly sorted (unique keys)"
        assert int(ctfs["kv_capacity"].sum()) == info["kv_count"]
        assert info["oversize_count"] == 0
        # sid of every record equals the CTF it sits in, and matches the object's sid
        base = 0
        for c, ctf in enumerate(ctfs):
            for k in sk[base:base + int(ctf["kv_capacity"])]:
                assert k >> 116 == c + 2
                assert run["sids"][(k >> 90) & (2**26 - 1)] == c + 2
            base += int(ctf["kv_capacity"])
        # packed keys decode back to the record fields
        off = 0
        base = 0
        for c, ctf in enumerate(ctfs):
            cap, kb = int(ctf["kv_capacity"]), int(ctf["key_bit"]) // 8
            if cap == 0:
                continue
            assert ctf["flags"] == 0
            assert int(ctf["id_bit"]) == O.load().orc_min_bits(w["cfg"]["num_objects"])
            assert int(ctf["duration_bit"]) == 12 and int(ctf["end_bit"]) == 12
            assert int(ctf["key_bit"]) % 8 == 0 and int(ctf["key_bit"]) >= 2 * int(ctf["id_bit"]) + 24 + int(ctf["mbr_bit"])
            dec = decode_ctf_keys(ctf, run["key_bytes"][off:off + cap * kb])
            for j, d in enumerate(dec):
                k = sk[base + j]
                pid, tgt, dur, end = (k >> 90) & (2**26 - 1), (k >> 64) & (2**26 - 1), (k >> 32) & (2**32 - 1), k & (2**32 - 1)
                # small ids: no sid bits spill into the stored pid field
                assert d[0] == pid and d[1] == tgt and d[2] == dur and d[3] == end
                box = run["sorted_boxes"][base + j].astype(np.float64)
                if not compat:
                    # the quantized MBR corner lies within one cell below the true corner
                    cell_x = (float(ctf["mbr"][2]) - float(ctf["mbr"][0])) / ((1 << int(ctf["low_x_bit"])) - 1)
                    assert d[4] <= box[0] + 1e-6 and box[0] - d[4] <= cell_x * 1.01 + 1e-6
            # the bitmap covers every record's box
            bm = run["bitmaps"][c * 8192:(c + 1) * 8192]
            if not compat:
                xg, yg = int(ctf["x_grid"]), int(ctf["y_grid"])
                m = ctf["mbr"].astype(np.float32)
                for j in range(cap):
                    b = run["sorted_boxes"][base + j].astype(np.float32)
                    i0 = int(np.float32((b[0] - m[0]) / (m[2] - m[0])) * np.float32(xg))
                    j0 = int(np.float32((b[1] - m[1]) / (m[3] - m[1])) * np.float32(yg))
                    bit = i0 + j0 * xg
                    assert bm[bit // 8] & (1 << (bit % 8)), (c, j)
            off += cap * kb
            base += cap
        assert off == info["key_bytes"]


def test_oracle_compat_reduction_drops_lanes():
    """compat bit 0: the CTF MBR covers only records whose in-CTF index mod 64 < 32"""
    w = workload("c1", kv_restriction=1500, ctf_count=4)
    a = run_oracle(w, 90)
    wc = workload("c1", kv_restriction=1500, ctf_count=4, compat=1)
    b = run_oracle(wc, 90)
    ra, rb = a.run(0), b.run(0)
    assert np.array_equal(ra["sorted_keys"], rb["sorted_keys"])
    base = 0
    for c in range(4):
        cap = int(ra["ctfs"][c]["kv_capacity"])
        boxes = ra["sorted_boxes"][base:base + cap]
        full = [boxes[:, 0].min(), boxes[:, 1].min(), boxes[:, 2].max(), boxes[:, 3].max()]
        sel = boxes[np.arange(cap) % 64 < 32]
        half = [sel[:, 0].min(), sel[:, 1].min(), sel[:, 2].max(), sel[:, 3].max()]
        assert np.array_equal(ra["ctfs"][c]["mbr"], np.array(full, dtype=np.float32))
        assert np.array_equal(rb["ctfs"][c]["mbr"], np.array(half, dtype=np.float32))
        base += cap


def test_oracle_oversize_objects_form_sorted_tail():
    # a tiny area threshold makes every object with a spread MBR oversize
    w = workload("c1", kv_restriction=400, ctf_count=16, area_threshold=1e-7)
    orc = run_oracle(w, 60)
    run = orc.run(0)
    info = run["info"]
    assert info["oversize_count"] > 0
    ok = [key128(k) for k in run["o_keys"]]
    assert all(a < b for a, b in zip(ok, ok[1:]))
    for k in ok:
        assert k >> 116 == 0
        assert run["sids"][(k >> 90) & (2**26 - 1)] == 1
    assert info["o_start_min"] <= info["o_start_max"]
