"""Runs on disk in the reference's layout (stcdb_amd/runfiles.py): an oracle
run written and read back equals itself field by field, and the files have
the sizes the reference's loaders read (load_CTB_meta / load_CTF_meta,
src/tracing/workbench.cpp:653-694)."""
import os

import numpy as np

import oracle_lib as O
from stcdb_amd import runfiles, workload
from stcdb_amd.engine import HostTrace


def dense_oracle(steps=40, **over):
    w = workload("c1", num_objects=3000, kv_restriction=300, min_meet_time=3, **over)
    w["gen"].update(n_hotspots=3, hotspot_frac=0.7, hotspot_sigma=150.0)
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(steps):
        orc.step(xy, t)
    return w, orc


def check_round_trip(run, back):
    C = len(run["ctfs"])
    assert len(back["ctfs"]) == C
    for k in runfiles.CTF_FIELDS:
        assert np.array_equal(back["ctfs"][k], run["ctfs"][k]), k
    assert np.array_equal(back["ctfs"]["mbr"].view(np.uint32), run["ctfs"]["mbr"].view(np.uint32))
    assert not back["ctfs"]["keys"].any() and not back["ctfs"]["bitmap"].any()  # pointers written as zero
    assert np.array_equal(back["key_bytes"], run["key_bytes"])
    for c in range(C):
        n = int(run["ctfs"][c]["bitmap_size"])
        assert np.array_equal(back["bitmaps"][c * 8192:c * 8192 + n], run["bitmaps"][c * 8192:c * 8192 + n])
    assert np.array_equal(back["sids"], run["sids"])
    assert np.array_equal(back["o_keys"], run["o_keys"])
    assert np.array_equal(back["o_boxes"].view(np.uint32), run["o_boxes"].view(np.uint32))
    assert np.array_equal(back["o_bitmap"], run["o_bitmap"])
    info = run["info"]
    assert back["ctb"]["o_count"] == info["oversize_count"] and back["ctb"]["ctf_count"] == C
    assert back["ctb"]["end_time_min"] == info["end_time_min"] and back["ctb"]["end_time_max"] == info["cur_time"]


def test_oracle_run_round_trips_through_files(tmp_path):
    w, orc = dense_oracle(area_threshold=1e-7)  # small threshold: an oversize buffer too
    assert orc.num_runs() >= 2
    seen_tail = False
    for r in range(orc.num_runs()):
        run = orc.run(r)
        seen_tail |= run["info"]["oversize_count"] > 0
        runfiles.dump_run(run, str(tmp_path), "6oversize_10", r)
        n, C = run["info"]["num_objects"], run["info"]["ctf_count"]
        sst, ctb, meta = runfiles.paths(str(tmp_path), "6oversize_10", r, C)
        assert os.path.getsize(ctb) == 128 + 2 * n + 32 * run["info"]["oversize_count"] + 8192
        for c in range(C):
            ctf = run["ctfs"][c]
            assert os.path.getsize(meta[c]) == 80 + int(ctf["bitmap_size"])
            assert os.path.getsize(sst[c]) == int(ctf["kv_capacity"]) * (int(ctf["key_bit"]) // 8)
        check_round_trip(run, runfiles.load_run(str(tmp_path), "6oversize_10", r, n, C))
    assert seen_tail
    orc.close()
