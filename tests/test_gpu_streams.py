"""Batched streams (the c4 family): independent streams, each one context of
the library with its own box, reach and leaf capacity, stepped together
through stc_step_batch (host threads, each context on its own streams, their
kernels overlapping on the GPU). Every stream must match its own oracle run
bit for bit, as a plain sequential run of that stream would: contexts share
no state.

The streams are scaled down from c4's (median 3,000 objects instead of 50,000,
boxes of 2-6 km instead of 10-50 km, about the same density) so that the
oracle checks every stream in seconds."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import HostTrace, Pipeline, batched_streams, step_batch
from stcdb_amd.tuner import tune_leaf_capacity

pytestmark = pytest.mark.gpu


def sorted_records(keys, boxes):
    order = np.lexsort((keys[:, 0], keys[:, 1]))
    return keys[order], boxes[order]


def test_batched_streams_parity():
    streams = batched_streams(12, median=3000, lo=500, hi=12000, side_km=(2.0, 6.0), steps=40, seed=4)
    runs = []
    for i, w in enumerate(streams):
        cfg = dict(w["cfg"], keep_pairs=1, kv_restriction=1 << 20)
        print(f"[streams] stream {i}: n={cfg['num_objects']} cells={w['gen']['grid_cells']} "
              f"r={cfg['reach_distance']} cap={cfg['grid_capacity']}", flush=True)
        runs.append((Pipeline(cfg, w["box"]), O.Oracle(cfg, w["box"]), HostTrace(w["gen"]).steps(w["steps"])))
    pairs = [0] * len(runs)
    records = [0] * len(runs)
    for t in range(streams[0]["steps"]):
        xys = []
        for gpu, orc, trace in runs:
            tt, xy = next(trace)
            assert tt == t
            xys.append(xy.copy())
        outs = step_batch([r[0] for r in runs], xys, t, threads=8)  # every stream's step at once
        for i, (gpu, orc, trace) in enumerate(runs):
            g = outs[i]
            o = orc.step(xys[i], t)
            assert g["leaves"] == o["leaves"], (i, t)
            assert g["cross_units"] == o["units_cross"], (i, t)
            assert g["candidates"] == o["candidates"], (i, t)
            assert g["active_meetings"] == o["active"], (i, t)
            assert g["expired"] == o["expired"], (i, t)
            assert g["kv_count"] == o["kv_count"], (i, t)
            pairs[i] += o["pairs"]
            if t % 5 == 4:
                assert np.array_equal(gpu.pairs(), orc.pairs()), (i, t)
                gk, gs, ge, gm = gpu.active()
                ok, os_, oe, om = orc.active()
                assert np.array_equal(gk, ok) and np.array_equal(gs, os_) and np.array_equal(ge, oe), (i, t)
                assert np.array_equal(gm.view(np.uint32), om.view(np.uint32)), (i, t)
    for i, (gpu, orc, _) in enumerate(runs):
        gk, gb = sorted_records(*gpu.records())
        ok, ob = sorted_records(*orc.records())
        assert np.array_equal(gk, ok), i
        assert np.array_equal(gb.view(np.uint32), ob.view(np.uint32)), i
        records[i] = len(ok)
        gpu.close()
        orc.close()
    print(f"[streams] ok: pairs {pairs} records {records}", flush=True)
    assert sum(p > 0 for p in pairs) >= len(runs) // 2
    assert sum(records) > 0


def run_stream(w, steps, trace=None):
    """one stream alone: its context, its steps; returns per-step pairs and the final state"""
    cfg = dict(w["cfg"], keep_pairs=1)
    gpu = Pipeline(cfg, w["box"])
    out = []
    for t, xy in (trace or HostTrace(w["gen"]).steps(steps)):
        gpu.step(xy, t)
        out.append(gpu.pairs())
    k, s, e, m = gpu.active()
    rk, rb = sorted_records(*gpu.records())
    gpu.close()
    return out, (k, s, e, m.view(np.uint32)), (rk, rb.view(np.uint32))


def test_c4_full_size_streams():
    """c4 at its size: 64 streams of ~6M objects in total (log-normal, median
    50k, 10-50 km boxes, reach 5-25 m), all contexts alive and stepped together
    through stc_step_batch on 16 host threads for a few steps; each stream
    equals its own sequential run alone, and four sampled streams equal their
    oracle runs"""
    steps = 6
    streams = batched_streams()  # the c4 parameters
    total = sum(w["cfg"]["num_objects"] for w in streams)
    print(f"[c4] {len(streams)} streams, {total} objects", flush=True)
    assert len(streams) == 64 and 3_000_000 < total < 12_000_000
    frames = [[xy.copy() for _, xy in HostTrace(w["gen"]).steps(steps)] for w in streams]
    ctxs = [Pipeline(dict(w["cfg"], keep_pairs=1), w["box"]) for w in streams]
    mixed_pairs = [[] for _ in streams]
    for t in range(steps):
        step_batch(ctxs, [frames[i][t] for i in range(len(streams))], t, threads=16)
        for i, gpu in enumerate(ctxs):
            mixed_pairs[i].append(gpu.pairs())
    mixed_state = []
    for gpu in ctxs:
        k, s, e, m = gpu.active()
        rk, rb = sorted_records(*gpu.records())
        mixed_state.append(((k, s, e, m.view(np.uint32)), (rk, rb.view(np.uint32))))
        gpu.close()
    n_pairs = 0
    for i, w in enumerate(streams):
        alone, act, rec = run_stream(w, steps, trace=enumerate(frames[i]))
        for t in range(steps):
            assert np.array_equal(alone[t], mixed_pairs[i][t]), (i, t)
            n_pairs += len(alone[t])
        for a, b in zip(act, mixed_state[i][0]):
            assert np.array_equal(a, b), i
        for a, b in zip(rec, mixed_state[i][1]):
            assert np.array_equal(a, b), i
    print(f"[c4] every stream alone equals the mix: {n_pairs} pairs", flush=True)
    assert n_pairs > 0
    # oracle parity on a sample: the smallest, the largest and two in between
    order = np.argsort([w["cfg"]["num_objects"] for w in streams])
    for i in [int(order[0]), int(order[21]), int(order[42]), int(order[-1])]:
        w = streams[i]
        orc = O.Oracle(dict(w["cfg"], keep_pairs=1), w["box"])
        for t in range(steps):
            orc.step(frames[i][t], t)
            assert np.array_equal(orc.pairs(), mixed_pairs[i][t]), (i, t)
        ok, os_, oe, om = orc.active()
        k, s, e, m = mixed_state[i][0]
        assert np.array_equal(ok, k) and np.array_equal(os_, s) and np.array_equal(oe, e), i
        assert np.array_equal(om.view(np.uint32), m), i
        print(f"[c4] stream {i} ({w['cfg']['num_objects']} objects) equals the oracle", flush=True)
        orc.close()


def test_tuned_capacity_runs_like_the_oracle():
    """the tuner picks one of its candidate capacities for a stream; the stream
    then runs at that capacity exactly as the oracle does"""
    w = batched_streams(4, median=3000, lo=2000, hi=8000, side_km=(2.0, 4.0), steps=12, seed=9)[2]
    frames = [xy.copy() for _, xy in HostTrace(w["gen"]).steps(w["steps"])]
    best, timings = tune_leaf_capacity(w["cfg"], w["box"], frames[:4], candidates=(25, 50, 100, 200))
    print(f"[tuner] n={w['cfg']['num_objects']} timings {timings} -> {best}", flush=True)
    assert best in (25, 50, 100, 200) and set(timings) == {25, 50, 100, 200}
    cfg = dict(w["cfg"], grid_capacity=best, keep_pairs=1)
    gpu, orc = Pipeline(cfg, w["box"]), O.Oracle(cfg, w["box"])
    for t, xy in enumerate(frames):
        g, o = gpu.step(xy, t), orc.step(xy, t)
        assert g["leaves"] == o["leaves"] and g["kv_count"] == o["kv_count"], t
        assert np.array_equal(gpu.pairs(), orc.pairs()), t
    gk, gb = sorted_records(*gpu.records())
    ok, ob = sorted_records(*orc.records())
    assert np.array_equal(gk, ok) and np.array_equal(gb.view(np.uint32), ob.view(np.uint32))
    gpu.close()
    orc.close()


def test_step_batch_reports_each_context():
    """a context whose step fails (a point outside its map box) reports its own
    code through stc_step_batch while the others step normally; the batch
    returns the failure and the failing context's message names it"""
    import ctypes as C
    from stcdb_amd import _lib as L
    streams = batched_streams(3, median=2000, lo=1000, hi=4000, side_km=(2.0, 3.0), steps=3, seed=11)
    pipes = [Pipeline(w["cfg"], w["box"]) for w in streams]
    frames = [next(iter(HostTrace(w["gen"]).steps(1)))[1].copy() for w in streams]
    outs = step_batch(pipes, frames, 0, threads=3)
    assert all(o["leaves"] > 0 for o in outs)
    bad = [f.copy() for f in frames]
    bad[1][0] = (streams[1]["box"][1][0] + 1.0, streams[1]["box"][1][1] + 1.0)  # outside stream 1's box
    lib = L.load()
    n = len(pipes)
    ctxs = (C.c_void_p * n)(*[p.ctx.value for p in pipes])
    ptrs = (C.c_void_p * n)(*[f.ctypes.data for f in bad])
    res = (L.StcStepOut * n)()
    rcs = (C.c_int32 * n)()
    rc = lib.stc_step_batch(ctxs, n, ptrs, None, 1, 3, res, rcs)
    assert rc == L.STC_E_INPUT and list(rcs) == [L.STC_OK, L.STC_E_INPUT, L.STC_OK]
    assert b"outside the map box" in lib.stc_last_error(pipes[1].ctx)
    assert res[0].leaves > 0 and res[2].leaves > 0
    with pytest.raises(L.StcError):
        step_batch(pipes, bad, 2, threads=3)
    for p in pipes:
        p.close()
