"""Host-side engine over the C ABI.

Pipeline mirrors the reference's tracer/workbench pair on the GPU path
(src/tracing/trace.h:24-54, src/tracing/process.cpp:378-634): it owns the
device context created by stc_create (cuda_create_device_bench) and runs one
stc_step (process_with_gpu) per timestep. Traces come from the seeded
generator, on the host (HostTrace) or resident in HBM (DeviceTrace).
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _check(rc, ctx=None):
    if rc != L.STC_OK:
        lib = L.load()
        msg = lib.stc_last_error(ctx) if ctx else b""
        raise L.StcError(rc, (msg or b"").decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _gen_struct(gp):
    g = L.StcGenParams()
    for k, v in gp.items():
        setattr(g, k, v)
    return g


# stc_query / stc_hit of include/stcdb.h
QUERY_DTYPE = np.dtype([("kind", "<u4"), ("oid", "<u4"), ("oid2", "<u4"), ("t0", "<u4"), ("t1", "<u4"),
                        ("flags", "<u4"), ("box", "<f8", (4,))])
HIT_DTYPE = np.dtype([("query", "<u4"), ("pid", "<u4"), ("target", "<u4"), ("start", "<u4"), ("end", "<u4"),
                      ("source", "<u4"), ("mbr", "<f8", (4,))])
Q_OID, Q_RANGE, Q_PAIR = 1, 2, 3
QF_COUNT = 1  # STC_QF_COUNT: count the query's hits, store none
# stc_meeting_unit: the reference's 32-byte meeting_unit (src/tracing/workbench.h:62-83)
MEETING_UNIT_DTYPE = np.dtype([("key", "<u8"), ("start", "<u4"), ("end", "<u4"), ("mbr", "<f4", (4,))])
assert MEETING_UNIT_DTYPE.itemsize == 32
# stc_qtnode: the reference's 80-byte QTSchema (src/index/QTree.h:49-62)
QTNODE_DTYPE = np.dtype([("grid_id", "<u4"), ("level", "<i2"), ("type", "<i2"), ("overflow_count", "<i2"),
                         ("underflow_count", "<i2"), ("pad", "<u4"), ("mid_x", "<f8"), ("mid_y", "<f8"),
                         ("mbr", "<f8", (4,)), ("children", "<u4", (4,))])
assert QTNODE_DTYPE.itemsize == 80


class HostTrace:
    """Seeded walk computed on the host; yields (N, 2) float64 arrays."""

    def __init__(self, gen):
        self.lib = L.load()
        self.gp = _gen_struct(gen)
        self.n = gen["num_objects"]
        self.w = (L.StcWalker * self.n)()
        _check(self.lib.stc_gen_init_host(C.byref(self.gp), self.w, 0, self.n))
        self.t = 0

    def positions(self):
        xy = np.empty((self.n, 2), dtype=np.float64)
        _check(self.lib.stc_gen_pos_host(C.byref(self.gp), self.w, self.n, _ptr(xy)))
        return xy

    def advance(self):
        self.t += 1
        _check(self.lib.stc_gen_step_host(C.byref(self.gp), self.w, self.n, self.t))

    def steps(self, count):
        for t in range(count):
            if t > 0:
                self.advance()
            yield t, self.positions()


class DeviceTrace:
    """Seeded walk resident in HBM; next(t) returns a device pointer."""

    def __init__(self, gen):
        self.lib = L.load()
        self.gp = _gen_struct(gen)
        self.h = C.c_void_p()
        _check(self.lib.stc_trace_create(C.byref(self.gp), C.byref(self.h)))

    def next(self, t):
        p = C.c_void_p()
        _check(self.lib.stc_trace_next(self.h, t, C.byref(p)))
        return p

    def close(self):
        if self.h:
            self.lib.stc_trace_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pipeline:
    """One device context: the per-timestep co-occurrence pipeline."""

    def __init__(self, cfg, box):
        self.lib = L.load()
        self.cfg = dict(cfg)
        c = L.StcConfig()
        for k, v in cfg.items():
            setattr(c, k, v)
        b = L.StcBox()
        b.low[0], b.low[1] = box[0]
        b.high[0], b.high[1] = box[1]
        self.ctx = C.c_void_p()
        rc = self.lib.stc_create(C.byref(c), C.byref(b), C.byref(self.ctx))
        if rc != L.STC_OK:
            raise L.StcError(rc, "stc_create failed: " + (self.lib.stc_last_error(None) or b"").decode())
        self.n = cfg["num_objects"]

    def step(self, xy, t):
        """one timestep from host points (N, 2) float64"""
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        assert xy.shape == (self.n, 2)
        out = L.StcStepOut()
        _check(self.lib.stc_step(self.ctx, _ptr(xy), t, C.byref(out)), self.ctx)
        return out.as_dict()

    def step_host_ptr(self, xy_host, t):
        """one timestep from a raw host pointer to N AoS float64 points (e.g. pinned memory)"""
        out = L.StcStepOut()
        _check(self.lib.stc_step(self.ctx, xy_host, t, C.byref(out)), self.ctx)
        return out.as_dict()

    def step_device(self, xy_dev, t):
        out = L.StcStepOut()
        _check(self.lib.stc_step_device(self.ctx, xy_dev, t, C.byref(out)), self.ctx)
        return out.as_dict()

    def step_meetings(self, units, t):
        """one timestep from persisted meeting units (MEETING_UNIT_DTYPE) instead of positions"""
        units = np.ascontiguousarray(units, dtype=MEETING_UNIT_DTYPE)
        out = L.StcStepOut()
        _check(self.lib.stc_step_meetings(self.ctx, _ptr(units), len(units), t, C.byref(out)), self.ctx)
        return out.as_dict()

    def step_meeting_units(self):
        """the last step's pairs as persisted meeting units (needs keep_pairs bit 0)"""
        n = self._count(self.lib.stc_get_step_meetings, None)
        u = np.empty(n, dtype=MEETING_UNIT_DTYPE)
        m = C.c_uint64()
        _check(self.lib.stc_get_step_meetings(self.ctx, _ptr(u), n, C.byref(m)), self.ctx)
        return u

    def _count(self, fn, *args):
        n = C.c_uint64()
        _check(fn(self.ctx, *args, 0, C.byref(n)), self.ctx)
        return n.value

    def pairs(self):
        n = self._count(self.lib.stc_get_pairs, None)
        keys = np.empty(n, dtype=np.uint64)
        m = C.c_uint64()
        _check(self.lib.stc_get_pairs(self.ctx, _ptr(keys), n, C.byref(m)), self.ctx)
        return keys

    def active(self):
        n = self._count(self.lib.stc_get_active, None, None, None, None)
        keys = np.empty(n, dtype=np.uint64)
        start = np.empty(n, dtype=np.uint32)
        end = np.empty(n, dtype=np.uint32)
        mbr = np.empty((n, 4), dtype=np.float32)
        m = C.c_uint64()
        _check(self.lib.stc_get_active(self.ctx, _ptr(keys), _ptr(start), _ptr(end), _ptr(mbr), n, C.byref(m)),
               self.ctx)
        return keys, start, end, mbr

    def records(self):
        n = self._count(self.lib.stc_get_records, None, None)
        keys = np.empty((n, 2), dtype=np.uint64)
        boxes = np.empty((n, 4), dtype=np.float32)
        m = C.c_uint64()
        _check(self.lib.stc_get_records(self.ctx, _ptr(keys), _ptr(boxes), n, C.byref(m)), self.ctx)
        return keys, boxes

    def build_schema(self, xy, on_device=False):
        """dynamic schema: the initial QTree from one timestep's points (host
        array (N, 2), or a device pointer with on_device)"""
        if on_device:
            _check(self.lib.stc_build_schema(self.ctx, xy, 1), self.ctx)
        else:
            xy = np.ascontiguousarray(xy, dtype=np.float64)
            assert xy.shape == (self.n, 2)
            _check(self.lib.stc_build_schema(self.ctx, _ptr(xy), 0), self.ctx)

    def set_schema(self, nodes):
        """dynamic schema: install a QTree (QTNODE_DTYPE array, node 0 the root)"""
        nodes = np.ascontiguousarray(nodes, dtype=QTNODE_DTYPE)
        _check(self.lib.stc_set_schema(self.ctx, _ptr(nodes), len(nodes)), self.ctx)

    def schema(self):
        """the current QTree as a QTNODE_DTYPE array (slots; type 0 = free)"""
        n = self._count(self.lib.stc_get_schema, None)
        out = np.zeros(n, dtype=QTNODE_DTYPE)
        m = C.c_uint64()
        _check(self.lib.stc_get_schema(self.ctx, _ptr(out), n, C.byref(m)), self.ctx)
        return out

    def leaves(self):
        n = self._count(self.lib.stc_get_leaves, None, None)
        boxes = np.empty((n, 4), dtype=np.float64)
        counts = np.empty(n, dtype=np.uint32)
        m = C.c_uint64()
        _check(self.lib.stc_get_leaves(self.ctx, _ptr(boxes), _ptr(counts), n, C.byref(m)), self.ctx)
        return boxes, counts

    def num_runs(self):
        n = C.c_uint32()
        _check(self.lib.stc_num_runs(self.ctx, C.byref(n)), self.ctx)
        return n.value

    def run_info(self, i):
        """stc_get_run_info as a dict (no wait for the run's host copy)"""
        info = L.StcRunInfo()
        _check(self.lib.stc_get_run_info(self.ctx, i, C.byref(info)), self.ctx)
        return {k: getattr(info, k) for k, _ in info._fields_}

    def run(self, i):
        info = L.StcRunInfo()
        _check(self.lib.stc_get_run_info(self.ctx, i, C.byref(info)), self.ctx)
        nc, no, nk, nf = info.ctf_count, info.oversize_count, info.kv_count, info.num_objects
        ctfs = (L.StcCtfMeta * nc)()
        key_bytes = np.empty(info.key_bytes, dtype=np.uint8)
        bitmaps = np.empty(nc * 8192, dtype=np.uint8)
        sids = np.empty(nf, dtype=np.uint16)
        o_keys = np.empty((no, 2), dtype=np.uint64)
        o_boxes = np.empty((no, 4), dtype=np.float32)
        o_bitmap = np.empty(8192, dtype=np.uint8)
        sk = np.empty((nk, 2), dtype=np.uint64)
        sb = np.empty((nk, 4), dtype=np.float32)
        _check(self.lib.stc_get_run(self.ctx, i, ctfs, _ptr(key_bytes), _ptr(bitmaps), _ptr(sids), _ptr(o_keys),
                                    _ptr(o_boxes), _ptr(o_bitmap), _ptr(sk), _ptr(sb)), self.ctx)
        return dict(info={k: getattr(info, k) for k, _ in info._fields_}, ctfs=ctfs_to_numpy(ctfs),
                    key_bytes=key_bytes, bitmaps=bitmaps, sids=sids, o_keys=o_keys, o_boxes=o_boxes,
                    o_bitmap=o_bitmap, sorted_keys=sk, sorted_boxes=sb)

    def query(self, q, cap=None, into=None):
        """answer a batch of queries (QUERY_DTYPE) over the state after the last
        step; returns (hits as HIT_DTYPE sorted by query, source, pid, target,
        start, end; counts per query). Queries flagged QF_COUNT store no hits.
        `into`: a HIT_DTYPE array of at least cap entries to store them in."""
        q = np.ascontiguousarray(q, dtype=QUERY_DTYPE)
        counts = np.zeros(len(q), dtype=np.uint64)
        n = C.c_uint64()
        if cap is None:
            _check(self.lib.stc_query_batch(self.ctx, _ptr(q), len(q), None, 0, C.byref(n), _ptr(counts)), self.ctx)
            cap = n.value
        hits = into if into is not None and len(into) >= cap else np.zeros(cap, dtype=HIT_DTYPE)
        _check(self.lib.stc_query_batch(self.ctx, _ptr(q), len(q), _ptr(hits), cap, C.byref(n), _ptr(counts)), self.ctx)
        return hits[:min(n.value, cap)], counts

    def query_at(self, q, into):
        """a query batch over the last published state (safe while another
        thread steps); returns (hits, counts, the state's timestep). `into` is
        a HIT_DTYPE buffer; hits beyond it are counted, not stored"""
        q = np.ascontiguousarray(q, dtype=QUERY_DTYPE)
        counts = np.zeros(len(q), dtype=np.uint64)
        n = C.c_uint64()
        st = C.c_uint32()
        _check(self.lib.stc_query_batch_at(self.ctx, _ptr(q), len(q), _ptr(into), len(into), C.byref(n),
                                           _ptr(counts), C.byref(st)), self.ctx)
        return into[:min(n.value, len(into))], counts, st.value

    def live_search(self, pids):
        """established live contacts of the given objects (stc_live_search);
        hits as HIT_DTYPE with `query` the index into pids"""
        pids = np.ascontiguousarray(pids, dtype=np.uint32)
        n = C.c_uint64()
        _check(self.lib.stc_live_search(self.ctx, _ptr(pids), len(pids), None, 0, C.byref(n)), self.ctx)
        hits = np.zeros(n.value, dtype=HIT_DTYPE)
        _check(self.lib.stc_live_search(self.ctx, _ptr(pids), len(pids), _ptr(hits), len(hits), C.byref(n)),
               self.ctx)
        return hits

    def dump_run(self, i, directory, prefix, ctb):
        """write run i in the reference's file layout (stc_dump_run)"""
        _check(self.lib.stc_dump_run(self.ctx, i, directory.encode(), prefix.encode(), ctb), self.ctx)

    def compact_runs(self, first, count):
        """merge runs [first, first + count) into one run in their place (stc_compact_runs)"""
        _check(self.lib.stc_compact_runs(self.ctx, first, count), self.ctx)

    def drop_runs(self):
        _check(self.lib.stc_drop_runs(self.ctx), self.ctx)

    def sync_runs(self):
        """wait for every background copy of a flushed run"""
        _check(self.lib.stc_sync(self.ctx), self.ctx)

    def close(self):
        if self.ctx:
            self.lib.stc_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CTF_DTYPE = np.dtype([
    ("mbr", np.float32, 4), ("kv_capacity", np.uint32), ("start_time_min", np.uint32),
    ("start_time_max", np.uint32), ("end_time_min", np.uint32), ("end_time_max", np.uint32),
    ("bitmap_size", np.uint32), ("key_bit", np.uint16), ("id_bit", np.uint16), ("duration_bit", np.uint16),
    ("end_bit", np.uint16), ("low_x_bit", np.uint16), ("low_y_bit", np.uint16), ("edge_bit", np.uint16),
    ("mbr_bit", np.uint16), ("x_grid", np.uint16), ("y_grid", np.uint16), ("flags", np.uint32),
])


def step_batch(pipes, frames, t, threads=16, on_device=False):
    """One timestep of independent contexts (batched streams) through
    stc_step_batch: frames[i] is pipes[i]'s points, host arrays (N_i, 2)
    float64 or, with on_device, device pointers. Returns each context's
    counters; raises on the first failing context."""
    n = len(pipes)
    lib = L.load()
    ctxs = (C.c_void_p * n)(*[p.ctx.value for p in pipes])
    keep = []
    ptrs = (C.c_void_p * n)()
    for i, f in enumerate(frames):
        if on_device:
            ptrs[i] = f.value if isinstance(f, C.c_void_p) else int(f)
        else:
            a = np.ascontiguousarray(f, dtype=np.float64)
            assert a.shape == (pipes[i].n, 2)
            keep.append(a)
            ptrs[i] = a.ctypes.data
    dev = (C.c_int32 * n)(*([1 if on_device else 0] * n))
    outs = (L.StcStepOut * n)()
    rcs = (C.c_int32 * n)()
    rc = lib.stc_step_batch(ctxs, n, ptrs, dev, t, threads, outs, rcs)
    if rc != L.STC_OK:
        bad = next(i for i in range(n) if rcs[i] != L.STC_OK)
        _check(rcs[bad], pipes[bad].ctx)
    return [outs[i].as_dict() for i in range(n)]


def ctfs_to_numpy(arr):
    raw = bytes(memoryview(arr))
    return np.frombuffer(raw, dtype=CTF_DTYPE).copy()
