"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference
shim (oracle/_ref/libstcref.so). Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# STC_ORACLE_LIB points the tests at another build of the oracle (mutation checks of the pins)
ORACLE_LIB = os.environ.get("STC_ORACLE_LIB", os.path.join(ORACLE_DIR, "liboracle.so"))
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libstcref.so")

ORC_C = ["nodes", "leaves", "border", "units_inside", "units_cross", "candidates", "pairs", "active",
         "expired", "new_records", "kv_count", "overflow", "flushed"]


class OrcConfig(C.Structure):
    _fields_ = [
        ("num_objects", C.c_uint32), ("grid_capacity", C.c_uint32), ("leaf_buffer", C.c_uint32),
        ("min_meet_time", C.c_uint32), ("max_meet_time", C.c_uint32), ("kv_restriction", C.c_uint32),
        ("ctf_count", C.c_uint32), ("start_time", C.c_uint32), ("compat", C.c_uint32),
        ("schema_mode", C.c_uint32),
        ("reach_distance", C.c_double), ("area_threshold", C.c_double),
        ("map_lo", C.c_double * 2), ("map_hi", C.c_double * 2),
        ("schema_update_delay", C.c_uint32), ("pad", C.c_uint32),
    ]


class OrcRunInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in
                ["kv_count", "oversize_count", "ctf_count", "num_objects", "o_start_min", "o_start_max",
                 "o_end_min", "o_end_max", "end_time_min", "cur_time", "sid_count", "oversize_oid_count"]] + \
               [("key_bytes", C.c_uint64)]


QUERY_DTYPE = np.dtype([("kind", "<u4"), ("oid", "<u4"), ("oid2", "<u4"), ("t0", "<u4"), ("t1", "<u4"),
                        ("flags", "<u4"), ("box", "<f8", (4,))])
HIT_DTYPE = np.dtype([("query", "<u4"), ("pid", "<u4"), ("target", "<u4"), ("start", "<u4"), ("end", "<u4"),
                      ("source", "<u4"), ("mbr", "<f8", (4,))])


def sort_hits(h):
    """canonical order of query results: (query, source, pid, target, start, end)"""
    return h[np.lexsort((h["end"], h["start"], h["target"], h["pid"], h["source"], h["query"]))]


CTF_DTYPE = np.dtype([
    ("mbr", np.float32, 4), ("kv_capacity", np.uint32), ("start_time_min", np.uint32),
    ("start_time_max", np.uint32), ("end_time_min", np.uint32), ("end_time_max", np.uint32),
    ("bitmap_size", np.uint32), ("key_bit", np.uint16), ("id_bit", np.uint16), ("duration_bit", np.uint16),
    ("end_bit", np.uint16), ("low_x_bit", np.uint16), ("low_y_bit", np.uint16), ("edge_bit", np.uint16),
    ("mbr_bit", np.uint16), ("x_grid", np.uint16), ("y_grid", np.uint16), ("flags", np.uint32),
])

_lib = None
_ref = None


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True, capture_output=True)
        lib = C.CDLL(ORACLE_LIB)
        V = C.c_void_p
        lib.orc_create.restype = V
        lib.orc_create.argtypes = [C.POINTER(OrcConfig)]
        lib.orc_destroy.argtypes = [V]
        lib.orc_build_schema.restype = C.c_int
        lib.orc_build_schema.argtypes = [V, V]
        lib.orc_step.restype = C.c_int
        lib.orc_step.argtypes = [V, V, C.c_uint32]
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_last_error.argtypes = [V]
        lib.orc_counts.argtypes = [V, V]
        lib.orc_last_flush_ms.restype = C.c_double
        lib.orc_last_flush_ms.argtypes = [V, C.POINTER(C.c_uint64)]
        for name in ["orc_num_pairs", "orc_num_leaves", "orc_num_nodes", "orc_num_units", "orc_num_active",
                     "orc_num_records"]:
            getattr(lib, name).restype = C.c_uint64
            getattr(lib, name).argtypes = [V]
        lib.orc_get_pairs.argtypes = [V, V]
        lib.orc_get_leaves.argtypes = [V, V, V]
        lib.orc_get_point_leaf.argtypes = [V, V, V]
        lib.orc_get_nodes.argtypes = [V, V, V, V, V]
        lib.orc_get_units.argtypes = [V, V, V, V]
        lib.orc_get_active.argtypes = [V, V, V, V, V]
        lib.orc_get_records.argtypes = [V, V, V]
        lib.orc_num_runs.restype = C.c_uint32
        lib.orc_num_runs.argtypes = [V]
        lib.orc_get_run_info.argtypes = [V, C.c_uint32, C.POINTER(OrcRunInfo)]
        lib.orc_get_run.argtypes = [V, C.c_uint32] + [V] * 9
        lib.orc_dpm_lon.restype = C.c_double
        lib.orc_dpm_lon.argtypes = [C.c_double]
        lib.orc_point_distance.restype = C.c_double
        lib.orc_point_distance.argtypes = [C.c_double] * 4
        lib.orc_box_distance.restype = C.c_double
        lib.orc_box_distance.argtypes = [V, C.c_double, C.c_double]
        lib.orc_min_bits.restype = C.c_uint32
        lib.orc_min_bits.argtypes = [C.c_uint32]
        lib.orc_conv_d2u32.restype = C.c_uint32
        lib.orc_conv_d2u32.argtypes = [C.c_double]
        lib.orc_conv_d2i32.restype = C.c_int32
        lib.orc_conv_d2i32.argtypes = [C.c_double]
        lib.orc_ctf_bits.argtypes = [V, V, C.c_uint32, C.c_uint32]
        lib.orc_query_batch.restype = C.c_uint64
        lib.orc_query_batch.argtypes = [V, V, C.c_uint32, V, C.c_uint64, V]
        lib.orc_end_time_min.restype = C.c_uint32
        lib.orc_end_time_min.argtypes = [V]
        lib.orc_get_step_meetings.argtypes = [V, V]
        lib.orc_step_meetings.restype = C.c_int
        lib.orc_step_meetings.argtypes = [V, V, C.c_uint64, C.c_uint32]
        lib.orc_compact_runs.restype = C.c_int
        lib.orc_compact_runs.argtypes = [V, C.c_uint32, C.c_uint32]
        lib.orc_brute_pairs.restype = C.c_uint64
        lib.orc_brute_pairs.argtypes = [V, C.c_uint32, C.c_double, C.c_int, V, C.c_uint64]
        _lib = lib
    return _lib


def load_ref():
    """the reference's own helpers compiled from /root/reference (None when unavailable)"""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        lib = C.CDLL(REF_LIB)
        D = C.c_double
        lib.ref_point_distance_geo.restype = D
        lib.ref_point_distance_geo.argtypes = [D] * 4
        lib.ref_box_distance_geo.restype = D
        lib.ref_box_distance_geo.argtypes = [C.c_void_p, D, D]
        lib.ref_dpm_lon.restype = D
        lib.ref_dpm_lon.argtypes = [D]
        lib.ref_dpm_lat.restype = D
        lib.ref_cantor.restype = C.c_uint64
        lib.ref_cantor.argtypes = [C.c_uint32, C.c_uint32]
        lib.ref_inverse_cantor.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p]
        lib.ref_min_bits.restype = C.c_uint32
        lib.ref_min_bits.argtypes = [C.c_uint32]
        lib.ref_key_fields.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        lib.ref_conv_d2u.restype = C.c_uint32
        lib.ref_conv_d2u.argtypes = [D]
        lib.ref_conv_d2i.restype = C.c_int32
        lib.ref_conv_d2i.argtypes = [D]
        _ref = lib
    return _ref


def orc_config(cfg, box):
    c = OrcConfig()
    c.num_objects = cfg["num_objects"]
    c.grid_capacity = cfg["grid_capacity"]
    c.leaf_buffer = int(cfg["grid_amplify"] * cfg["grid_capacity"])
    c.min_meet_time = cfg["min_meet_time"]
    c.max_meet_time = cfg["max_meet_time"]
    c.kv_restriction = cfg["kv_restriction"]
    c.ctf_count = cfg["ctf_count"]
    c.start_time = cfg["start_time"]
    c.compat = cfg["compat"]
    # 0 rebuild per step; 1 dynamic schema with the GPU's update rules; 2 with the CPU path's rules
    c.schema_mode = cfg.get("schema_mode", 1 if cfg.get("dynamic_schema", 0) else 0)
    c.schema_update_delay = cfg.get("schema_update_delay", 1)
    c.reach_distance = cfg["reach_distance"]
    c.area_threshold = cfg["area_threshold"]
    c.map_lo[0], c.map_lo[1] = box[0]
    c.map_hi[0], c.map_hi[1] = box[1]
    return c


class Oracle:
    def __init__(self, cfg, box):
        self.lib = load()
        self.cfg = cfg
        self.box = box
        self._c = orc_config(cfg, box)
        self.h = self.lib.orc_create(C.byref(self._c))
        self.n = cfg["num_objects"]

    def step(self, xy, t):
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        rc = self.lib.orc_step(self.h, _p(xy), t)
        if rc != 0:
            raise RuntimeError("oracle: " + self.lib.orc_last_error(self.h).decode())
        return self.counts()

    def last_flush(self):
        """(wall ms, records) of the last flush a step ran"""
        r = C.c_uint64()
        ms = self.lib.orc_last_flush_ms(self.h, C.byref(r))
        return ms, r.value

    def build_schema(self, xy):
        """dynamic schema: the initial tree from one timestep's points"""
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        if self.lib.orc_build_schema(self.h, _p(xy)) != 0:
            raise RuntimeError("oracle: " + self.lib.orc_last_error(self.h).decode())

    def step_meetings(self, units, t):
        """a step from persisted meeting units instead of positions"""
        from stcdb_amd.engine import MEETING_UNIT_DTYPE
        units = np.ascontiguousarray(units, dtype=MEETING_UNIT_DTYPE)
        if self.lib.orc_step_meetings(self.h, _p(units), len(units), t) != 0:
            raise RuntimeError("oracle: " + self.lib.orc_last_error(self.h).decode())
        return self.counts()

    def step_meeting_units(self):
        """the last step's pairs as meeting units, in pair order"""
        from stcdb_amd.engine import MEETING_UNIT_DTYPE
        u = np.empty(self.lib.orc_num_pairs(self.h), dtype=MEETING_UNIT_DTYPE)
        self.lib.orc_get_step_meetings(self.h, _p(u))
        return u

    def counts(self):
        a = np.zeros(len(ORC_C), dtype=np.uint64)
        self.lib.orc_counts(self.h, _p(a))
        return {k: int(v) for k, v in zip(ORC_C, a)}

    def pairs(self):
        n = self.lib.orc_num_pairs(self.h)
        k = np.empty(n, dtype=np.uint64)
        self.lib.orc_get_pairs(self.h, _p(k))
        return k

    def leaves(self):
        n = self.lib.orc_num_leaves(self.h)
        b = np.empty((n, 4), dtype=np.float64)
        c = np.empty(n, dtype=np.uint32)
        self.lib.orc_get_leaves(self.h, _p(b), _p(c))
        return b, c

    def point_leaf(self):
        leaf = np.empty(self.n, dtype=np.uint32)
        border = np.empty(self.n, dtype=np.uint8)
        self.lib.orc_get_point_leaf(self.h, _p(leaf), _p(border))
        return leaf, border

    def nodes(self):
        n = self.lib.orc_num_nodes(self.h)
        boxes = np.empty((n, 4), dtype=np.float64)
        mids = np.empty((n, 2), dtype=np.float64)
        typ = np.empty(n, dtype=np.uint32)
        ch = np.empty((n, 4), dtype=np.uint32)
        self.lib.orc_get_nodes(self.h, _p(boxes), _p(mids), _p(typ), _p(ch))
        return boxes, mids, typ, ch

    def units(self):
        n = self.lib.orc_num_units(self.h)
        pid = np.empty(n, dtype=np.uint32)
        node = np.empty(n, dtype=np.uint32)
        inside = np.empty(n, dtype=np.uint8)
        self.lib.orc_get_units(self.h, _p(pid), _p(node), _p(inside))
        return pid, node, inside

    def active(self):
        n = self.lib.orc_num_active(self.h)
        keys = np.empty(n, dtype=np.uint64)
        start = np.empty(n, dtype=np.uint32)
        end = np.empty(n, dtype=np.uint32)
        mbr = np.empty((n, 4), dtype=np.float32)
        self.lib.orc_get_active(self.h, _p(keys), _p(start), _p(end), _p(mbr))
        return keys, start, end, mbr

    def records(self):
        n = self.lib.orc_num_records(self.h)
        keys = np.empty((n, 2), dtype=np.uint64)
        boxes = np.empty((n, 4), dtype=np.float32)
        self.lib.orc_get_records(self.h, _p(keys), _p(boxes))
        return keys, boxes

    def num_runs(self):
        return self.lib.orc_num_runs(self.h)

    def run(self, i):
        if i >= self.num_runs():
            raise IndexError(f"run {i} of {self.num_runs()}")
        info = OrcRunInfo()
        self.lib.orc_get_run_info(self.h, i, C.byref(info))
        nc, no, nk, nf = info.ctf_count, info.oversize_count, info.kv_count, info.num_objects
        ctfs = np.empty(nc, dtype=CTF_DTYPE)
        key_bytes = np.empty(info.key_bytes, dtype=np.uint8)
        bitmaps = np.empty(nc * 8192, dtype=np.uint8)
        sids = np.empty(nf, dtype=np.uint16)
        o_keys = np.empty((no, 2), dtype=np.uint64)
        o_boxes = np.empty((no, 4), dtype=np.float32)
        o_bitmap = np.empty(8192, dtype=np.uint8)
        sk = np.empty((nk, 2), dtype=np.uint64)
        sb = np.empty((nk, 4), dtype=np.float32)
        self.lib.orc_get_run(self.h, i, _p(ctfs), _p(key_bytes), _p(bitmaps), _p(sids), _p(o_keys), _p(o_boxes),
                             _p(o_bitmap), _p(sk), _p(sb))
        return dict(info={k: getattr(info, k) for k, _ in info._fields_}, ctfs=ctfs, key_bytes=key_bytes,
                    bitmaps=bitmaps, sids=sids, o_keys=o_keys, o_boxes=o_boxes, o_bitmap=o_bitmap,
                    sorted_keys=sk, sorted_boxes=sb)

    def end_time_min(self):
        return self.lib.orc_end_time_min(self.h)

    def compact_runs(self, first, count):
        """merge runs [first, first + count) into one run in their place"""
        if self.lib.orc_compact_runs(self.h, first, count) != 0:
            raise RuntimeError("oracle: " + self.lib.orc_last_error(self.h).decode())

    def query(self, q):
        """answer a batch of queries (QUERY_DTYPE); returns (hits sorted, counts)"""
        q = np.ascontiguousarray(q, dtype=QUERY_DTYPE)
        counts = np.zeros(len(q), dtype=np.uint64)
        n = self.lib.orc_query_batch(self.h, _p(q), len(q), None, 0, _p(counts))
        hits = np.zeros(n, dtype=HIT_DTYPE)
        self.lib.orc_query_batch(self.h, _p(q), len(q), _p(hits), n, _p(counts))
        return sort_hits(hits), counts

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def brute_pairs(xy, r, probe_low):
    lib = load()
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    cnt = lib.orc_brute_pairs(_p(xy), n, r, probe_low, None, 0)
    keys = np.empty(cnt, dtype=np.uint64)
    lib.orc_brute_pairs(_p(xy), n, r, probe_low, _p(keys), cnt)
    return np.unique(keys)


def dpm_lon(lat):
    return load().orc_dpm_lon(lat)


def point_distance(x1, y1, x2, y2):
    return load().orc_point_distance(x1, y1, x2, y2)


def box_distance(box4, px, py):
    b = np.ascontiguousarray(box4, dtype=np.float64)
    return load().orc_box_distance(_p(b), px, py)


CPATH_LIB = os.path.join(ORACLE_DIR, "libcpupath.so")
_cpath = None


def load_cpath():
    global _cpath
    if _cpath is None:
        if not os.path.exists(CPATH_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "libcpupath.so"], check=True, capture_output=True)
        lib = C.CDLL(CPATH_LIB)
        V = C.c_void_p
        lib.cpath_create.restype = V
        lib.cpath_create.argtypes = [C.POINTER(OrcConfig), C.c_uint32, C.c_uint64, C.c_uint32]
        lib.cpath_destroy.argtypes = [V]
        lib.cpath_step.restype = C.c_int
        lib.cpath_step.argtypes = [V, V, C.c_uint32]
        lib.cpath_last_error.restype = C.c_char_p
        lib.cpath_last_error.argtypes = [V]
        lib.cpath_stage_ms.argtypes = [V, V]
        lib.cpath_counts.argtypes = [V, V]
        for name in ["cpath_num_pairs", "cpath_num_meetings", "cpath_num_leaves"]:
            getattr(lib, name).restype = C.c_uint64
            getattr(lib, name).argtypes = [V]
        lib.cpath_get_pairs.argtypes = [V, V]
        lib.cpath_get_meetings.argtypes = [V, V]
        lib.cpath_get_leaves.argtypes = [V, V, V]
        _cpath = lib
    return _cpath


class CpuPath:
    """the reference's CPU path (filter, reachability, update_meetings,
    update_schema on `threads` pthreads; oracle/stc_cpu_path.c)"""

    STAGES = ["filter", "reachability", "update_meetings", "update_schema"]

    def __init__(self, cfg, box, threads, buckets=0, keep_pairs=True):
        self.lib = load_cpath()
        c = dict(cfg)
        c.setdefault("schema_mode", 2 if c.get("dynamic_schema", 1) else 0)
        self._c = orc_config(c, box)
        self.h = self.lib.cpath_create(C.byref(self._c), threads, buckets, 1 if keep_pairs else 0)
        if not self.h:
            raise MemoryError("cpath_create")

    def step(self, xy, t):
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        if self.lib.cpath_step(self.h, _p(xy), t) != 0:
            raise RuntimeError("cpu path: " + self.lib.cpath_last_error(self.h).decode())
        ms = np.zeros(4)
        self.lib.cpath_stage_ms(self.h, _p(ms))
        cnt = np.zeros(6, dtype=np.uint64)
        self.lib.cpath_counts(self.h, _p(cnt))
        out = dict(zip(["units", "candidates", "pairs", "taken", "active", "identified"], map(int, cnt)))
        out["ms"] = dict(zip(self.STAGES, ms.tolist()))
        return out

    def pairs(self):
        k = np.empty(self.lib.cpath_num_pairs(self.h), dtype=np.uint64)
        self.lib.cpath_get_pairs(self.h, _p(k))
        return k

    def meetings(self):
        m = np.empty((self.lib.cpath_num_meetings(self.h), 4), dtype=np.uint32)
        self.lib.cpath_get_meetings(self.h, _p(m))
        return m

    def leaves(self):
        n = self.lib.cpath_num_leaves(self.h)
        b = np.empty((n, 4), dtype=np.float64)
        c = np.empty(n, dtype=np.uint32)
        self.lib.cpath_get_leaves(self.h, _p(b), _p(c))
        order = np.lexsort((b[:, 0], b[:, 1]))
        return b[order], c[order]

    def close(self):
        if self.h:
            self.lib.cpath_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
