"""ctypes binding of libstcdb_hip.so (the C ABI of include/stcdb.h).

The library is built in-tree by stcdb_amd/build.py. There is no fallback: if
the library is missing, loading raises, so a GPU run never silently executes
anything but the HIP path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstcdb_hip.so")

STC_OK = 0
STC_E_ARG = -1
STC_E_HIP = -2
STC_E_CAPACITY = -3
STC_E_INPUT = -4
STC_E_STATE = -5
STC_COMPAT_HALF_REDUCTION = 1


class StcConfig(C.Structure):
    _fields_ = [
        ("num_objects", C.c_uint32),
        ("grid_capacity", C.c_uint32),
        ("grid_amplify", C.c_double),
        ("min_meet_time", C.c_uint32),
        ("max_meet_time", C.c_uint32),
        ("start_time", C.c_uint32),
        ("kv_restriction", C.c_uint32),
        ("ctf_count", C.c_uint32),
        ("compat", C.c_uint32),
        ("num_meeting_buckets", C.c_uint64),
        ("reach_distance", C.c_double),
        ("area_threshold", C.c_double),
        ("keep_pairs", C.c_uint32),
        ("device_runs", C.c_uint32),
        ("dynamic_schema", C.c_uint32),
        ("schema_update_delay", C.c_uint32),
        ("async_flush", C.c_uint32),
    ]


class StcBox(C.Structure):
    _fields_ = [("low", C.c_double * 2), ("high", C.c_double * 2)]


class StcStepOut(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint64),
        ("leaves", C.c_uint64),
        ("border_points", C.c_uint64),
        ("cross_units", C.c_uint64),
        ("candidates", C.c_uint64),
        ("pairs", C.c_uint64),
        ("active_meetings", C.c_uint64),
        ("expired", C.c_uint64),
        ("new_records", C.c_uint64),
        ("kv_count", C.c_uint64),
        ("overflow_points", C.c_uint64),
        ("ambiguous_pairs", C.c_uint64),
        ("ambiguous_nodes", C.c_uint64),
        ("table_buckets", C.c_uint64),
        ("table_occupied", C.c_uint64),
        ("flushed", C.c_uint32),
        ("rebuilt", C.c_uint32),
        ("ms_partition", C.c_float),
        ("ms_filter", C.c_float),
        ("ms_refine", C.c_float),
        ("ms_meetings", C.c_float),
        ("ms_flush", C.c_float),
        ("ms_total", C.c_float),
        ("ms_refine_kernel", C.c_float),
        ("ms_identify_kernel", C.c_float),
        ("continuing", C.c_uint64),
        ("new_keys", C.c_uint64),
        ("ms_count_kernel", C.c_float),
        ("ms_schema", C.c_float),
        ("ms_flush_done", C.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class StcCtfMeta(C.Structure):
    _fields_ = [
        ("mbr", C.c_float * 4),
        ("kv_capacity", C.c_uint32),
        ("start_time_min", C.c_uint32),
        ("start_time_max", C.c_uint32),
        ("end_time_min", C.c_uint32),
        ("end_time_max", C.c_uint32),
        ("bitmap_size", C.c_uint32),
        ("key_bit", C.c_uint16),
        ("id_bit", C.c_uint16),
        ("duration_bit", C.c_uint16),
        ("end_bit", C.c_uint16),
        ("low_x_bit", C.c_uint16),
        ("low_y_bit", C.c_uint16),
        ("edge_bit", C.c_uint16),
        ("mbr_bit", C.c_uint16),
        ("x_grid", C.c_uint16),
        ("y_grid", C.c_uint16),
        ("flags", C.c_uint32),
    ]


class StcQtNode(C.Structure):
    """QTSchema (src/index/QTree.h:49-62), 80 bytes"""
    _fields_ = [
        ("grid_id", C.c_uint32),
        ("level", C.c_int16),
        ("type", C.c_int16),
        ("overflow_count", C.c_int16),
        ("underflow_count", C.c_int16),
        ("mid_x", C.c_double),
        ("mid_y", C.c_double),
        ("mbr", C.c_double * 4),
        ("children", C.c_uint32 * 4),
    ]


class StcRunInfo(C.Structure):
    _fields_ = [
        ("kv_count", C.c_uint32),
        ("oversize_count", C.c_uint32),
        ("ctf_count", C.c_uint32),
        ("num_objects", C.c_uint32),
        ("o_start_min", C.c_uint32),
        ("o_start_max", C.c_uint32),
        ("o_end_min", C.c_uint32),
        ("o_end_max", C.c_uint32),
        ("end_time_min", C.c_uint32),
        ("cur_time", C.c_uint32),
        ("sid_count", C.c_uint32),
        ("oversize_oid_count", C.c_uint32),
        ("key_bytes", C.c_uint64),
    ]


class StcDeviceInfo(C.Structure):
    _fields_ = [
        ("device_id", C.c_int32),
        ("name", C.c_char * 256),
        ("mem_size", C.c_uint64),
        ("num_cus", C.c_int32),
        ("clock_khz", C.c_int32),
    ]


class StcGenParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("num_objects", C.c_uint32),
        ("grid_cells", C.c_uint32),
        ("spacing", C.c_double),
        ("vmin", C.c_double),
        ("vmax", C.c_double),
        ("n_hotspots", C.c_uint32),
        ("pad", C.c_uint32),
        ("hotspot_frac", C.c_double),
        ("hotspot_sigma", C.c_double),
        ("lon0", C.c_double),
        ("lat0", C.c_double),
        ("dlon", C.c_double),
        ("dlat", C.c_double),
    ]


class StcWalker(C.Structure):
    _fields_ = [
        ("ix", C.c_int32),
        ("iy", C.c_int32),
        ("dir", C.c_int32),
        ("pad", C.c_int32),
        ("off", C.c_double),
        ("speed", C.c_double),
    ]


P = C.c_void_p
U64P = C.POINTER(C.c_uint64)

# name -> (restype, argtypes); every symbol declared in include/stcdb.h
SIGNATURES = {
    "stc_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "stc_device_info_get": (C.c_int, [C.c_int32, C.POINTER(StcDeviceInfo)]),
    "stc_create": (C.c_int, [C.POINTER(StcConfig), C.POINTER(StcBox), C.POINTER(C.c_void_p)]),
    "stc_destroy": (C.c_int, [P]),
    "stc_build_schema": (C.c_int, [P, P, C.c_int32]),
    "stc_set_schema": (C.c_int, [P, P, C.c_uint32]),
    "stc_get_schema": (C.c_int, [P, P, C.c_uint64, U64P]),
    "stc_last_error": (C.c_char_p, [P]),
    "stc_step": (C.c_int, [P, P, C.c_uint32, C.POINTER(StcStepOut)]),
    "stc_step_device": (C.c_int, [P, P, C.c_uint32, C.POINTER(StcStepOut)]),
    "stc_step_batch": (C.c_int, [P, C.c_uint32, P, P, C.c_uint32, C.c_uint32, C.POINTER(StcStepOut), P]),
    "stc_get_pairs": (C.c_int, [P, P, C.c_uint64, U64P]),
    "stc_get_step_meetings": (C.c_int, [P, P, C.c_uint64, U64P]),
    "stc_step_meetings": (C.c_int, [P, P, C.c_uint64, C.c_uint32, C.POINTER(StcStepOut)]),
    "stc_get_active": (C.c_int, [P, P, P, P, P, C.c_uint64, U64P]),
    "stc_get_records": (C.c_int, [P, P, P, C.c_uint64, U64P]),
    "stc_get_leaves": (C.c_int, [P, P, P, C.c_uint64, U64P]),
    "stc_num_runs": (C.c_int, [P, C.POINTER(C.c_uint32)]),
    "stc_get_run_info": (C.c_int, [P, C.c_uint32, C.POINTER(StcRunInfo)]),
    "stc_get_run": (C.c_int, [P, C.c_uint32, P, P, P, P, P, P, P, P, P]),
    "stc_drop_runs": (C.c_int, [P]),
    "stc_compact_runs": (C.c_int, [P, C.c_uint32, C.c_uint32]),
    "stc_sync": (C.c_int, [P]),
    "stc_ctf_widths": (C.c_int, [C.POINTER(StcCtfMeta), C.POINTER(StcBox), C.c_uint32, C.c_uint32]),
    "stc_live_search": (C.c_int, [P, P, C.c_uint32, P, C.c_uint64, U64P]),
    "stc_live_search_noalloc": (C.c_int, [P, P, C.c_uint32, P, C.c_uint64, U64P]),
    "stc_dump_run": (C.c_int, [P, C.c_uint32, C.c_char_p, C.c_char_p, C.c_uint32]),
    "stc_query_batch": (C.c_int, [P, P, C.c_uint32, P, C.c_uint64, U64P, P]),
    "stc_query_batch_at": (C.c_int, [P, P, C.c_uint32, P, C.c_uint64, U64P, P, C.POINTER(C.c_uint32)]),
    "stc_gen_init_host": (C.c_int, [C.POINTER(StcGenParams), P, C.c_uint32, C.c_uint32]),
    "stc_gen_step_host": (C.c_int, [C.POINTER(StcGenParams), P, C.c_uint32, C.c_uint32]),
    "stc_gen_pos_host": (C.c_int, [C.POINTER(StcGenParams), P, C.c_uint32, P]),
    "stc_trace_create": (C.c_int, [C.POINTER(StcGenParams), C.POINTER(C.c_void_p)]),
    "stc_trace_next": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_void_p)]),
    "stc_trace_destroy": (C.c_int, [P]),
}

_lib = None


def load():
    """Load libstcdb_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python stcdb_amd/build.py` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class StcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"stcdb error {code}: {msg}")
        self.code = code
