"""Flushed runs on disk, in the reference's file layout.

The reference writes each run (a CTB) with straight_dump
(src/tracing/process.cpp:254-348) and reads it back with load_CTB_meta /
load_CTF_meta (src/tracing/workbench.cpp:653-694):

  <prefix>_SSTable_<ctb>-<c>  the packed keys of CTF c (CTF::dump_keys, step_merge.cpp:652-658)
  <prefix>_CTB<ctb>           the CTB struct image, then num_objects u16 sids, the oversize
                              buffer's keys (16 B each), boxes (f_box, 16 B each) and its
                              8 KB bitmap (workbench::dump_CTB_meta, workbench.cpp:597-614)
  <prefix>_STcL<ctb>-<c>      the CTF struct image, then ctf_bitmap_size bitmap bytes
                              (CTF::dump_meta, step_merge.cpp:644-650)

The struct images have the x86-64 layout of the reference classes (CTF: 80
bytes, step_merge.h:64-109; CTB: 128 bytes, step_merge.h:111-131). Their
pointer fields are written as zero. CTB fields that the reference's GPU path
leaves unset are written with the run's values: the CTF count, the start-time
bounds over its CTFs, and the run's end-time base and flush time.

stc_dump_run (include/stcdb.h) writes the same files natively; this module
writes and reads them from the arrays stc_get_run returns.
"""
import os

import numpy as np

CTF_IMAGE = np.dtype([
    ("mbr", "<f4", (4,)), ("keys", "<u8"), ("bitmap", "<u8"), ("kv_capacity", "<u4"),
    ("start_time_min", "<u4"), ("start_time_max", "<u4"), ("end_time_min", "<u4"), ("end_time_max", "<u4"),
    ("bitmap_size", "<u4"), ("key_bit", "<u2"), ("id_bit", "<u2"), ("duration_bit", "<u2"), ("end_bit", "<u2"),
    ("low_x_bit", "<u2"), ("low_y_bit", "<u2"), ("edge_bit", "<u2"), ("mbr_bit", "<u2"), ("x_grid", "<u2"),
    ("y_grid", "<u2"), ("pad", "<u4"),
])
CTB_IMAGE = np.dtype([
    ("ctfs", "<u8"), ("ctf_count", "<u4"), ("start_time_min", "<u4"), ("start_time_max", "<u4"),
    ("end_time_min", "<u4"), ("end_time_max", "<u4"), ("pad0", "<u4"), ("first_widpid", "<u8"), ("sids", "<u8"),
    ("bitmaps", "<u8"), ("bitmap_mbrs", "<u8"), ("ctf_capacity", "<u8"),
    ("o_count", "<u4"), ("pad1", "<u4"), ("o_keys", "<u8"), ("o_boxes", "<u8"), ("o_bitmaps", "<u8"),
    ("o_start_time_min", "<u4"), ("o_start_time_max", "<u4"), ("o_end_time_min", "<u4"), ("o_end_time_max", "<u4"),
    ("box_rtree", "<u8"),
])
assert CTF_IMAGE.itemsize == 80 and CTB_IMAGE.itemsize == 128

CTF_FIELDS = ["kv_capacity", "start_time_min", "start_time_max", "end_time_min", "end_time_max", "bitmap_size",
              "key_bit", "id_bit", "duration_bit", "end_bit", "low_x_bit", "low_y_bit", "edge_bit", "mbr_bit",
              "x_grid", "y_grid"]
BITMAP_BYTES = 8192


def paths(directory, prefix, ctb, ctf_count):
    sst = [os.path.join(directory, f"{prefix}_SSTable_{ctb}-{c}") for c in range(ctf_count)]
    meta = [os.path.join(directory, f"{prefix}_STcL{ctb}-{c}") for c in range(ctf_count)]
    return sst, os.path.join(directory, f"{prefix}_CTB{ctb}"), meta


def dump_run(run, directory, prefix, ctb):
    """write one run (the dict of Pipeline.run / the oracle's run) as files"""
    ctfs, info = run["ctfs"], run["info"]
    C = len(ctfs)
    sst, ctb_path, meta = paths(directory, prefix, ctb, C)
    off = 0
    for c in range(C):
        nbytes = int(ctfs[c]["kv_capacity"]) * (int(ctfs[c]["key_bit"]) // 8)
        with open(sst[c], "wb") as f:
            f.write(run["key_bytes"][off:off + nbytes].tobytes())
        off += nbytes
        img = np.zeros(1, dtype=CTF_IMAGE)
        img["mbr"] = ctfs[c]["mbr"]
        for k in CTF_FIELDS:
            img[k] = ctfs[c][k]
        bsz = int(ctfs[c]["bitmap_size"])
        with open(meta[c], "wb") as f:
            f.write(img.tobytes())
            f.write(run["bitmaps"][c * BITMAP_BYTES:c * BITMAP_BYTES + bsz].tobytes())
    img = np.zeros(1, dtype=CTB_IMAGE)
    img["ctf_count"] = C
    img["start_time_min"] = int(ctfs["start_time_min"].min()) if C else 0
    img["start_time_max"] = int(ctfs["start_time_max"].max()) if C else 0
    img["end_time_min"] = info["end_time_min"]
    img["end_time_max"] = info["cur_time"]
    img["o_count"] = info["oversize_count"]
    img["o_start_time_min"] = info["o_start_min"]
    img["o_start_time_max"] = info["o_start_max"]
    img["o_end_time_min"] = info["o_end_min"]
    img["o_end_time_max"] = info["o_end_max"]
    with open(ctb_path, "wb") as f:
        f.write(img.tobytes())
        f.write(np.ascontiguousarray(run["sids"], dtype="<u2").tobytes())
        f.write(np.ascontiguousarray(run["o_keys"], dtype="<u8").tobytes())
        f.write(np.ascontiguousarray(run["o_boxes"], dtype="<f4").tobytes())
        f.write(np.ascontiguousarray(run["o_bitmap"], dtype=np.uint8).tobytes())


def load_run(directory, prefix, ctb, num_objects, ctf_count):
    """read one run back: load_CTB_meta + load_CTF_meta + the SSTables"""
    sst, ctb_path, meta = paths(directory, prefix, ctb, ctf_count)
    raw = np.fromfile(ctb_path, dtype=np.uint8)
    img = raw[:128].view(CTB_IMAGE)[0]
    p = 128
    sids = raw[p:p + 2 * num_objects].view("<u2").copy()
    p += 2 * num_objects
    no = int(img["o_count"])
    o_keys = raw[p:p + 16 * no].view("<u8").reshape(no, 2).copy()
    p += 16 * no
    o_boxes = raw[p:p + 16 * no].view("<f4").reshape(no, 4).copy()
    p += 16 * no
    o_bitmap = raw[p:p + BITMAP_BYTES].copy()
    ctfs = []
    bitmaps = np.zeros(ctf_count * BITMAP_BYTES, dtype=np.uint8)
    keys = []
    for c in range(ctf_count):
        m = np.fromfile(meta[c], dtype=np.uint8)
        ci = m[:80].view(CTF_IMAGE)[0]
        bsz = int(ci["bitmap_size"])
        bitmaps[c * BITMAP_BYTES:c * BITMAP_BYTES + bsz] = m[80:80 + bsz]
        ctfs.append(ci)
        keys.append(np.fromfile(sst[c], dtype=np.uint8))
    return dict(ctb=img, ctfs=np.array(ctfs, dtype=CTF_IMAGE), key_bytes=np.concatenate(keys) if keys else
                np.zeros(0, np.uint8), bitmaps=bitmaps, sids=sids, o_keys=o_keys, o_boxes=o_boxes, o_bitmap=o_bitmap)
