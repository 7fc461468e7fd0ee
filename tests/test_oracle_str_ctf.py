"""Pins the flush's STR sid assignment and the CTF bit widths with independent
restatements.

STR (set_oid, narrow_xy_to_y, make_sid: src/cuda/geometry.cu:739-806, sorts at
:1858-1905): each object with records and an MBR area below the oversize
threshold gets the key x = lo0/2 + hi0/2 over its quantised box, y = hi1/2 +
hi1/2 (high[1] twice, as the reference writes it); objects are ranked by a
stable sort on x << 32 | y from oid order, cut into split_num strips of
(M / split_num + 1) objects, ranked within a strip by a stable sort on y, and
cut into cells of (strip capacity / split_num + 1). sid = 2 + strip *
split_num + cell. Here that is recomputed in numpy from the run's own sorted
records (the quantised box of an object is the union of float_to_uint of its
records' MBRs, cuda_util.cuh:115-119, geometry.cu:116-121, 729-733), and the
test asserts that the y key with low[1] would have given different sids.

CTF widths (CTF::get_ctf_bits, src/tracing/step_merge.cpp:140-186): the grid
counts low_x_grid = floor(width / (map width / 3000)) and low_y_grid (4200)
are computed in exact rational arithmetic, the widths with int.bit_length
(min_bits_to_store, util.h:527-533), the bitmap grid from 300 x 420 cells of
the map clamped to [32, 255] (and to 255 only when either side reaches 256).
Crafted CTF boxes sit just above powers of two and at the clamps, so a wrong
divisor or clamp changes a width. The product's host routine
(stc_ctf_widths) is checked with the same cases as the oracle's.
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import HostTrace, workload

OID_MASK = (1 << 26) - 1


def min_bits(v):
    return 1 if v == 0 else int(v).bit_length()


def expected_ctf(mbr, map4, n, max_meet):
    """(low_x_bit, low_y_bit, edge_bit, mbr_bit, key_bit, x_grid, y_grid, bitmap_size), or None
    where a grid count lies within 1e-6 of an integer (double rounding could go either way)"""
    m = np.float32(mbr)
    wx = Fraction(float(np.float32(m[2] - m[0])))
    wy = Fraction(float(np.float32(m[3] - m[1])))
    W, H = Fraction(map4[2]) - Fraction(map4[0]), Fraction(map4[3]) - Fraction(map4[1])
    vals = [wx * 3000 / W, wy * 4200 / H, wx / W * 300, wy / H * 420]
    for v in vals:
        if abs(v - round(v)) < Fraction(1, 10 ** 6) and v > 0:
            return None
    lxg, lyg, ow, oh = (math.floor(v) for v in vals)
    lxb = min_bits(max(lxg, 1) - 1)
    lyb = min_bits(max(lyg, 1) - 1)
    eb = max(lxb, lyb)
    mb = lxb + lyb + 2 * eb
    dur = min_bits(max_meet - 1)
    temp = 2 * min_bits(n) + 2 * dur + mb
    key_bit = -(-temp // 8) * 8
    xg, yg = max(ow + 1, 32), max(oh + 1, 32)
    if xg >= 256 or yg >= 256:
        xg, yg = min(ow + 1, 255), min(oh + 1, 255)
    return (lxb, lyb, eb, mb, key_bit, xg, yg, (xg * yg + 7) // 8)


def crafted_boxes(map4):
    """CTF boxes whose grid counts sit 0.1 above 2^k + 1 (x: 3000 cells, y: 4200)
    and whose bitmap grids sit at 31, 32, 101 and 256 cells"""
    W, H = map4[2] - map4[0], map4[3] - map4[1]
    x0, y0 = map4[0] + 0.1 * W, map4[1] + 0.05 * H
    out = []
    for kx, ky in [(10, 10), (11, 3), (3, 11), (10, 11), (8, 8), (0, 0)]:
        gx, gy = (2 ** kx + 1.1), (2 ** ky + 1.1)
        out.append([x0, y0, x0 + gx * W / 3000, y0 + gy * H / 4200])
    for cells_x, cells_y in [(30.5, 30.5), (31.5, 100.1), (100.1, 31.3), (255.5, 20.5), (20.5, 255.5),
                             (256.6, 300.2)]:
        out.append([x0 - 0.05 * W, y0, x0 - 0.05 * W + cells_x * W / 300, y0 + cells_y * H / 420])
    return np.float32(out)


MAPS = [
    [-87.94068, 41.64413, -87.81, 41.734],          # the c1 map
    [-88.1, 41.5, -87.5, 42.05],
    [2.1, 48.7, 2.6, 49.0],
]


def product_widths(mbr, map4, n, max_meet):
    import stcdb_amd._lib as L
    lib = L.load()
    c = L.StcCtfMeta()
    for i in range(4):
        c.mbr[i] = float(mbr[i])
    b = L.StcBox()
    b.low[0], b.low[1], b.high[0], b.high[1] = map4
    assert lib.stc_ctf_widths(ctypes.byref(c), ctypes.byref(b), n, max_meet) == 0
    return (c.low_x_bit, c.low_y_bit, c.edge_bit, c.mbr_bit, c.key_bit, c.x_grid, c.y_grid, c.bitmap_size)


def oracle_widths(mbr, map4, n, max_meet):
    lib = O.load()
    c = np.zeros(1, dtype=O.CTF_DTYPE)
    c["mbr"][0] = mbr
    m = np.array(map4, dtype=np.float64)
    lib.orc_ctf_bits(c.ctypes.data_as(ctypes.c_void_p), m.ctypes.data_as(ctypes.c_void_p), n, max_meet)
    c = c[0]
    return tuple(int(c[k]) for k in ["low_x_bit", "low_y_bit", "edge_bit", "mbr_bit", "key_bit", "x_grid",
                                      "y_grid", "bitmap_size"])


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_ctf_widths_crafted_and_random(impl):
    f = oracle_widths if impl == "oracle" else product_widths
    rng = np.random.default_rng(7)
    checked = 0
    for map4 in MAPS:
        W, H = map4[2] - map4[0], map4[3] - map4[1]
        boxes = list(crafted_boxes(map4))
        for _ in range(300):
            x0, y0 = map4[0] + rng.uniform(0, 0.5) * W, map4[1] + rng.uniform(0, 0.5) * H
            boxes.append(np.float32([x0, y0, x0 + rng.uniform(0, 0.5) ** 2 * W, y0 + rng.uniform(0, 0.5) ** 2 * H]))
        for mbr in boxes:
            for n, mm in [(10_000, 4096), (10_000_000, 4096), (600_000, 100)]:
                e = expected_ctf(mbr, map4, n, mm)
                if e is None:
                    continue
                assert f(mbr, map4, n, mm) == e, (map4, mbr.tolist(), n, mm)
                checked += 1
    assert checked > 2500


def test_crafted_boxes_exercise_each_rule():
    """the crafted boxes cross the thresholds a wrong constant would move"""
    map4 = MAPS[0]
    got = [expected_ctf(b, map4, 10_000, 4096) for b in crafted_boxes(map4)]
    assert all(g is not None for g in got)
    lx = {g[0] for g in got}
    assert {11, 12} <= lx  # 2^10 + 1 and 2^11 + 1 grid cells need 11 and 12 bits
    xg = [g[5] for g in got[6:]]
    yg = [g[6] for g in got[6:]]
    assert 32 in xg and 101 in yg and 255 in xg and 255 in yg and 21 in yg and 21 in xg


def quantise(a):
    """float_to_uint (cuda_util.cuh:115-119) in float32 arithmetic"""
    v = (np.asarray(a, dtype=np.float32) + np.float32(180)) * np.float32(100000)
    assert (v > 0).all() and (v < 2 ** 32).all()
    return v.astype(np.uint64)


def str_sids(n, split, keys_hi, boxes, oversize, y_low=False):
    """sids of the objects with records, restated with numpy sorts"""
    pid = ((keys_hi >> np.uint64(26)) & np.uint64(OID_MASK)).astype(np.int64)
    q = quantise(boxes)
    big = np.uint64(1 << 62)
    lo0 = np.full(n, big, np.uint64); lo1 = np.full(n, big, np.uint64)
    hi0 = np.zeros(n, np.uint64); hi1 = np.zeros(n, np.uint64)
    np.minimum.at(lo0, pid, q[:, 0]); np.minimum.at(lo1, pid, q[:, 1])
    np.maximum.at(hi0, pid, q[:, 2]); np.maximum.at(hi1, pid, q[:, 3])
    objs = np.array(sorted(set(pid.tolist()) - set(oversize)), dtype=np.int64)
    xk = lo0[objs] // 2 + hi0[objs] // 2
    yk = (lo1[objs] // 2 + hi1[objs] // 2) if y_low else (hi1[objs] // 2 + hi1[objs] // 2)
    m = len(objs)
    xpc = m // split + 1
    pkc = xpc // split + 1
    xr = np.empty(m, np.int64)
    # the strip sort orders the whole 64-bit key x << 32 | y (sort_by_key(mid_xys, oids))
    xr[np.lexsort((objs, yk, xk))] = np.arange(m)
    strip = xr // xpc
    sid = np.empty(m, np.int64)
    for s in range(split):
        sel = np.nonzero(strip == s)[0]
        order = sel[np.lexsort((xr[sel], yk[sel]))]
        sid[order] = 2 + s * split + np.arange(len(order)) // pkc
    return objs, sid


@pytest.mark.parametrize("ctf_count", [16, 64])
def test_str_sids_from_records(ctf_count):
    n = 10_000
    w = workload("c1", num_objects=n, reach_distance=30.0, min_meet_time=2, kv_restriction=2000,
                 ctf_count=ctf_count)
    orc = O.Oracle(w["cfg"], w["box"])
    tr = HostTrace(w["gen"])
    for t, xy in tr.steps(100):
        orc.step(xy, t)
        if orc.num_runs():
            break
    assert orc.num_runs() >= 1
    run = orc.run(0)
    info = run["info"]
    split = int(round(math.sqrt(ctf_count)))
    oversize = np.nonzero(run["sids"] == 1)[0].tolist()
    objs, sid = str_sids(n, split, run["sorted_keys"][:, 1], run["sorted_boxes"], oversize)
    assert len(objs) == info["sid_count"] - info["oversize_oid_count"] > 300
    assert np.array_equal(run["sids"][objs].astype(np.int64), sid)
    # objects without records carry sid 0
    rest = np.setdiff1d(np.arange(n), np.concatenate([objs, np.array(oversize, dtype=np.int64)]))
    assert (run["sids"][rest] == 0).all()
    # the record keys carry the sid of their pid at bit 116 (bit 52 of the high word)
    pid = ((run["sorted_keys"][:, 1] >> np.uint64(26)) & np.uint64(OID_MASK)).astype(np.int64)
    assert np.array_equal((run["sorted_keys"][:, 1] >> np.uint64(52)).astype(np.int64), run["sids"][pid].astype(np.int64))
    # the y key with low[1] would place objects in other CTFs: the case is exercised
    _, sid_low = str_sids(n, split, run["sorted_keys"][:, 1], run["sorted_boxes"], oversize, y_low=True)
    assert (sid_low != sid).sum() > 0
