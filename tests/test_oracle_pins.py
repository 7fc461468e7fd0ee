"""Pins the oracle's arithmetic against the reference's own code (compiled from
/root/reference into oracle/_ref by oracle/Makefile) and against closed forms."""
import math
import struct

import numpy as np
import pytest

import oracle_lib as O

ref = O.load_ref()
needs_ref = pytest.mark.skipif(ref is None, reason="reference shim not built (reference tree absent)")


def rng(seed=0):
    return np.random.default_rng(seed)


@needs_ref
def test_dpm_lon_matches_reference_bit_exact():
    r = rng(1)
    lats = np.concatenate([r.uniform(-89.9, 89.9, 20000), r.uniform(41.5, 42.5, 20000), [0.0, 41.644584027118924]])
    for lat in lats:
        assert O.dpm_lon(lat) == ref.ref_dpm_lon(lat)


@needs_ref
def test_point_distance_matches_reference_bit_exact():
    r = rng(2)
    x = r.uniform(-88.0, -87.5, (20000, 2))
    y = r.uniform(41.6, 42.1, (20000, 2))
    # include near pairs (metres apart) as in the pipeline
    x[:10000, 1] = x[:10000, 0] + r.normal(0, 1e-4, 10000)
    y[:10000, 1] = y[:10000, 0] + r.normal(0, 1e-4, 10000)
    for i in range(x.shape[0]):
        a = O.point_distance(x[i, 0], y[i, 0], x[i, 1], y[i, 1])
        b = ref.ref_point_distance_geo(x[i, 0], y[i, 0], x[i, 1], y[i, 1])
        assert a == b


@needs_ref
def test_box_distance_matches_reference_outside_boxes():
    # the oracle uses the device form (no containment shortcut); outside the
    # box both forms evaluate the same expression
    r = rng(3)
    for _ in range(20000):
        lo = np.array([r.uniform(-88, -87.6), r.uniform(41.6, 42.0)])
        hi = lo + r.uniform(1e-5, 1e-2, 2)
        px, py = r.uniform(lo[0] - 0.01, hi[0] + 0.01), r.uniform(lo[1] - 0.01, hi[1] + 0.01)
        box = np.array([lo[0], lo[1], hi[0], hi[1]])
        a = O.box_distance(box, px, py)
        b = ref.ref_box_distance_geo(box.ctypes.data, px, py)
        inside = lo[0] <= px <= hi[0] and lo[1] <= py <= hi[1]
        if inside:
            assert b == 0.0 and a < 1e-6
        else:
            assert a == b


@needs_ref
def test_min_bits_matches_reference():
    vals = [0, 1, 2, 3, 4, 7, 8, 4095, 4096, 9999, 10000, 10001, 999999, 10**6, 10**7, 2**24 - 1, 2**24,
            2**31, 2**32 - 1]
    vals += list(rng(4).integers(0, 2**32, 5000))
    for v in vals:
        assert O.load().orc_min_bits(int(v)) == ref.ref_min_bits(int(v))


@needs_ref
def test_host_conversions_match_compiler():
    vals = [0.0, 0.5, 1.0, 1.9, -0.5, -1.0, -5.0e9, -3.3, 4294967295.0, 4294967296.0, 5e9, 1e19, -1e19,
            float("nan"), 2147483647.0, 2147483648.0, -2147483649.0, 3e9]
    vals += list(rng(5).uniform(-1e10, 1e10, 2000))
    lib = O.load()
    for v in vals:
        assert lib.orc_conv_d2u32(v) == ref.ref_conv_d2u(v), v
        assert lib.orc_conv_d2i32(v) == ref.ref_conv_d2i(v), v


@needs_ref
def test_key_layout_matches_reference_extractors():
    r = rng(6)
    out = (np.zeros(5, dtype=np.uint32))
    for _ in range(2000):
        sid, pid, tgt = int(r.integers(0, 4096)), int(r.integers(0, 2**26)), int(r.integers(0, 2**26))
        dur, end = int(r.integers(0, 2**32)), int(r.integers(0, 2**32))
        k = (sid << 116) | (pid << 90) | (tgt << 64) | (dur << 32) | end
        ref.ref_key_fields(k & (2**64 - 1), k >> 64, out.ctypes.data)
        assert list(out) == [sid, pid, tgt, dur, end]


@needs_ref
def test_cantor_pairing_is_a_bijection_on_ordered_pairs():
    # the reference keys its meeting table by Cantor(min, max); this build keys
    # by (oid1 << 32 | oid2) — both are injective on ordered pairs
    r = rng(7)
    a_out = np.zeros(1, dtype=np.uint32)
    b_out = np.zeros(1, dtype=np.uint32)
    for _ in range(3000):
        a, b = sorted(int(v) for v in r.integers(0, 10**7, 2))
        z = ref.ref_cantor(a, b)
        assert z == (a + b) * (a + b + 1) // 2 + b
        ref.ref_inverse_cantor(z, a_out.ctypes.data, b_out.ctypes.data)
        assert (int(a_out[0]), int(b_out[0])) == (a, b)


def test_dpm_lon_closed_form():
    # at the equator one degree of longitude is 40076 km / 360
    assert O.dpm_lon(0.0) == 360.0 / (math.sin(90 * 3.14159265 / 180) * 40076.0 * 1000.0)
    # monotone in |lat| and symmetric
    assert O.dpm_lon(41.6) < O.dpm_lon(41.7)
    assert O.dpm_lon(-41.6) == O.dpm_lon(41.6)


def test_point_distance_closed_form():
    # along a meridian: metres = dlat_degrees / dpm_lat
    dpm_lat = 360.0 / (40076.0 * 1000.0)
    d = O.point_distance(-87.9, 41.7, -87.9, 41.7 + 10 * dpm_lat)
    assert abs(d - 10.0) < 1e-9
    # along a parallel, scaled by the probe's latitude
    dl = O.dpm_lon(41.7)
    d = O.point_distance(-87.9 + 7 * dl, 41.7, -87.9, 41.7)
    assert abs(d - 7.0) < 1e-9
