"""Pins the oracle against committed vectors of the reference's own helpers
(tests/golden/ref_vectors.npz, written by tests/golden/make_ref_vectors.py from
the reference code compiled into oracle/_ref). Runs without the reference tree."""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


def test_dpm_lon_golden(g):
    got = np.array([O.dpm_lon(v) for v in g["lat"]])
    assert np.array_equal(got.view(np.uint64), g["dpm_lon"].view(np.uint64))


def test_point_distance_golden(g):
    got = np.array([O.point_distance(*p) for p in g["pts"]])
    assert np.array_equal(got.view(np.uint64), g["dist"].view(np.uint64))


def test_box_distance_golden(g):
    got = np.array([O.box_distance(np.ascontiguousarray(b), px, py) for b, (px, py) in zip(g["boxes"], g["probes"])])
    assert np.array_equal(got.view(np.uint64), g["box_dist"].view(np.uint64))


def test_min_bits_golden(g):
    lib = O.load()
    got = np.array([lib.orc_min_bits(int(v)) for v in g["mb_in"]], dtype=np.uint32)
    assert np.array_equal(got, g["mb_out"])


def test_conversions_golden(g):
    lib = O.load()
    assert np.array_equal(np.array([lib.orc_conv_d2u32(v) for v in g["conv_in"]], dtype=np.uint32), g["conv_u"])
    assert np.array_equal(np.array([lib.orc_conv_d2i32(v) for v in g["conv_in"]], dtype=np.int32), g["conv_i"])


def test_key_fields_golden(g):
    # the 128-bit record key layout the oracle and the device write: sid | pid | target | duration | end
    for (lo, hi), f in zip(g["keys"], g["fields"]):
        k = (int(hi) << 64) | int(lo)
        assert [k >> 116, (k >> 90) & (2**26 - 1), (k >> 64) & (2**26 - 1), (k >> 32) & (2**32 - 1),
                k & (2**32 - 1)] == list(f)
