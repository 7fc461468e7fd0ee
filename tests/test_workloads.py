"""The seeded workload definitions (host only)."""
import numpy as np

from stcdb_amd import batched_streams, workload


def test_batched_streams_shape():
    s = batched_streams()
    n = np.array([w["cfg"]["num_objects"] for w in s])
    assert len(s) == 64
    assert n.min() >= 1_000 and n.max() <= 2_000_000
    assert 2e6 < n.sum() < 1.2e7 and 25_000 < np.median(n) < 100_000
    for w in s:
        side_km = w["gen"]["grid_cells"] * w["gen"]["spacing"] / 1000.0
        assert 10.0 <= side_km <= 50.0
        assert 5.0 <= w["cfg"]["reach_distance"] <= 25.0
        assert w["cfg"]["grid_capacity"] in (50, 100, 200)
        assert w["gen"]["num_objects"] == w["cfg"]["num_objects"]
    # independent seeds per stream, and the list depends only on the arguments
    assert len({w["gen"]["seed"] for w in s}) == 64
    assert [w["cfg"] for w in batched_streams()] == [w["cfg"] for w in s]
    assert [w["cfg"] for w in batched_streams(seed=5)] != [w["cfg"] for w in s]


def test_workload_overrides_rescale_generator():
    w = workload("c1", num_objects=1234, reach_distance=7.0)
    assert w["gen"]["num_objects"] == 1234 and w["cfg"]["num_objects"] == 1234
    assert w["cfg"]["reach_distance"] == 7.0 and w["steps"] == 120
