"""Host-side logic without a GPU: the leaf-capacity tuner's choice, the bench
command line's refusal of more than one GPU, and the batched-streams draw."""
import subprocess
import sys
import os

import pytest

import stcdb_amd.tuner as T
from stcdb_amd import batched_streams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class FakePipeline:
    """stands in for a context: a step takes the time the table gives for its capacity"""
    times = {}
    made = []

    def __init__(self, cfg, box):
        self.cap = cfg["grid_capacity"]
        FakePipeline.made.append(self.cap)

    def step(self, xy, t):
        return {"ms_total": FakePipeline.times[self.cap][t]}

    def close(self):
        pass


def test_tuner_picks_the_fastest_median_and_the_smaller_on_ties(monkeypatch):
    monkeypatch.setattr(T, "Pipeline", FakePipeline)
    FakePipeline.made = []
    # the first step (which builds the tree) is not counted; medians 2.0, 1.5, 1.5, 3.0
    FakePipeline.times = {50: [9, 2.0, 2.0, 1.0], 100: [9, 1.5, 1.4, 1.6], 200: [0.1, 1.5, 1.5, 1.5],
                          400: [9, 3.0, 3.0, 3.0]}
    best, timings = T.tune_leaf_capacity({"grid_capacity": 10}, None, [None] * 4, candidates=(50, 100, 200, 400))
    assert FakePipeline.made == [50, 100, 200, 400]
    assert timings == {50: 2.0, 100: 1.5, 200: 1.5, 400: 3.0}
    assert best == 100
    with pytest.raises(AssertionError):
        T.tune_leaf_capacity({"grid_capacity": 10}, None, [None])


def test_batched_streams_are_seeded_and_within_bounds():
    a = batched_streams(16, steps=3, seed=7)
    b = batched_streams(16, steps=3, seed=7)
    assert [w["cfg"] for w in a] == [w["cfg"] for w in b]
    for w in a:
        assert 1_000 <= w["cfg"]["num_objects"] <= 2_000_000
        assert 5.0 <= w["cfg"]["reach_distance"] <= 25.0
        assert w["cfg"]["grid_capacity"] in (50, 100, 200)
        assert w["cfg"]["dynamic_schema"] == 1
        side_km = w["gen"]["grid_cells"] * w["gen"]["spacing"] / 1000.0
        assert 9.9 <= side_km <= 50.1


def test_bench_rejects_a_single_tuner_frame():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "c4", "--tune-steps", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "--tune-steps must be at least 2" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one MI355X" in (r.stdout + r.stderr)
