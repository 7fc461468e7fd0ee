"""Trace chunks in the reference's .tr format round-trip (tracer::loadData,
src/tracing/process.cpp:128-144: 100 steps x num_objects x {double x, y})."""
import os

import numpy as np

from stcdb_amd import traces, workload
from stcdb_amd.engine import HostTrace


def test_trace_chunks_round_trip(tmp_path):
    w = workload("c1", num_objects=500)
    prefix = str(tmp_path / "trace_")
    paths = traces.write_trace(prefix, HostTrace(w["gen"]), 230)
    assert [os.path.basename(p) for p in paths] == ["trace_0_500.tr", "trace_100_500.tr", "trace_200_500.tr"]
    assert os.path.getsize(paths[0]) == 100 * 500 * 16 and os.path.getsize(paths[2]) == 30 * 500 * 16
    ref = [xy.copy() for _, xy in HostTrace(w["gen"]).steps(230)]
    got = list(traces.iter_trace(prefix, 500, 230))
    assert [t for t, _ in got] == list(range(230))
    for (t, xy), r in zip(got, ref):
        assert np.array_equal(xy.view(np.uint64), r.view(np.uint64)), t
