"""Leaf-capacity tuner for batched streams.

A stream's leaf capacity (the reference's grid_capacity, config.h) sets the
QTree's leaf size and with it how the step's work divides between the border
lookup, the pair enumeration and the per-leaf overheads; the best value
depends on the stream's density and reach. The reference takes it from the
command line; here each stream is timed at a few candidate capacities on its
own first frames and the fastest is kept. The chosen capacity is part of the
stream's configuration from then on: results are those of the reference run
with that capacity (the pair decisions depend on the tree through the probe
point of each unit, geometry.cu:258-392, so runs at different capacities are
not compared with each other).

The design has no meeting-bucket count to tune (the meetings are sorted
arrays, not a hash table) and its tile shapes are fixed by each kernel's
register budget, so the leaf capacity is the one knob.
"""
import statistics

from .engine import Pipeline

CANDIDATES = (100, 200, 400, 800, 1600)


def tune_leaf_capacity(cfg, box, frames, candidates=CANDIDATES, on_device=False):
    """Step a scratch context per candidate capacity over `frames` (host
    arrays, or device pointers with on_device) and return (best capacity,
    {capacity: median ms per step after the first}). Ties go to the smaller
    capacity. Needs at least two frames."""
    assert len(frames) >= 2
    timings = {}
    for cap in candidates:
        p = Pipeline(dict(cfg, grid_capacity=int(cap)), box)
        try:
            ms = []
            for t, f in enumerate(frames):
                r = p.step_device(f, t) if on_device else p.step(f, t)
                if t:
                    ms.append(r["ms_total"])
        finally:
            p.close()
        timings[int(cap)] = statistics.median(ms)
    best = min(timings, key=lambda c: (timings[c], c))
    return best, timings
