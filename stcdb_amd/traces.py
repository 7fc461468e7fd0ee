"""Trace chunks in the reference's file format.

tracer::loadData (src/tracing/process.cpp:128-144) reads
<path><st>_<num_objects>.tr: 100 timesteps of num_objects points, each point
the reference's Point {double x, y} (src/geometry/geometry.h:23-56) in
degrees, timestep after timestep. write_chunk writes the seeded generator's
positions in that format, so a reference-style run and this library consume
identical bytes; read_chunk returns the (steps, num_objects, 2) float64 array
that Pipeline.step takes one timestep at a time.
"""
import numpy as np

STEPS_PER_CHUNK = 100


def chunk_path(prefix, st, num_objects):
    return f"{prefix}{st}_{num_objects}.tr"


def write_chunk(prefix, st, frames):
    """frames: (steps <= 100, num_objects, 2) float64 positions starting at timestep st"""
    frames = np.ascontiguousarray(frames, dtype="<f8")
    assert frames.ndim == 3 and frames.shape[2] == 2 and frames.shape[0] <= STEPS_PER_CHUNK
    path = chunk_path(prefix, st, frames.shape[1])
    with open(path, "wb") as f:
        f.write(frames.tobytes())
    return path


def read_chunk(prefix, st, num_objects):
    """the chunk starting at timestep st; a short final chunk holds fewer steps"""
    path = chunk_path(prefix, st, num_objects)
    raw = np.fromfile(path, dtype="<f8")
    if raw.size % (2 * num_objects):
        raise ValueError(f"{path}: {raw.size * 8} bytes is not a whole number of {num_objects}-point steps")
    return raw.reshape(-1, num_objects, 2)


def write_trace(prefix, trace, steps):
    """write `steps` timesteps of a HostTrace as consecutive 100-step chunks; returns the paths"""
    out, buf, st = [], [], 0
    for t, xy in trace.steps(steps):
        buf.append(xy.copy())
        if len(buf) == STEPS_PER_CHUNK:
            out.append(write_chunk(prefix, st, np.stack(buf)))
            st, buf = t + 1, []
    if buf:
        out.append(write_chunk(prefix, st, np.stack(buf)))
    return out


def iter_trace(prefix, num_objects, steps):
    """yield (t, xy) for t < steps from consecutive chunk files"""
    t = 0
    while t < steps:
        chunk = read_chunk(prefix, t, num_objects)
        for xy in chunk:
            if t >= steps:
                return
            yield t, xy
            t += 1
        if len(chunk) < STEPS_PER_CHUNK and t < steps:
            raise EOFError(f"trace ends at step {t}")
