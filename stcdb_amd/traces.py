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


def meetings_path(prefix, st, num_objects):
    """workbench::dump_meetings names the file <trace_path>meeting<st>_<num_objects>.tr"""
    return f"{prefix}meeting{st}_{num_objects}.tr"


def write_meetings(prefix, st, num_objects, steps):
    """Per-second meeting units of up to 100 timesteps starting at st, in the
    layout of workbench::dump_meetings (src/tracing/workbench.cpp:446-455): 100
    u32 counts indexed by timestep % 100, then the 32-byte units of each
    timestep, blocks in that index order. steps[i] holds the units of
    timestep st + i (MEETING_UNIT_DTYPE, as Pipeline.step_meeting_units
    returns them). The reference's own writer never advances its offset
    (geometry.cu:1720 writes every second at h_meetings_count, which stays 0);
    this writes the layout its reader expects."""
    from .engine import MEETING_UNIT_DTYPE
    assert len(steps) <= STEPS_PER_CHUNK
    counts = np.zeros(STEPS_PER_CHUNK, dtype="<u4")
    blocks = [np.zeros(0, dtype=MEETING_UNIT_DTYPE)] * STEPS_PER_CHUNK
    for i, u in enumerate(steps):
        k = (st + i) % STEPS_PER_CHUNK
        blocks[k] = np.ascontiguousarray(u, dtype=MEETING_UNIT_DTYPE)
        counts[k] = len(u)
    path = meetings_path(prefix, st, num_objects)
    with open(path, "wb") as f:
        f.write(counts.tobytes())
        for b in blocks:
            f.write(b.tobytes())
    return path


def read_meetings(prefix, st, num_objects):
    """the 100 unit blocks of a meetings file, indexed by timestep % 100
    (workbench::load_meetings, workbench.cpp:457-470; the block of timestep t
    starts after the blocks of the indices below t % 100 — the reference's
    loader, geometry.cu:1612-1615, sums one index short)"""
    from .engine import MEETING_UNIT_DTYPE
    path = meetings_path(prefix, st, num_objects)
    raw = np.fromfile(path, dtype=np.uint8)
    counts = raw[:4 * STEPS_PER_CHUNK].view("<u4")
    units = raw[4 * STEPS_PER_CHUNK:]
    if units.size != 32 * int(counts.sum()):
        raise ValueError(f"{path}: {units.size} unit bytes for {int(counts.sum())} units")
    units = units.view(MEETING_UNIT_DTYPE)
    off = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
    return [units[off[k]:off[k + 1]] for k in range(STEPS_PER_CHUNK)]
