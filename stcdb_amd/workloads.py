"""Seeded synthetic workloads for the co-occurrence pipeline.

Objects walk a Manhattan street grid with 100 m spacing (stcdb_amd/csrc/stc_gen.h).
Coordinates are produced in metres and mapped to degrees about the
reference map's south-west corner (the MBR of data/streets.map starts at
lon -87.9401, lat 41.6446): x = lon0 + X * dpm_lon(lat0), y = lat0 + Y * dpm_lat.

  c1  10,000 objects, 10 km x 10 km, uniform starts, 120 steps, r = 10 m,
      leaf capacity 200, minimum meeting 5 s
  c2  1,000,000 objects, 50 km x 50 km, 70 % of starts near 20 Gaussian
      hotspots (sigma 500 m), 3,600 steps, r = 20 m, minimum meeting 10 s
  c3  10,000,000 objects, c2's walk and hotspot mix, 600 steps, memtable
      runs of 8 Mi records, 256 CTFs
  c4  batched streams: 64 independent streams, object counts log-normal
      (median 50k, clipped to 1k-2M, ~6M in total), each with its own
      10-50 km box, reach uniform in 5-25 m and its own leaf capacity
      (batched_streams below), the reference's default dynamic schema; each
      stream is one context of the library
  c5  5,000,000 objects with c2's mix, runs of 8 Mi records kept in HBM for
      queries: the serving workload (ingest at 1 Hz plus 2,000 queries/s)
"""
import math

import numpy as np

LON0 = -87.94008647664191
LAT0 = 41.644584027118924
PI_REF = 3.14159265
DPM_LAT = 360.0 / (40076.0 * 1000.0)
MARGIN_M = 50.0


def dpm_lon(lat):
    """degree_per_meter_longitude (src/util/util.h:59-66), host form"""
    a = abs(lat)
    if a == 90:
        a = 89
    return 360.0 / (math.sin((90 - a) * PI_REF / 180) * 40076.0 * 1000.0)


def gen_params(num_objects, grid_cells, seed, hotspots=0, hotspot_frac=0.7, hotspot_sigma=500.0,
               spacing=100.0, vmin=1.0, vmax=15.0):
    return dict(seed=seed, num_objects=num_objects, grid_cells=grid_cells, spacing=spacing, vmin=vmin,
                vmax=vmax, n_hotspots=hotspots, hotspot_frac=hotspot_frac, hotspot_sigma=hotspot_sigma,
                lon0=LON0, lat0=LAT0, dlon=dpm_lon(LAT0), dlat=DPM_LAT)


def map_box(gp):
    """map MBR: the street grid plus a 50 m margin on every side"""
    w = gp["grid_cells"] * gp["spacing"]
    return ((LON0 - MARGIN_M * gp["dlon"], LAT0 - MARGIN_M * gp["dlat"]),
            (LON0 + (w + MARGIN_M) * gp["dlon"], LAT0 + (w + MARGIN_M) * gp["dlat"]))


def pipeline_config(num_objects, reach_distance, min_meet_time, grid_capacity=200, kv_restriction=67108864,
                    ctf_count=100, compat=0, area_threshold=100.0, max_meet_time=4096, grid_amplify=2.0,
                    start_time=0, num_meeting_buckets=0, keep_pairs=0, dynamic_schema=0, schema_update_delay=1,
                    async_flush=0):
    """The reference's configuration defaults (src/util/config.h:15-86):
    kv_restriction = G_bytes(2) * 33554432, CTF_count 100, grid_amplify 2,
    max_meet_time 4096; area_threshold 100 is the file_size = 10 entry of the
    oversize table (src/tracing/process.cpp:408-411), i.e. no object is cut."""
    return dict(num_objects=num_objects, grid_capacity=grid_capacity, grid_amplify=grid_amplify,
                min_meet_time=min_meet_time, max_meet_time=max_meet_time, start_time=start_time,
                kv_restriction=kv_restriction, ctf_count=ctf_count, compat=compat,
                num_meeting_buckets=num_meeting_buckets, reach_distance=reach_distance,
                area_threshold=area_threshold, keep_pairs=keep_pairs, dynamic_schema=dynamic_schema,
                schema_update_delay=schema_update_delay, async_flush=async_flush)


WORKLOADS = {
    "c1": dict(gen=gen_params(10_000, 100, seed=1), steps=120,
               cfg=pipeline_config(10_000, 10.0, 5, grid_capacity=200)),
    "c2": dict(gen=gen_params(1_000_000, 500, seed=2, hotspots=20), steps=3600,
               cfg=pipeline_config(1_000_000, 20.0, 10, grid_capacity=200)),
    "c3": dict(gen=gen_params(10_000_000, 500, seed=3, hotspots=20), steps=600,
               cfg=pipeline_config(10_000_000, 20.0, 10, grid_capacity=200, kv_restriction=8 * 1024 * 1024,
                                   ctf_count=256)),
    "c5": dict(gen=gen_params(5_000_000, 500, seed=5, hotspots=20), steps=3600,
               cfg=dict(pipeline_config(5_000_000, 20.0, 10, grid_capacity=200, kv_restriction=8 * 1024 * 1024,
                                        ctf_count=256), device_runs=1)),
}


def workload(name, **overrides):
    """a copy of a named workload; overrides apply to the pipeline config
    (num_objects also rescales the generator)"""
    w = WORKLOADS[name]
    gen = dict(w["gen"])
    cfg = dict(w["cfg"])
    for k, v in overrides.items():
        if k == "steps":
            continue
        cfg[k] = v
        if k == "num_objects":
            gen["num_objects"] = v
    return dict(gen=gen, cfg=cfg, steps=overrides.get("steps", w["steps"]), box=map_box(gen))


def batched_streams(n_streams=64, median=50_000, lo=1_000, hi=2_000_000, side_km=(10.0, 50.0),
                    reach_m=(5.0, 25.0), capacities=(50, 100, 200), min_meet_time=10, steps=600, seed=4,
                    dynamic_schema=1):
    """The c4 family: independent streams, one workload each (num_objects
    log-normal about `median` with sigma 1, clipped to [lo, hi]; a square box
    of side uniform in `side_km`; reach uniform in `reach_m`; a leaf capacity
    drawn from `capacities`, the starting point of the leaf-capacity tuner;
    the reference's default dynamic schema unless dynamic_schema is 0). All
    draws come from one seeded numpy generator, so the list depends only on
    the arguments."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_streams):
        n = int(np.clip(round(median * math.exp(rng.normal(0.0, 1.0))), lo, hi))
        cells = max(int(round(rng.uniform(*side_km) * 10.0)), 1)  # 100 m street spacing
        reach = float(round(rng.uniform(*reach_m), 3))
        cap = int(capacities[int(rng.integers(len(capacities)))])
        gen = gen_params(n, cells, seed=seed * 1000 + i)
        cfg = pipeline_config(n, reach, min_meet_time, grid_capacity=cap, dynamic_schema=dynamic_schema)
        out.append(dict(gen=gen, cfg=cfg, steps=steps, box=map_box(gen)))
    return out
