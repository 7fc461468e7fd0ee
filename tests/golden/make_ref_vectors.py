"""Writes ref_vectors.npz: input/output vectors of the reference's own
distance, bit-width, conversion and key-field helpers, computed by the
reference code compiled from its sources into oracle/_ref/libstcref.so
(oracle/Makefile; src/geometry/geometry.h:40-48, src/util/util.h:59-66,
527-553). The committed vectors pin the oracle where the reference tree is
absent. Run from the repository root: python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def main():
    ref = O.load_ref()
    if ref is None:
        sys.exit("oracle/_ref/libstcref.so is missing: run `make -C oracle` with the reference tree present")
    r = np.random.default_rng(20261202)
    lat = np.concatenate([r.uniform(-89.9, 89.9, 1500), r.uniform(41.5, 42.5, 1500), [0.0, 41.644584027118924]])
    dpm = np.array([ref.ref_dpm_lon(v) for v in lat])

    pts = np.empty((3000, 4))
    pts[:, 0] = r.uniform(-88.0, -87.5, 3000)
    pts[:, 1] = r.uniform(41.6, 42.1, 3000)
    pts[:, 2] = pts[:, 0] + r.normal(0, 2e-4, 3000)
    pts[:, 3] = pts[:, 1] + r.normal(0, 2e-4, 3000)
    dist = np.array([ref.ref_point_distance_geo(*p) for p in pts])

    boxes, probes, bdist = [], [], []
    while len(boxes) < 3000:
        lo = np.array([r.uniform(-88, -87.6), r.uniform(41.6, 42.0)])
        hi = lo + r.uniform(1e-5, 1e-2, 2)
        px, py = r.uniform(lo[0] - 0.01, hi[0] + 0.01), r.uniform(lo[1] - 0.01, hi[1] + 0.01)
        if lo[0] <= px <= hi[0] and lo[1] <= py <= hi[1]:
            continue  # inside: the reference returns 0 by a containment test the device form does not make
        b = np.array([lo[0], lo[1], hi[0], hi[1]])
        boxes.append(b)
        probes.append((px, py))
        bdist.append(ref.ref_box_distance_geo(b.ctypes.data, px, py))

    mb_in = np.array([0, 1, 2, 3, 4, 7, 8, 4095, 4096, 9999, 10000, 10001, 999999, 10**6, 10**7, 2**24 - 1, 2**24,
                      2**31, 2**32 - 1] + list(r.integers(0, 2**32, 2000)), dtype=np.uint64)
    mb_out = np.array([ref.ref_min_bits(int(v)) for v in mb_in], dtype=np.uint32)

    conv_in = np.array([0.0, 0.5, 1.0, 1.9, -0.5, -1.0, -5.0e9, -3.3, 4294967295.0, 4294967296.0, 5e9, 1e19, -1e19,
                        2147483647.0, 2147483648.0, -2147483649.0, 3e9] + list(r.uniform(-1e10, 1e10, 1000)))
    conv_u = np.array([ref.ref_conv_d2u(v) for v in conv_in], dtype=np.uint32)
    conv_i = np.array([ref.ref_conv_d2i(v) for v in conv_in], dtype=np.int32)

    keys = np.empty((1000, 2), dtype=np.uint64)
    fields = np.zeros((1000, 5), dtype=np.uint32)
    for i in range(1000):
        sid, pid, tgt = int(r.integers(0, 4096)), int(r.integers(0, 2**26)), int(r.integers(0, 2**26))
        dur, end = int(r.integers(0, 2**32)), int(r.integers(0, 2**32))
        k = (sid << 116) | (pid << 90) | (tgt << 64) | (dur << 32) | end
        keys[i] = (k & (2**64 - 1), k >> 64)
        ref.ref_key_fields(int(keys[i, 0]), int(keys[i, 1]), fields[i].ctypes.data)

    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), lat=lat, dpm_lon=dpm, pts=pts, dist=dist,
                        boxes=np.array(boxes), probes=np.array(probes), box_dist=np.array(bdist), mb_in=mb_in,
                        mb_out=mb_out, conv_in=conv_in, conv_u=conv_u, conv_i=conv_i, keys=keys, fields=fields)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))


if __name__ == "__main__":
    main()
