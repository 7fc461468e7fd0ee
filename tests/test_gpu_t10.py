"""Flush and compaction parity at T10 widths.

The packed key of a record is pid << (id + duration + end + mbr bits) + ...
where the pid comes from extract_fields (src/cuda/cuda_util.cuh:163-169) with
the sid still above its 26 id bits, so once id_bit is large enough the low sid
bits fall inside key_bit and are stored (write_key_and_mbr,
src/cuda/geometry.cu:1073-1107). With 600,000 objects id_bit is 20
(min_bits_to_store), and the CTFs whose widths round key_bit up far enough
carry sid bits. Runs and a compaction of two runs are compared with the
oracle byte for byte, and the test checks that stored sid bits occur."""
import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import workload
from test_gpu_compact import lockstep, same_runs
from test_gpu_query import compare
from test_oracle_query import make_queries

pytestmark = pytest.mark.gpu


def stored_sid_bits(run):
    """CTFs whose packed keys hold bits of the sid above the 26 pid bits"""
    ctfs, raw = run["ctfs"], run["key_bytes"]
    off, hits = 0, 0
    for c in ctfs:
        kb, cap = int(c["key_bit"]) // 8, int(c["kv_capacity"])
        shift = int(c["id_bit"]) + int(c["duration_bit"]) + int(c["end_bit"]) + int(c["mbr_bit"])
        if cap and shift + 26 < 8 * kb:
            keys = raw[off:off + cap * kb].reshape(cap, kb)
            vals = [int.from_bytes(bytes(k), "little") >> shift for k in keys[:64]]
            hits += any(v >> 26 for v in vals)
        off += cap * kb
    return hits


def test_flush_and_compaction_parity_at_id_bit_20():
    w = workload("c2", num_objects=600_000, min_meet_time=3, kv_restriction=150_000, ctf_count=64)
    gpu, orc = lockstep(w, 9, "t10")
    assert orc.num_runs() >= 2
    r0 = orc.run(0)
    assert int(r0["ctfs"]["id_bit"][0]) == 20
    same_runs(gpu, orc, "t10 runs")
    n_sid = stored_sid_bits(r0)
    print(f"[t10] {orc.num_runs()} runs, {n_sid} CTFs of run 0 store sid bits", flush=True)
    assert n_sid > 0
    gpu.compact_runs(0, 2)
    orc.compact_runs(0, 2)
    same_runs(gpu, orc, "t10 merged")
    assert stored_sid_bits(orc.run(0)) > 0
    compare(gpu, orc, make_queries(orc, w, np.random.default_rng(4), 120), "t10 after merge")
    gpu.close()
    orc.close()
