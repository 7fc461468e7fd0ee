"""stcdb_amd — MI355X-native per-timestep co-occurrence pipeline of STCDB.

The hot path (QTree partition, pair refinement, meeting table, identify and
the sort-based LSM flush) runs as hand-written HIP kernels in
libstcdb_hip.so behind the C ABI of include/stcdb.h; this package is the
host-side mirror of the reference's tracer/workbench interface.
"""
from ._lib import load, StcError, LIB_PATH  # noqa: F401
from .engine import Pipeline, HostTrace, DeviceTrace, step_batch  # noqa: F401
from .workloads import WORKLOADS, workload, batched_streams, map_box, gen_params, pipeline_config  # noqa: F401
