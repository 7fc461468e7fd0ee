"""The device's runs written by stc_dump_run are byte-identical to the
oracle's runs written by stcdb_amd/runfiles.py, file by file, and read back
into the arrays stc_get_run returns."""
import os

import numpy as np
import pytest

import oracle_lib as O
from stcdb_amd import Pipeline, runfiles, workload
from stcdb_amd.engine import HostTrace
from test_runfiles import check_round_trip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("compat", [0, 1])
def test_device_run_files_equal_oracle_run_files(tmp_path, compat):
    w = workload("c1", num_objects=3000, kv_restriction=300, min_meet_time=3, compat=compat, area_threshold=1e-7)
    w["gen"].update(n_hotspots=3, hotspot_frac=0.7, hotspot_sigma=150.0)
    gpu = Pipeline(w["cfg"], w["box"])
    orc = O.Oracle(w["cfg"], w["box"])
    for t, xy in HostTrace(w["gen"]).steps(40):
        gpu.step(xy, t)
        orc.step(xy, t)
    assert gpu.num_runs() == orc.num_runs() >= 2
    gdir, odir = tmp_path / "gpu", tmp_path / "orc"
    gdir.mkdir()
    odir.mkdir()
    for r in range(gpu.num_runs()):
        gpu.dump_run(r, str(gdir), "6oversize_10", r)
        runfiles.dump_run(orc.run(r), str(odir), "6oversize_10", r)
        run = gpu.run(r)
        check_round_trip(run, runfiles.load_run(str(gdir), "6oversize_10", r, run["info"]["num_objects"],
                                                run["info"]["ctf_count"]))
    names = sorted(os.listdir(gdir))
    assert names == sorted(os.listdir(odir)) and len(names) > 100
    for name in names:
        assert (gdir / name).read_bytes() == (odir / name).read_bytes(), name
    print(f"[runfiles] {len(names)} files equal", flush=True)
    gpu.close()
    orc.close()
