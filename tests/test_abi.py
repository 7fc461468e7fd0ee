"""The C-ABI library loads and exports every symbol that include/stcdb.h declares."""
import ctypes
import os
import re
import subprocess

import stcdb_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "stcdb.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(stc_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ["stc_create", "stc_destroy", "stc_step", "stc_step_device", "stc_get_run", "stc_device_count"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """every ctypes mirror has the size and field offsets gcc gives the C structs"""
    structs = {"stc_config": L.StcConfig, "stc_box": L.StcBox, "stc_step_out": L.StcStepOut,
               "stc_qtnode": L.StcQtNode,
               "stc_ctf_meta": L.StcCtfMeta, "stc_run_info": L.StcRunInfo, "stc_device_info": L.StcDeviceInfo,
               "stc_gen_params": L.StcGenParams, "stc_walker": L.StcWalker}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "stcdb_amd/csrc/stc_gen.h"',
             '#include "include/stcdb.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", ROOT, "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_query_record_layouts_match_the_c_compiler(tmp_path):
    """the numpy records used for stc_query / stc_hit have gcc's layout"""
    from stcdb_amd.engine import HIT_DTYPE, QUERY_DTYPE
    dtypes = {"stc_query": QUERY_DTYPE, "stc_hit": HIT_DTYPE}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "include/stcdb.h"', "int main(void){"]
    for cname, dt in dtypes.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in dt.names:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "qlayout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "qlayout"
    subprocess.run(["gcc", "-I", ROOT, "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for cname, dt in dtypes.items():
        assert int(got[cname]) == dt.itemsize, cname
        for f in dt.names:
            assert int(got[f"{cname}.{f}"]) == dt.fields[f][1], (cname, f)
