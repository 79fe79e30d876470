"""The C-ABI structs as the C compiler lays them out (include/eafp.h,
compiled with gcc) against the ctypes mirrors of the Python binding: every
size and every member offset.  The Fortran bind(C) mirrors are exercised
end to end by tests/test_fortran_shim.py."""
import ctypes
import os
import shutil
import subprocess

import pytest

from eafp import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "eafp_desc": _lib.Desc,
    "eafp_stats": _lib.Stats,
    "eafp_mixed_opts": _lib.MixedOpts,
    "eafp_pp_fields": _lib.PPFields,
    "eafp_pp_config": _lib.PPConfig,
    "eafp_pp_exports": _lib.PPExports,
    "eafp_pp_update": _lib.PPUpdate,
    "eafp_pp_step_opts": _lib.PPStepOpts,
    "eafp_tdma_system": _lib.TdmaSystem,
    "eafp_counters": _lib.Counters,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layouts_match_the_c_compiler(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "eafp.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'    printf("%s %zu\\n", "{cname}", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'    printf("%s.%s %zu\\n", "{cname}", "{fname}", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True, capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout.split("\n")
    got = dict(line.split() for line in out if line.strip())
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"
