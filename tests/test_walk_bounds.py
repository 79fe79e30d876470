"""Host-side replay of the kernels' tile walk and addressing (CPU only).

Builds tests/walk_check.cpp against the library's own kernel header with
hipcc and runs it: every address an iteration kernel forms must stay inside
its system, every cell must be computed exactly once and every partial-sum
slot written exactly once, for every launch shape the library can choose.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_tile_walk_and_addressing(tmp_path):
    exe = tmp_path / "walk_check"
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-O2",
           "-o", str(exe), os.path.join(HERE, "walk_check.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok "), out.stdout
    assert int(out.stdout.split()[1]) > 100
