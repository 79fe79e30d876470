"""Host replay of the GPU line TDMA's schedule (CPU only).

Builds tests/tdma_check.cpp against the library's own per-line code
(csrc/tdma_line.hpp) and the oracle, with AddressSanitizer on the host side,
and runs it: the anti-diagonal schedule, with the lines of each diagonal in
reverse and rotated order, must give phi bit-identical to the sequential
restatement, every line must be solved once per pass, the pivot pass must
cover every interior cell once, and no access may leave its array.
"""
import os
import shutil
import subprocess

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_tdma_schedule_replay(tmp_path):
    libdir = os.path.dirname(oracle.build())
    exe = tmp_path / "tdma_check"
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-O1", "-g", "-ffp-contract=off",
           "-Xarch_host", "-fsanitize=address",
           "-o", str(exe), os.path.join(HERE, "tdma_check.cpp"), "-L" + libdir, "-loracle", "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok "), out.stdout
    assert int(out.stdout.split()[1]) > 100
