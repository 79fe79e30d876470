"""Build step: compiles libeafp.so for gfx950 with the committed Makefile."""
from __future__ import annotations

import os
import subprocess

from ._lib import CSRC, LIB_PATH


def build_library(force: bool = True, jobs: int = 4) -> str:
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"build finished but {LIB_PATH} is missing")
    return LIB_PATH
