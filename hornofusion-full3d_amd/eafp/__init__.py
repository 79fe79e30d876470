"""eafp -- MI355X-native p' (pressure-correction) solver for EAF3D.

The compute path is libeafp.so (hand-written HIP kernels for gfx950 behind
the C-ABI of include/eafp.h).  This package holds the host-side mirror of
the reference interface (``solver.cg_3d_mpi``), device-resident batches
(``solver.Batch``), the seeded synthetic system generator (``systems``) and
the build step (``build``).
"""
from . import systems  # noqa: F401
from ._lib import EafpError  # noqa: F401
from .solver import (Batch, Mesh, cg_3d_mpi, counters, pp_assemble, pp_correct, pp_step, release,  # noqa: F401
                     tdma_3d_mpi, tdma_batch)

__all__ = ["Batch", "EafpError", "Mesh", "cg_3d_mpi", "counters", "pp_assemble", "pp_correct", "pp_step", "release",
           "systems",
           "tdma_3d_mpi", "tdma_batch"]
