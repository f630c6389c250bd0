"""delfem_amd: MI355X-native DelFEM FEM hot path (P1 assembly -> block-CRS -> ILU(0) -> PCG/PBiCGSTAB).

The compute path is libdfm_gpu.so (hand-written HIP for gfx950 behind the C-ABI in include/dfm_gpu.h);
this package is its Python binding (`_native`) plus the synthetic workload definitions (`workloads`)."""
from . import _native  # noqa: F401
from ._native import (DIFFUSION2D, FP64, MIXED, ORDER_BLOCKCOLOR, ORDER_MULTICOLOR, ORDER_NATURAL, POISSON,  # noqa
                      SOLID2D, SOLID3D, STVK2D, STVK3D, Context, DfmError, System)
