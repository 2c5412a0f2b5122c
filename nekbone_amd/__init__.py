"""nekbone_amd: MI355X-native implementation of Nekbone's CG Poisson hot path.

The compute path is the HIP library libnekbone_amd.so (C ABI declared in
include/nekbone_amd.h); this package is its host-side binding.
"""
from .core import (ARITH_EXACT, ARITH_FMA, MIXED_F16_MFMA, MIXED_F32, MIXED_G16, KCLASS_AX,
                   KCLASS_GS, KCLASS_POST, KCLASS_REDUCE, KCLASS_UPDATE, LIB_PATH, MASK_DIRICHLET,
                   MASK_REF, NOT_CONVERGED, PRECOND_HSMG, PRECOND_IDENTITY, PRECOND_JACOBI, Batch,
                   DeviceVec, NekboneError, NotConverged, Problem, bw_probe, cubic, init, lib,
                   set_f_raw, sync, tuning_load, tuning_save)

__all__ = [
    "ARITH_EXACT", "ARITH_FMA", "MIXED_F16_MFMA", "MIXED_F32", "MIXED_G16", "KCLASS_AX",
    "KCLASS_GS", "KCLASS_POST", "KCLASS_REDUCE", "KCLASS_UPDATE", "LIB_PATH", "MASK_DIRICHLET",
    "MASK_REF", "NOT_CONVERGED", "PRECOND_HSMG", "PRECOND_IDENTITY", "PRECOND_JACOBI", "Batch",
    "DeviceVec", "NekboneError", "NotConverged", "Problem", "bw_probe", "cubic", "init", "lib",
    "set_f_raw", "sync", "tuning_load", "tuning_save",
]
