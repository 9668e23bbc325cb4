"""gasoline_amd: MI355X force hot path for gasoline (tree gravity + SPH).

The compute lives in libgasoline_amd.so (HIP kernels for gfx950 behind the C
ABI of include/gasoline_amd.h); host/libggmsr.so holds the C host sequences
over that ABI (the master's force steps and block time steps).  This package
only binds them.  There is
no CPU fallback: importing the binding without the built library raises.
"""
from .pkd import (GGError, Pkd, Smf, lib, msr_lib, BlockSteps, StepParam, FIELDS, PREC_FP32, PREC_FP64,
                  SMX_DENSITY, SMX_DENDVDX, SMX_SPHPRESSURETERMS,
                  TYPE_ACTIVE, TYPE_TREEACTIVE, TYPE_SMOOTHACTIVE, TYPE_SMOOTHDONE,
                  TYPE_NbrOfACTIVE, TYPE_GAS, TYPE_DARK, TYPE_STAR, tipsy_header)

__all__ = ["GGError", "Pkd", "Smf", "lib", "msr_lib", "BlockSteps", "StepParam", "FIELDS", "PREC_FP32", "PREC_FP64",
           "SMX_DENSITY", "SMX_DENDVDX", "SMX_SPHPRESSURETERMS", "TYPE_ACTIVE",
           "TYPE_TREEACTIVE", "TYPE_SMOOTHACTIVE", "TYPE_SMOOTHDONE",
           "TYPE_NbrOfACTIVE", "TYPE_GAS", "TYPE_DARK", "TYPE_STAR", "tipsy_header"]
