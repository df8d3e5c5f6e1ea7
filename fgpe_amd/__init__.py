"""MI355X-native time-stepping hot path of the fGPE Gross-Pitaevskii solver.

The compute path is libfgpe.so (hand-written HIP kernels for gfx950, C ABI in
include/fgpe_c.h); this package holds its build recipe, a ctypes binding and
the workload definitions.
"""
from ._lib import (FGPE_IMAG_HEAVY_BALL, FGPE_IMAG_RK4_RENORM, FGPE_PREC_F64,  # noqa: F401
                   FGPE_PREC_MIXED, INIT_HARMONIC, INIT_TUBE, INIT_VOID, LATTICES, Context, FgpeError, Geom,
                   Graph, Group, load)
