// The per-length launchers (and so the pass kernels) of one precision and
// pencil length, both plan variants: the build recipe compiles this unit once
// per (FGPE_INST_T, FGPE_INST_N) so the kernels of different lengths compile
// in parallel.
#include "launch.hpp"

#if !defined(FGPE_INST_T) || !defined(FGPE_INST_N)
#error "compile with -DFGPE_INST_T=<double|float> -DFGPE_INST_N=<length>"
#endif

namespace fgpe {
FGPE_LAUNCH_INST(, FGPE_INST_T, FGPE_INST_N, 0)
FGPE_LAUNCH_INST(, FGPE_INST_T, FGPE_INST_N, 1)
// the x class's third plan (the lengths of FGPE_FOR_EACH_N_X2)
#if FGPE_INST_N == 128 || FGPE_INST_N == 256
FGPE_LAUNCH_INST_X(, FGPE_INST_T, FGPE_INST_N, 2)
#endif
}  // namespace fgpe
