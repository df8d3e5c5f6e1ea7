// FP64 dispatch of the pass-kernel launchers (the per-length launchers they
// call are instantiated in launch_inst.hip, one length per unit).
#include "launch.hpp"

namespace fgpe {
#define FGPE_DECL(N_) FGPE_LAUNCH_INST(extern, double, N_, 0) FGPE_LAUNCH_INST(extern, double, N_, 1)
FGPE_FOR_EACH_N(FGPE_DECL)
#undef FGPE_DECL
#define FGPE_DECL(N_) FGPE_LAUNCH_INST_X(extern, double, N_, 2)
FGPE_FOR_EACH_N_X2(FGPE_DECL)
#undef FGPE_DECL
}  // namespace fgpe

template struct fgpe::Launch<double>;
