// FP64 instantiation of the pass-kernel launchers.
#include "launch.hpp"
template struct fgpe::Launch<double>;
