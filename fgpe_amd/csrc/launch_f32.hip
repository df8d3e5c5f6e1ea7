// FP32 (mixed-precision bulk) instantiation of the pass-kernel launchers.
#include "launch.hpp"
template struct fgpe::Launch<float>;
