// header.hpp -- constants, types and the index macro shared by the host API
// (the reference's hpp/header.hpp:20-31, without any GPU runtime header: the
// host layer reaches the device only through include/fgpe_c.h).
#ifndef FGPE_HOST_HEADER_HPP
#define FGPE_HOST_HEADER_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../../../include/fgpe_c.h"

// 1D index of (x, y, z), x fastest (hpp/header.hpp:20)
#define ID(x, y, z, NX, NY) ((z) * (NX) * (NY) + (y) * (NX) + (x))

#define NO_HAMIL_PARAMS 32

constexpr double pi = 3.141592653589793;
constexpr double hbar = 63507.799;
constexpr double kB = 0.831466 * 1.E7;
constexpr double abohr = 5.2917721E-5;

typedef std::complex<double> complex;

// Prints the library's message for a failed C-ABI call and exits (the
// reference prints CUDA errors and carries on, src/parallel.cu:7-11; a failed
// call here leaves no valid state to carry on with).
void FCE(int rc, const char* what);

#endif
