// psi.cpp -- Psi initialisers (src/psi.cu:7-116, 135-176, 351-395).
#include "../hpp/psi.hpp"

#include <cstdio>

Psi::Psi(const Params& p, const std::string&) : generator(p.getInt("rndseed")), distribution(-1.0, 1.0)
{
    psi = (complex*)malloc(p.Npoints * sizeof(complex));
    if (p.getBool("read_psi")) {
        std::cout << "# Reading initial wavefunction from: " << p.inpath << std::endl;
        readPsi(p.inpath, p.Npoints);
    } else {
        bool periodic = std::any_of(p.omega, p.omega + 3, [](double x) { return std::abs(x) < 1E-6; });
        if (periodic && std::abs(p.omega[0]) < 1E-6) {
            std::cout << "# Initializing tubular psi_init..." << std::endl;
            initTube(p, 0);
        } else {
            std::cout << "# Initializing harmonic psi_init..." << std::endl;
            initHarmonic(p, 0);
        }
        normalize(p);
    }
}

Psi::Psi(const ParamsBase& p, const int& type) : generator(p.getInt("rndseed") + type), distribution(-1.0, 1.0)
{
    psi = (complex*)malloc(p.Npoints * sizeof(complex));
    if (p.getBool("read_psi")) {
        std::cout << "# Reading initial wavefunction from: " << p.inpath << std::endl;
        readPsi(p.inpath + std::to_string(type), p.Npoints);
    } else {
        const double* omega = p._omega(type);
        bool periodic = std::any_of(omega, omega + 3, [](double x) { return std::abs(x) < 1E-6; });
        if (periodic && std::abs(omega[0]) < 1E-6) {
            std::cout << "# Initializing tubular psi_init..." << std::endl;
            initTube(p, type);
        } else {
            std::cout << "# Initializing harmonic psi_init..." << std::endl;
            initHarmonic(p, type);
        }
        normalize(p, type);
    }
}

void Psi::initTube(const ParamsBase& p, const int& type)
{
    double rrnd = p.getDouble("rand_pct");
    double crnd = p.getDouble("crand_pct");
    const double* ax = p._ax(type);
    for (int i = 0; i < p.Npoints; ++i) {
        double y = getY(p, i) / ax[1];
        double z = getZ(p, i) / ax[2];
        complex noise = complex(rrnd * distribution(generator), crnd * distribution(generator));
        psi[i] = (complex(1, 0) + noise) * complex(std::exp(-0.5 * 0.1 * ((y * y) + (z * z))), 0);
    }
}

void Psi::initHarmonic(const ParamsBase& p, const int& type)
{
    double rrnd = p.getDouble("rand_pct");
    double crnd = p.getDouble("crand_pct");
    const double* ax = p._ax(type);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) / ax[0];
        double y = getY(p, i) / ax[1];
        double z = getZ(p, i) / ax[2];
        complex noise = complex(rrnd * distribution(generator), crnd * distribution(generator));
        psi[i] = (complex(1, 0) + noise) * complex(std::exp(-0.5 * 0.1 * ((y * y) + (z * z) + (x * x))), 0);
    }
}

void Psi::initVoid(const ParamsBase& p, const int&)
{
    double rrnd = p.getDouble("rand_pct");
    double crnd = p.getDouble("crand_pct");
    for (int i = 0; i < p.Npoints; ++i) {
        complex noise = complex(rrnd * distribution(generator), crnd * distribution(generator));
        psi[i] = (complex(1, 0) + noise);
    }
}

double Psi::getX(const ParamsBase& p, int i) { return -p.XMAX[0] + iX(i, p.NX[0], p.NX[1]) * p.DX[0] + 0.5 * p.DX[0]; }
double Psi::getY(const ParamsBase& p, int i) { return -p.XMAX[1] + iY(i, p.NX[0], p.NX[1]) * p.DX[1] + 0.5 * p.DX[1]; }
double Psi::getZ(const ParamsBase& p, int i) { return -p.XMAX[2] + iZ(i, p.NX[0], p.NX[1]) * p.DX[2] + 0.5 * p.DX[2]; }

void Psi::imprintVortexTube(const ParamsBase& p, const double& Ncirc)
{
    double xLen = 2 * p.XMAX[0];
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i);
        psi[i] = complex(std::cos(Ncirc * 2 * pi / xLen * x), std::sin(Ncirc * 2 * pi / xLen * x)) * psi[i];
    }
}

void Psi::addNoise(complex& noise_amp, const int& length)
{
    for (int i = 0; i < length; ++i) {
        complex noise = complex(std::real(noise_amp) * distribution(generator),
                                std::imag(noise_amp) * distribution(generator));
        psi[i] = (complex(1, 0) + noise) * psi[i];
    }
}

void Psi::normalize(const ParamsBase& p, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    if (!p.isTwoComponent) npart = p._npart(1);
    double norm = 0;
    for (int i = 0; i < p.Npoints; ++i) norm += std::pow(std::abs(psi[i]), 2);
    norm = std::pow(npart / (norm * p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    for (int i = 0; i < p.Npoints; ++i) psi[i] = psi[i] * norm;
    norm = 0;
    for (int i = 0; i < p.Npoints; ++i) norm += std::pow(std::abs(psi[i]), 2) * p.DX[0] * p.DX[1] * p.DX[2];
    std::cout << "# Init Psi norm: " << norm << std::endl;
}

// A missing file is an error here (the reference reads from a NULL FILE*,
// src/psi.cu:383-394).
void Psi::readPsi(const std::string& prefix, const int& length)
{
    std::string name = prefix + "psi_final.wdat";
    FILE* file = fopen(name.c_str(), "rb");
    if (file == NULL) {
        fprintf(stderr, "Error: Could not open file %s!\n", name.c_str());
        exit(2);
    }
    size_t n = fread(psi, sizeof(complex), length, file);
    if (n != (size_t)length)
        fprintf(stderr, "Error: Only %zu elements were read instead of %d!\n", n, length);
    fclose(file);
}
