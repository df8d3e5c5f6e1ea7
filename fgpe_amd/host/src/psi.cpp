// psi.cpp -- Psi initialisers (src/psi.cu:7-395).
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

// Lattice profiles added to the real part with the mean amplitude
// sqrt(npart / dV), then normalised to component 1's particle number, as the
// reference does for either component.
void Psi::force1DLattice(const ParamsBase& p, const double& amp, const double& k, const double& off, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    double avg = std::pow(npart / (p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) - off;
        double force = amp * (1 - std::cos(k * x)) / 2.;
        psi[i] = force * avg + psi[i];
    }
    normalize(p, 1);
}

void Psi::force2DTriangularLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                                   const double& offy, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    double avg = std::pow(npart / (p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    const double h = std::sqrt(3.0) / 2;
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) - offx;
        double y = getY(p, i) - offy;
        double force = amp * (std::cos(k * x) + std::cos(k * (-0.5 * x + h * y)) + std::cos(k * (-0.5 * x - h * y))) / 3;
        psi[i] = force * avg + psi[i];
    }
    normalize(p, 1);
}

void Psi::force2DSquareLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                               const double& offy, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    double avg = std::pow(npart / (p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) - offx;
        double y = getY(p, i) - offy;
        double force = amp * (std::cos(k * y) + std::cos(k * x)) / 2;
        psi[i] = force * avg + psi[i];
    }
    normalize(p, 1);
}

void Psi::force3DHEXLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                            const double& offy, const double& offz, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    double avg = std::pow(npart / (p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    const double r3 = std::sqrt(3.);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) - offx, y = getY(p, i) - offy, z = getZ(p, i) - offz;
        double x1 = (-x + r3 * y) / 2., x4 = (-x - r3 * y) / 2.;
        double force = amp * (std::cos(k * x1) + std::cos(k * x) + std::cos(k * z) + std::cos(k * x4)) / 4.0;
        psi[i] = force * avg + psi[i];
    }
    normalize(p, 1);
}

void Psi::force3DFCCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                            const double& offy, const double& offz, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    double avg = std::pow(npart / (p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    const double r3 = std::sqrt(3.);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) - offx, y = getY(p, i) - offy, z = getZ(p, i) - offz;
        // the nine projections, (x+y+z) three times over as in the reference
        const double u[9] = {(x + y - z) / r3, (y + z - x) / r3, (z + x - y) / r3,
                             (x - y - z) / r3, (y - z - x) / r3, (z - x - y) / r3,
                             (x + y + z) / r3, (y + z + x) / r3, (z + x + y) / r3};
        double s = std::cos(k * u[0]);
        for (int j = 1; j < 9; ++j) s = s + std::cos(k * u[j]);
        double force = amp * s / 9.0;
        psi[i] = force * avg + psi[i];
    }
    normalize(p, 1);
}

void Psi::force3DBCCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                            const double& offy, const double& offz, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    double avg = std::pow(npart / (p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    const double r2 = std::sqrt(2.);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) - offx, y = getY(p, i) - offy, z = getZ(p, i) - offz;
        const double u[6] = {(x + y) / r2, (z + y) / r2, (z + x) / r2, (x - y) / r2, (z - y) / r2, (z - x) / r2};
        double s = std::cos(k * u[0]);
        for (int j = 1; j < 6; ++j) s = s + std::cos(k * u[j]);
        double force = amp * s / 6;
        psi[i] = force * avg + psi[i];
    }
    normalize(p, 1);
}

void Psi::force3DSCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                           const double& offy, const double& offz, const int& type)
{
    double npart = (type == 1) ? p._npart(1) : p._npart(2);
    double avg = std::pow(npart / (p.DX[0] * p.DX[1] * p.DX[2]), 0.5);
    for (int i = 0; i < p.Npoints; ++i) {
        double x = getX(p, i) - offx, y = getY(p, i) - offy, z = getZ(p, i) - offz;
        double force = amp * (std::cos(k * z) + std::cos(k * y) + std::cos(k * x)) / 3;
        psi[i] = force * avg + psi[i];
    }
    normalize(p, 1);
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
