// psi.hpp -- initial wave function (hpp/psi.hpp): Gaussian or tube with
// seeded noise, vortex phase imprint, lattice profiles, normalisation, read
// from psi_final.wdat.
#ifndef FGPE_HOST_PSI_HPP
#define FGPE_HOST_PSI_HPP

#include <random>

#include "params.hpp"

class Psi {
    complex* psi;
    std::mt19937 generator;
    std::uniform_real_distribution<double> distribution;

public:
    Psi(const Params& p, const std::string& type);
    Psi(const ParamsBase& p, const int& type);
    Psi(const Psi&) = delete;
    Psi& operator=(const Psi&) = delete;

    void initTube(const ParamsBase& p, const int& type);
    void initHarmonic(const ParamsBase& p, const int& type);
    void initVoid(const ParamsBase& p, const int& type);
    void normalize(const ParamsBase& p, const int& type = -1);
    void addNoise(complex& noise_amp, const int& length);
    void imprintVortexTube(const ParamsBase& p, const double& Ncirc);
    void readPsi(const std::string& prefix, const int& length);

    // additive lattice profiles, each followed by normalize(p, 1)
    void force3DHEXLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                           const double& offy, const double& offz, const int& type = 1);
    void force3DFCCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                           const double& offy, const double& offz, const int& type = 1);
    void force3DBCCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                           const double& offy, const double& offz, const int& type = 1);
    void force3DSCLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                          const double& offy, const double& offz, const int& type = 1);
    void force2DSquareLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                              const double& offy, const int& type = 1);
    void force2DTriangularLattice(const ParamsBase& p, const double& amp, const double& k, const double& offx,
                                  const double& offy, const int& type = 1);
    void force1DLattice(const ParamsBase& p, const double& amp, const double& k, const double& off,
                        const int& type = 1);

    inline int iX(int index, int NX, int) { return index % NX; }
    inline int iY(int index, int NX, int NY) { return (index / NX) % NY; }
    inline int iZ(int index, int NX, int NY) { return index / (NX * NY); }
    double getX(const ParamsBase& p, int i);
    double getY(const ParamsBase& p, int i);
    double getZ(const ParamsBase& p, int i);

    complex* getPsi() const { return psi; }
    ~Psi() { free(psi); }
};

#endif
