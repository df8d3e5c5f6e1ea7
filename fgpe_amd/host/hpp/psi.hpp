// psi.hpp -- the initial wave function built on the host (the reference's
// Psi, hpp/psi.hpp): Gaussian or tube envelope with seeded noise, a vortex
// phase imprint, lattice profiles, normalisation, or a state read from
// psi_final.wdat.  The public members are the ones a plugin calls.
#ifndef FGPE_HOST_PSI_HPP
#define FGPE_HOST_PSI_HPP

#include <functional>
#include <random>

#include "params.hpp"

class Psi {
public:
    Psi(const Params& p, const std::string& type);   // one component
    Psi(const ParamsBase& p, const int& type);       // component type (1 or 2) of two
    Psi(const Psi&) = delete;
    Psi& operator=(const Psi&) = delete;
    ~Psi() { free(psi); }

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

    // grid index -> axis index, and cell centres
    inline int iX(int index, int NX, int) { return index % NX; }
    inline int iY(int index, int NX, int NY) { return (index / NX) % NY; }
    inline int iZ(int index, int NX, int NY) { return index / (NX * NY); }
    double getX(const ParamsBase& p, int i);
    double getY(const ParamsBase& p, int i);
    double getZ(const ParamsBase& p, int i);

    complex* getPsi() const { return psi; }

private:
    enum class Envelope { Tube, Gaussian, Flat };
    // The seeded noise: std::mt19937 with uniform numbers on [-1, 1).
    class Noise {
    public:
        explicit Noise(int seed) : gen_(seed), u_(-1.0, 1.0) {}
        complex draw(double re_amp, double im_amp);
    private:
        std::mt19937 gen_;
        std::uniform_real_distribution<double> u_;
    };

    complex* psi = nullptr;
    Noise noise;

    Psi(const ParamsBase& p, int type, int seed, const std::string& read_prefix);
    void fill(const ParamsBase& p, int type, Envelope env);
    // psi += f(x - off) sqrt(npart / dV) at every point, then normalize(p, 1)
    void addProfile(const ParamsBase& p, int type, const double off[3],
                    const std::function<double(double, double, double)>& f);
};

#endif
