// vext.hpp -- external potential (hpp/vext.hpp): harmonic, tube or box trap,
// cosine seeding protocols, read from vext.wdat.
#ifndef FGPE_HOST_VEXT_HPP
#define FGPE_HOST_VEXT_HPP

#include "params.hpp"

class Vext {
    complex* vext;

public:
    void initTube(const ParamsBase& p, const int& type);
    void initHarmonic(const ParamsBase& p, const int& type);
    void initBox(const ParamsBase& p, const int& type);
    void initVoid(const ParamsBase& p, const int& type);

    Vext(const Params& p, const std::string& type);
    Vext(const Params2& p, const int& type);
    Vext(const Vext&) = delete;
    Vext& operator=(const Vext&) = delete;
    ~Vext() { free(vext); }

    complex* getVext() { return vext; }
    void readVext(const std::string& prefix, const int& length);
    void addProtocolPotential(const ParamsBase& p, const int& type, const int sign = 1);
    void addWeightedProtocolPotential(const Params2& p, const int& type);
    // the amplitude A of addWeightedProtocolPotential (prints it, as the reference)
    static double weightedProtocolAmplitude(const Params2& p, const int& type);
    void addOpticalLattice(const ParamsBase& p, const int& type);

    inline int iX(int index, int NX, int) { return index % NX; }
    inline int iY(int index, int NX, int NY) { return (index / NX) % NY; }
    inline int iZ(int index, int NX, int NY) { return index / (NX * NY); }
    double getX(const ParamsBase& p, int i);
    double getY(const ParamsBase& p, int i);
    double getZ(const ParamsBase& p, int i);
};

#endif
