// vext.hpp -- the external potential built on the host (the reference's
// Vext, hpp/vext.hpp): harmonic, tube or box trap, cosine seeding
// protocols, or a trap read from vext.wdat.  The public members are the ones
// a plugin calls.
#ifndef FGPE_HOST_VEXT_HPP
#define FGPE_HOST_VEXT_HPP

#include "params.hpp"

class Vext {
public:
    Vext(const Params& p, const std::string& type);   // one component
    Vext(const Params2& p, const int& type);          // component type (1 or 2) of two
    Vext(const Vext&) = delete;
    Vext& operator=(const Vext&) = delete;
    ~Vext() { free(vext); }

    void initTube(const ParamsBase& p, const int& type);
    void initHarmonic(const ParamsBase& p, const int& type);
    void initBox(const ParamsBase& p, const int& type);
    void initVoid(const ParamsBase& p, const int& type);

    complex* getVext() { return vext; }
    void readVext(const std::string& prefix, const int& length);
    // cosine seeds A cos(2 pi n x / (2 XMAX_x)) added to the trap
    void addProtocolPotential(const ParamsBase& p, const int& type, const int sign = 1);
    void addWeightedProtocolPotential(const Params2& p, const int& type);
    // the amplitude A of addWeightedProtocolPotential (printed, as the reference does)
    static double weightedProtocolAmplitude(const Params2& p, const int& type);
    void addOpticalLattice(const ParamsBase& p, const int& type);

    // grid index -> axis index, and cell centres
    inline int iX(int index, int NX, int) { return index % NX; }
    inline int iY(int index, int NX, int NY) { return (index / NX) % NY; }
    inline int iZ(int index, int NX, int NY) { return index / (NX * NY); }
    double getX(const ParamsBase& p, int i);
    double getY(const ParamsBase& p, int i);
    double getZ(const ParamsBase& p, int i);

private:
    complex* vext = nullptr;

    Vext(const ParamsBase& p, int type, const std::string& read_prefix, bool say);
    void harmonic(const ParamsBase& p, int type, bool along_x);
    void addCosine(const ParamsBase& p, double A, double n);
};

#endif
