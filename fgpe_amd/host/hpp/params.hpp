// params.hpp -- parsed parameters in code units (hpp/params.hpp:11-108,
// hpp/params2.hpp:20-60): grid, box, k mesh, DV, oscillator length and
// frequency, couplings a (in a_ho) and the run controls.
#ifndef FGPE_HOST_PARAMS_HPP
#define FGPE_HOST_PARAMS_HPP

#include "header.hpp"
#include "input_variables.hpp"

class ParamsBase {
protected:
    InputVariables iv;

public:
    std::string outpath = "", inpath = "", vpath = "";
    bool isTwoComponent = false;
    int NX[3] = {0, 0, 0};
    int DIM = 3;
    int Npoints = 0;
    double V = 0;
    double XMAX[3] = {0, 0, 0};
    double DX[3] = {0, 0, 0};
    double DV = 0;
    double KXMAX[3] = {0, 0, 0};
    double DKX[3] = {0, 0, 0};
    double hamiltonianParams[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    virtual const double* _omega(int type) const = 0;
    virtual const double* _ax(int type) const = 0;
    virtual double _aho(int type) const = 0;
    virtual double _omho(int type) const = 0;
    virtual double _mass(int type) const = 0;
    virtual double _npart(int type) const = 0;
    virtual double _a(int type) const = 0;
    virtual double _add(int type) const = 0;
    virtual double _edd(int type) const = 0;

    double dt = 0, epsilon_e = 0, time0 = 0;
    int niter = 0, itmod = 0, iter0 = 0;

    int getInt(const std::string& varName) const { return iv.getInt(varName); }
    double getDouble(const std::string& varName) const { return iv.getDouble(varName); }
    std::string getString(const std::string& varName) const { return iv.getString(varName); }
    bool getBool(const std::string& varName) const { return iv.getBool(varName); }
    InputVariables getIV() const { return iv; }
    virtual ~ParamsBase() = default;
};

class Params : public ParamsBase {
private:
    void calculateLengthScales();
    void InitializeParams(const Params* par);

public:
    // the soft-core droplet-lattice box (hamiltonian "soft"): lengths made a
    // whole number of inter-droplet distances
    void RedefineLatticeSoft(const Params* par);
    double omega[3] = {0, 0, 0};
    double ax[3] = {0, 0, 0};
    double aho = 0, omho = 0;
    double mass = 0;
    double npart = 0;
    double a = 0, add = 0, edd = 0;

    const double* _omega(int) const override { return omega; }
    const double* _ax(int) const override { return ax; }
    double _aho(int) const override { return aho; }
    double _omho(int) const override { return omho; }
    double _mass(int) const override { return mass; }
    double _npart(int) const override { return npart; }
    double _a(int) const override { return a; }
    double _add(int) const override { return add; }
    double _edd(int) const override { return edd; }

    Params() {}
    Params(const Params& other);
    Params& operator=(const Params& other);
    Params(const InputVariables& iv);
};

class Params2 : public ParamsBase {
private:
    void calculateLengthScales();
    void InitializeParams(const Params2* par);

public:
    double omega1[3] = {0, 0, 0};
    double omega2[3] = {0, 0, 0};
    double ax1[3] = {0, 0, 0};
    double ax2[3] = {0, 0, 0};
    double aho1 = 0, omho1 = 0;
    double aho2 = 0, omho2 = 0;
    double mass1 = 0, mass2 = 0;
    double npart1 = 0, npart2 = 0;
    double a11 = 0, a22 = 0, a12 = 0;

    const double* _omega(int type) const override { return (type == 1) ? omega1 : omega2; }
    const double* _ax(int type) const override { return (type == 1) ? ax1 : ax2; }
    double _aho(int type) const override { return (type == 1) ? aho1 : aho2; }
    double _omho(int type) const override { return (type == 1) ? omho1 : omho2; }
    double _mass(int type) const override { return (type == 1) ? mass1 : mass2; }
    double _npart(int type) const override { return (type == 1) ? npart1 : npart2; }
    double _a(int type) const override { return (type == 1) ? a11 : a22; }
    double _add(int) const override { return 0; }
    double _edd(int) const override { return 0; }
    double _a12(int) const { return a12; }

    Params2() {}
    Params2(const Params2& other);
    Params2& operator=(const Params2& other);
    Params2(const InputVariables& iv);
};

#endif
