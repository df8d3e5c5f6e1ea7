// baseHamiltonian.hpp -- the device context and the reference's algorithm
// building blocks (hpp/baseHamiltonian.hpp:11-77, with the device runtime of
// hpp/parallel.hpp:11-128 folded in).  Device work goes through the C ABI
// (include/fgpe_c.h); no GPU runtime header is included here.
//
// The reference's alg_* functions take raw device pointers.  Here they take
// DeviceArray tags naming the solver's arrays (d_psi, d_hpsi, ...), so plugin
// code keeps its shape: alg_calcNorm(&PARS[NORM], this->d_psi).
#ifndef FGPE_HOST_BASE_HAMILTONIAN_HPP
#define FGPE_HOST_BASE_HAMILTONIAN_HPP

#include "output.hpp"
#include "params.hpp"

struct DeviceArray {
    enum Kind { PSI = 0, HPSI = 1, HPSI_EN = 2, AUX = 3 } kind;
    int comp;
};

class BaseHamiltonian {
public:
    ParamsBase* pars;
    Output output;
    fgpe_ctx* ctx = nullptr;
    int precision;

    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *kx = nullptr, *ky = nullptr, *kz = nullptr;

    // The plugin's Hamiltonian: set in the plugin constructor (the role of
    // d_h_params, hamiltonians/BEC_oneComponent.cu:10-15).
    fgpe_hparams hparams{};

    DeviceArray d_psi{DeviceArray::PSI, 0}, d_psi2{DeviceArray::PSI, 1};
    DeviceArray d_hpsi{DeviceArray::HPSI, 0}, d_hpsi2{DeviceArray::HPSI, 1};
    DeviceArray d_hpsi_en{DeviceArray::HPSI_EN, 0}, d_hpsi2_en{DeviceArray::HPSI_EN, 1};
    DeviceArray d_aux{DeviceArray::AUX, 0}, d_aux2{DeviceArray::AUX, 1};
    DeviceArray d_psi_new{DeviceArray::AUX, 0}, d_psi2_new{DeviceArray::AUX, 1};

    explicit BaseHamiltonian(ParamsBase* par, int precision = FGPE_PREC_F64);
    BaseHamiltonian(const BaseHamiltonian&) = delete;
    BaseHamiltonian& operator=(const BaseHamiltonian&) = delete;
    virtual ~BaseHamiltonian();

    inline int iX(int index) { return (index % pars->NX[0]); }
    inline int iY(int index) { return ((index / pars->NX[0]) % pars->NX[1]); }
    inline int iZ(int index) { return (index / (pars->NX[0] * pars->NX[1])); }

    void InitializeKspace();
    void InitializeSpace();

    // reductions of src/baseHamiltonian.cu:173-236 on the current state
    void alg_calcNorm(double* norm, DeviceArray psi);
    void alg_calc2Observables(double norm, double* o1, double* o2, DeviceArray psi, DeviceArray o1psi,
                              DeviceArray o2psi);
    void alg_calcPx(double* px, DeviceArray psi, DeviceArray aux);
    void alg_calcCosine(double* avgcos, DeviceArray psi);

    // Runs the time steps queued by the solver loop, so that device state is
    // current; every alg_* call does this first.
    virtual void syncState() {}
    double obsTime = 0.0;   // time argument of H in the observables

protected:
    double obs[6] = {0, 0, 0, 0, 0, 0};
    bool obsValid = false;
    void currentObservables();
    void invalidate() { obsValid = false; }
};

#endif
