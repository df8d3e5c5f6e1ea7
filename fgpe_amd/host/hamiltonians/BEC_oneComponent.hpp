// BEC_oneComponent.hpp -- the single-component BEC plugin
// (hamiltonians/BEC_oneComponent.hpp:23-120).  Its Hamiltonian is the
// library's fused contact kernel with the d_quench_as ramp of a(t)
// (hamiltonians/BEC_oneComponent.cu:128-178), configured in the constructor.
#ifndef FGPE_HOST_BEC_ONE_COMPONENT_HPP
#define FGPE_HOST_BEC_ONE_COMPONENT_HPP

#include "../hpp/oneComponentGPSolver.hpp"

class BEConeComponent : public OneComponentGPSolver {
public:
    Params p;
    BEConeComponent(Params& p, int precision = FGPE_PREC_F64, bool run = true);
    // The reference returns the Psi/Vext objects by value (a shallow copy of
    // their arrays); the copies are unused, so these return nothing.
    void InitPsi();
    void InitVext();
    void run();

    void call_IM_init() override;
    void call_IM_loop_before_step() override;
    void call_IM_loop_after_step() override;
    void call_IM_loop_convergence() override;
    void call_IM_end() override;
    void call_RE_init() override;
    void call_RE_loop_before_step() override;
    void call_RE_loop_after_step() override;
    void call_RE_end() override;

    void alg_calcCos(double* _avgcos, DeviceArray _psi) { alg_calcCosine(_avgcos, _psi); }
};

// A constant-a one-component plugin with the dipolar and soft-core k-space
// terms switched on from the input file (keys add, dipole_x/y/z, sc_U, sc_R,
// sc_dim): the physics the reference's Vt generators (src/parallel.cu:326-477)
// were written for.  Hooks are those of BEConeComponent.
class BECDipolar : public BEConeComponent {
public:
    BECDipolar(Params& p, int precision = FGPE_PREC_F64);
};

#endif
