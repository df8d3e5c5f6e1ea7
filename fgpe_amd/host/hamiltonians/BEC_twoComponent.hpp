// BEC_twoComponent.hpp -- the two-component BEC plugin
// (hamiltonians/BEC_twoComponent.hpp): coupled contact terms 4 pi (a11 n1 +
// a12 n2), 4 pi (a22 n2 + a12 n1) in the library's fused kernel.
#ifndef FGPE_HOST_BEC_TWO_COMPONENT_HPP
#define FGPE_HOST_BEC_TWO_COMPONENT_HPP

#include "../hpp/twoComponentGPSolver.hpp"

class BECtwoComponent : public TwoComponentGPSolver {
public:
    Params2 p;
    BECtwoComponent(Params2& p, int precision = FGPE_PREC_F64);
    void InitPsi();
    void InitVext();

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

#endif
