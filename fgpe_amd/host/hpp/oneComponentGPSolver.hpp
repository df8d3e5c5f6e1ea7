// oneComponentGPSolver.hpp -- the one-component time loops and the plugin
// virtuals (hpp/oneComponentGPSolver.hpp:11-86).
#ifndef FGPE_HOST_ONE_COMPONENT_GP_SOLVER_HPP
#define FGPE_HOST_ONE_COMPONENT_GP_SOLVER_HPP

#include "baseHamiltonian.hpp"
#include "psi.hpp"
#include "vext.hpp"

class OneComponentGPSolver : public BaseHamiltonian {
protected:
    complex* psi;
    complex* vext;

public:
    Params p;
    double PARS[32] = {0.0};
    enum Variables { T, EN, MU, NORM, PX, LZ, EN_PREV, AVG_TIME, ITER, ITER_TIME, EN_0, MU_0, NCYCLES,
                     AUX1, AUX2, AUX3, AUX4, AUX5 };

    OneComponentGPSolver(Params& p, int precision = FGPE_PREC_F64);
    complex* getPsi() { return psi; }
    complex* getVext() { return vext; }

    void runImagTimeEvol();
    void runRealTimeEvol();


    // H is the library's fused kernel, configured by hparams (and a user term
    // of the library, hparams.user_id), and every observable applies it
    // itself, so calling these is a no-op.  A reference plugin that defines
    // its own alg_calcHpsi / alg_calcHpsiMU (hamiltonians/BEC_oneComponent.cu
    // :92-111) would have those kernels silently skipped: as final members
    // they make such a definition a compile error instead, so its terms move
    // into hparams or a user term.
    virtual void alg_calcHpsi() final {}
    virtual void alg_calcHpsiMU() final {}

    virtual void call_IM_init() = 0;
    virtual void call_IM_loop_before_step() = 0;
    virtual void call_IM_loop_after_step() = 0;
    virtual void call_IM_loop_convergence() = 0;
    virtual void call_IM_end() = 0;
    virtual void call_RE_init() = 0;
    virtual void call_RE_loop_before_step() = 0;
    virtual void call_RE_loop_after_step() = 0;
    virtual void call_RE_end() = 0;

    void syncState() override;
    // psi (host) <- device state (the D2H copies of the hooks)
    void downloadPsi();
    ~OneComponentGPSolver() override;

private:
    int pending = 0;
    double pendingT0 = 0.0;
    double dtCode = 0.0;
    void upload();
};

#endif
