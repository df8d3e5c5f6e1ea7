// user_term.hpp -- the user's pointwise Hamiltonian term.
//
// This is the one place a custom Hamiltonian is written, the counterpart of
// the kernels a plugin author edits in the reference
// (SingleTemplate_template / SingleTemplate_MU_template,
// hamiltonians/single_template.cu:333-373).  The library compiles
// UserTerm::eval into the fused pointwise stage of every H application (RK4
// stages, heavy-ball iterations, fgpe_apply_h and the observables), where it
// adds
//     w psi_c      to  H psi_c            (w = wr + i wi)
//     w_en psi_c   to  the energy operator (w_en = er + i ei)
// on top of the kinetic, trap, contact and k-space terms the library already
// has.  Its inputs are the stage time, the grid index, the component, the
// stage state psi_c, the densities of both components and the eight
// parameters fgpe_hparams::user.  A plugin sets those parameters; editing
// eval and rebuilding libfgpe.so changes the physics.
//
// The shipped library has no user term (enabled = false): eval is not
// compiled into the kernels and costs nothing.  Building with
// -DFGPE_USER_TERM_EXAMPLE selects the example below (the build recipe makes
// that variant as libfgpe_user_example.so for the tests):
//     w  = dg(t) n_c + s iz - i (K3 / 2) n_c^2
//     w_en = dg(t) n_c / 2 + s iz - i (K3 / 2) n_c^2
// with dg(t) = user[0] sin(user[1] t) (a modulated contact coupling),
// K3 = user[2] (three-body loss) and s = user[3] (a linear ramp along z in
// grid cells).
#pragma once
#include <hip/hip_runtime.h>

namespace fgpe {

struct UserTermIn {
    double t;           // time of this H application
    int ix, iy, iz;     // grid index, x fastest
    int comp;           // component 0 or 1
    double pr, pi;      // psi_c at this point (the stage state)
    double n;           // |psi_c|^2
    double n_other;     // |psi_c'|^2 of the other component (0 with one)
    const double* prm;  // fgpe_hparams::user[0..7]
};

struct UserTermOut {
    double wr = 0.0, wi = 0.0;   // added to H:       H psi_c += w psi_c
    double er = 0.0, ei = 0.0;   // added to H_en:  H_en psi_c += w_en psi_c
};

struct UserTerm {
#ifdef FGPE_USER_TERM_EXAMPLE
    static constexpr bool enabled = true;
    __device__ static UserTermOut eval(const UserTermIn& in)
    {
        UserTermOut o;
        const double dg = in.prm[0] * sin(in.prm[1] * in.t);
        const double loss = -0.5 * in.prm[2] * in.n * in.n;
        const double ramp = in.prm[3] * in.iz;
        o.wr = dg * in.n + ramp;
        o.wi = loss;
        o.er = 0.5 * dg * in.n + ramp;
        o.ei = loss;
        return o;
    }
#else
    static constexpr bool enabled = false;
    __device__ static UserTermOut eval(const UserTermIn&) { return UserTermOut{}; }
#endif
};

}  // namespace fgpe
