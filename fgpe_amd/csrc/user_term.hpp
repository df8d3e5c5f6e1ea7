// user_term.hpp -- user pointwise Hamiltonian terms.
//
// This is where a custom Hamiltonian is written, the counterpart of the
// kernels a plugin author edits in the reference (SingleTemplate_template /
// SingleTemplate_MU_template, hamiltonians/single_template.cu:333-373; a
// reference binary holds several plugins and main.cu picks one by name,
// main.cu:41-69).  A library build compiles a list of terms, each with an id
// >= 1; fgpe_hparams::user_id picks one per context (0: none).  The chosen
// term's eval runs inside the fused pointwise stage of every H application
// (RK4 stages, heavy-ball iterations, fgpe_apply_h and the observables),
// where it adds
//     w psi_c      to  H psi_c            (w = wr + i wi)
//     w_en psi_c   to  the energy operator (w_en = er + i ei)
// on top of the kinetic, trap, contact and k-space terms the library already
// has.  Its inputs are the stage time, the grid index, the component, the
// stage state psi_c, the densities of both components and the eight
// parameters fgpe_hparams::user.
//
// The x stage is compiled once per listed id (and once without a term), so a
// context without a term runs the plain kernels.  The shipped library lists
// no term.  Building with -DFGPE_USER_TERM_EXAMPLE lists the two example
// terms below (the build recipe makes that variant as libfgpe_user_example.so
// for the tests).
//
// To add a term: write a struct with `static constexpr bool enabled = true`
// and a __device__ `eval`, and add X(id, Struct) to FGPE_USER_TERMS.
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

// Example 1: a modulated contact coupling, three-body loss and a linear ramp
// along z,
//     w    = dg(t) n_c + s iz - i (K3 / 2) n_c^2
//     w_en = dg(t) n_c / 2 + s iz - i (K3 / 2) n_c^2
// with dg(t) = user[0] sin(user[1] t), K3 = user[2], s = user[3].
struct ExampleTerm {
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
};

// Example 2: a decaying optical lattice in the xy plane that the two
// components feel with opposite signs,
//     w = w_en = sgn_c A e^(-gamma t) cos(kx ix + ky iy)
// with A = user[4], gamma = user[5], kx = user[6], ky = user[7] (per grid
// cell), sgn_0 = +1, sgn_1 = -1.
struct DriveTerm {
    static constexpr bool enabled = true;
    __device__ static UserTermOut eval(const UserTermIn& in)
    {
        UserTermOut o;
        const double sgn = in.comp == 0 ? 1.0 : -1.0;
        const double w = sgn * in.prm[4] * exp(-in.prm[5] * in.t) * cos(in.prm[6] * in.ix + in.prm[7] * in.iy);
        o.wr = w;
        o.er = w;
        return o;
    }
};

#ifdef FGPE_USER_TERM_EXAMPLE
#define FGPE_USER_TERMS(X) X(1, ExampleTerm) X(2, DriveTerm)
#else
#define FGPE_USER_TERMS(X)
#endif

// UserTerm<id>: the term of that id; UserTerm<0> is none.
struct NoTerm {
    static constexpr bool enabled = false;
    __device__ static UserTermOut eval(const UserTermIn&) { return UserTermOut{}; }
};
template <int ID> struct UserTerm : NoTerm {};
#define FGPE_USER_TERM_SPEC(ID, S) \
    template <> struct UserTerm<ID> : S {};
FGPE_USER_TERMS(FGPE_USER_TERM_SPEC)
#undef FGPE_USER_TERM_SPEC

// the ids this build compiles (host side)
inline bool user_term_compiled(int id)
{
#define FGPE_USER_TERM_HAS(ID, S) if (id == ID) return true;
    FGPE_USER_TERMS(FGPE_USER_TERM_HAS)
#undef FGPE_USER_TERM_HAS
    return false;
}
inline int user_term_count()
{
    int n = 0;
#define FGPE_USER_TERM_COUNT(ID, S) ++n;
    FGPE_USER_TERMS(FGPE_USER_TERM_COUNT)
#undef FGPE_USER_TERM_COUNT
    return n;
}

}  // namespace fgpe
