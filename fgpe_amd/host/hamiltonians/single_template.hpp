// single_template.hpp -- the plugin template (hamiltonians/single_template.*):
// kinetic term only; a user adds terms by editing the hparams block of the
// constructor (contact a, a(t) ramp, dipolar a_dd and direction, soft-core
// U, R), the initial state / trap, and any other pointwise term in the
// library's user functor (fgpe_amd/csrc/user_term.hpp, parameters
// hparams.user from the input keys user_prm0..7).  Status and output hooks
// are those of BEConeComponent.
#ifndef FGPE_HOST_SINGLE_TEMPLATE_HPP
#define FGPE_HOST_SINGLE_TEMPLATE_HPP

#include "BEC_oneComponent.hpp"

class SingleTemplate : public BEConeComponent {
public:
    SingleTemplate(Params& p, int precision = FGPE_PREC_F64);
};

#endif
