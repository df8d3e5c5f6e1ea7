// single_template.cpp -- see single_template.hpp
#include "single_template.hpp"

SingleTemplate::SingleTemplate(Params& par, int precision) : BEConeComponent(par, precision, false)
{
    hparams = fgpe_hparams{};
    hparams.ncomp = 1;
    /*
        MODIFY: define the terms entering your Hamiltonian, e.g.
        hparams.a[0] = p.a;                      contact 4 pi a |psi|^2
        hparams.a_dd = p.add;                    dipolar kernel along hparams.dip
        hparams.sc_U = ...; hparams.sc_R = ...;  soft-core kernel
    */
    run();
}
