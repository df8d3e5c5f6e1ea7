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
        Any other pointwise term is written once in the library's
        fgpe_amd/csrc/user_term.hpp (a functor listed in FGPE_USER_TERMS
        with an id, rebuild libfgpe.so); hparams.user_id selects it, here
        from the input key user_term (default 1), and its eight parameters
        come from hparams.user, here from the input keys user_prm0 ...
        user_prm7 when present.
    */
    const InputVariables iv = p.getIV();
    hparams.user_id = iv.getVariable("user_term") != "None" ? p.getInt("user_term") : 1;
    for (int i = 0; i < 8; ++i) {
        const std::string key = "user_prm" + std::to_string(i);
        if (iv.getVariable(key) != "None") hparams.user[i] = p.getDouble(key);
    }
    run();
}
