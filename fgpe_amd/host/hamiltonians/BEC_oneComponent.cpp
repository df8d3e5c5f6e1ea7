// BEC_oneComponent.cpp -- constructor, initial state, trap and hooks of
// hamiltonians/BEC_oneComponent.cu:4-83, 189-397 (status rows, psi dumps,
// info.wtxt) over the host API.
#include "BEC_oneComponent.hpp"

#include <cstring>

#include "../hpp/device_init.hpp"

BEConeComponent::BEConeComponent(Params& par, int precision, bool do_run) : OneComponentGPSolver(par, precision)
{
    p = par;
    InitPsi();
    InitVext();
    // h_params[0] = p.a with the ramp of d_quench_as: prms[1..5] are zero, so
    // a(t) is prms[2] = 0 at every t, exactly as in the reference
    // (hamiltonians/BEC_oneComponent.cu:14, :167-178)
    hparams.ncomp = 1;
    hparams.quench_on = 1;
    hparams.quench[0] = p.a;
    // moving frame: alg_addVPx with 2 pi omega / (omho aho)
    // (hamiltonians/BEC_oneComponent.cu:98-101)
    if (p.getBool("trotate")) hparams.vpx[0] = 2 * pi * p.getDouble("omega") / (p.omho * p.aho);
    if (do_run) run();
}

void BEConeComponent::run()
{
    if (p.getBool("treal")) {
        std::cout << "# Running real time evolution" << std::endl;
        runRealTimeEvol();
    } else {
        std::cout << "# Running imaginary time evolution" << std::endl;
        runImagTimeEvol();
    }
}

BECDipolar::BECDipolar(Params& par, int precision) : BEConeComponent(par, precision, false)
{
    hparams.quench_on = 0;
    hparams.a[0] = p.a;
    hparams.a_dd = p.add;
    const InputVariables iv = p.getIV();
    hparams.dip[0] = iv.getVariable("dipole_x") != "None" ? p.getDouble("dipole_x") : 0.0;
    hparams.dip[1] = iv.getVariable("dipole_y") != "None" ? p.getDouble("dipole_y") : 0.0;
    hparams.dip[2] = iv.getVariable("dipole_z") != "None" ? p.getDouble("dipole_z") : 1.0;
    if (iv.getVariable("sc_U") != "None") {
        hparams.sc_U = p.getDouble("sc_U");
        hparams.sc_R = p.getDouble("sc_R") / p.aho;
        hparams.sc_dim = iv.getVariable("sc_dim") != "None" ? p.getInt("sc_dim") : p.DIM;
    }
    run();
}

void BEConeComponent::InitPsi()
{
    if (deviceInitPsi(p)) {
        devicePsi(ctx, p, 0, 0, p.getBool("tpi_vortex"), this->psi);
        output.WritePsiInit(this->psi, p.Npoints);
        return;
    }
    Psi psi_(p, "");
    if (p.getBool("tpi_vortex")) psi_.imprintVortexTube(p, p.getDouble("api_prc"));
    output.WritePsiInit(psi_.getPsi(), p.Npoints);
    std::memcpy(this->psi, psi_.getPsi(), p.Npoints * sizeof(complex));
}

void BEConeComponent::InitVext()
{
    if (deviceInitVext(p)) {
        // addProtocolPotential(p, 0): A = aseed
        const bool seed = p.getBool("tvseed");
        deviceVext(ctx, p, 0, 0, seed, seed ? p.getDouble("aseed") : 0.0, this->vext);
        output.WriteVext(this->vext, p.Npoints);
        return;
    }
    Vext vext_(p, "");
    if (p.getBool("tvseed")) vext_.addProtocolPotential(p, 0);
    output.WriteVext(vext_.getVext(), p.Npoints);
    std::memcpy(this->vext, vext_.getVext(), p.Npoints * sizeof(complex));
}

void BEConeComponent::call_IM_init()
{
    static char status_buffer[256];
    sprintf(status_buffer, "%-7s %-10s %-10s %-15s %-15s %-10s %-15s\n", "Iter.", "Time", "Norm", "mu", "Energy",
            "Ediff", "<Px>");
    output.CreateStatusFile(status_buffer);
    output.WritePsiInit(this->psi, p.Npoints);
}

void BEConeComponent::call_IM_loop_before_step() {}

void BEConeComponent::call_IM_loop_after_step()
{
    static char status_buffer[256];
    int iteration = (int)PARS[ITER];
    if (iteration % p.getInt("itStatMod") == 0) {
        this->alg_calcPx(&PARS[PX], this->d_psi, this->d_aux);
        PARS[PX] /= PARS[NORM] / p.aho;
        sprintf(status_buffer, "%-7d %-10.4f %-10.1f %-15.12f %-15.12f %-10.3E %-15.12f\n", (int)PARS[ITER],
                PARS[ITER_TIME], PARS[NORM], PARS[MU], PARS[EN], PARS[EN_PREV] - PARS[EN], PARS[PX] / hbar);
        output.WriteStatus(status_buffer);
    }
    if ((iteration + 1) % p.itmod == 0) {
        downloadPsi();
        output.WritePsi(this->psi, p.Npoints);
    }
}

void BEConeComponent::call_IM_loop_convergence()
{
    static char status_buffer[256];
    this->alg_calcPx(&PARS[PX], this->d_psi, this->d_aux);
    PARS[PX] /= PARS[NORM] / p.aho * hbar;
    sprintf(status_buffer, "%-7d %-10.4f %-10.1f %-15.12f %-15.12f %-10.3E %-15.12f\n", (int)PARS[ITER],
            PARS[ITER_TIME], PARS[NORM], PARS[MU], PARS[EN], PARS[EN_PREV] - PARS[EN], PARS[PX]);
    output.WriteStatus(status_buffer);
    std::cout << "Convergence reached: dE = " << PARS[EN_PREV] - PARS[EN] << std::endl;
}

void BEConeComponent::call_IM_end()
{
    output.WritePsiFinal(this->psi, p.Npoints);
    output.CloseStatusFile();
    output.WriteWtxt(p, 1);
    output.WriteVariable2WTXT("en", PARS[EN], "code");
    output.WriteVariable2WTXT("mu", PARS[MU], "code");
    output.WriteVariable2WTXT("px", PARS[PX], "hbar");
    output.WriteVariable2WTXT("convergence", PARS[EN_PREV] - PARS[EN], "code");
    std::cout << "Average iteration time: " << PARS[AVG_TIME] / PARS[ITER] << std::endl;
    std::cout << "Files saved under prefix: " << p.outpath << std::endl;
}

void BEConeComponent::call_RE_init()
{
    static char status_buffer[256];
    sprintf(status_buffer, "%-15s %-10s %-10s %-15s %-15s %-15s %-15s %-10s\n", "Time", "ItTime", "Norm", "mu",
            "Energy", "Px", "<cos>", "Ediff");
    output.CreateRealStatusFile(status_buffer);
    output.WritePsiInit(this->psi, p.Npoints);
    PARS[NORM] = p.npart;
    this->alg_calcHpsi();
    this->alg_calc2Observables(PARS[NORM], &PARS[MU_0], &PARS[EN_0], this->d_psi, this->d_hpsi, this->d_hpsi_en);
    this->alg_calcNorm(&PARS[NORM], this->d_psi);
    this->alg_calcCos(&PARS[AUX1], this->d_psi);
    PARS[AUX1] /= PARS[NORM];
    this->alg_calcPx(&PARS[PX], this->d_psi, this->d_aux);
    PARS[PX] /= (PARS[NORM] * hbar) / p.aho;
    sprintf(status_buffer, "%-15.5f %-10.4f %-10.1f %-15.12f %-15.12f %-15.12f %-15.12f %-10.3E\n", 0.0, 0.0,
            PARS[NORM], PARS[MU_0], PARS[EN_0], PARS[PX], PARS[AUX1], PARS[EN_0] - PARS[EN]);
    output.WriteStatusReal(status_buffer);
}

void BEConeComponent::call_RE_loop_before_step() {}

void BEConeComponent::call_RE_loop_after_step()
{
    static char status_buffer[256];
    if (((int)PARS[ITER]) % p.getInt("itStatMod") == 0) {
        this->alg_calcNorm(&PARS[NORM], this->d_psi);
        this->alg_calc2Observables(PARS[NORM], &PARS[MU], &PARS[EN], this->d_psi, this->d_hpsi, this->d_hpsi_en);
        this->alg_calcCos(&PARS[AUX1], this->d_psi);
        PARS[AUX1] /= PARS[NORM];
        this->alg_calcPx(&PARS[PX], this->d_psi, this->d_aux);
        PARS[PX] /= (PARS[NORM] * hbar) / p.aho;
        double time = p.time0 + PARS[ITER] * p.dt;
        sprintf(status_buffer, "%-15.5f %-10.4f %-10.1f %-15.12f %-15.12f %-15.12f %-15.12f %-10.3E\n", time,
                PARS[ITER_TIME], PARS[NORM], PARS[MU], PARS[EN], PARS[PX], PARS[AUX1], PARS[EN] - PARS[EN_0]);
        output.WriteStatusReal(status_buffer);
    }
    if (((int)PARS[ITER] + 1) % p.itmod == 0) {
        downloadPsi();
        output.WritePsi(this->psi, p.Npoints);
        PARS[NCYCLES]++;
    }
}

void BEConeComponent::call_RE_end()
{
    output.WritePsiFinal(this->psi, p.Npoints);
    output.CloseRealStatusFile();
    output.WriteWtxt(p, PARS[NCYCLES]);
    output.WriteVariable2WTXT("en0", PARS[EN_0], "code");
    output.WriteVariable2WTXT("mu0", PARS[MU_0], "code");
    output.WriteVariable2WTXT("en", PARS[EN], "code");
    output.WriteVariable2WTXT("mu", PARS[MU_0], "code");
    output.WriteVariable2WTXT("px", PARS[PX], "hbar");
    std::cout << "Average iteration time: " << PARS[AVG_TIME] / PARS[ITER] << std::endl;
    std::cout << "Files saved under prefix: " << p.outpath << std::endl;
}
