// BEC_twoComponent.cpp -- constructor, initial states, traps and hooks of
// hamiltonians/BEC_twoComponent.cu:4-124, 250-493 over the host API.
#include "BEC_twoComponent.hpp"

#include <cstring>

#include "../hpp/device_init.hpp"

BECtwoComponent::BECtwoComponent(Params2& par, int precision) : TwoComponentGPSolver(par, precision)
{
    p = par;
    InitPsi();
    InitVext();
    hparams.ncomp = 2;
    hparams.a[0] = p.a11;
    hparams.a[1] = p.a22;
    hparams.a[2] = p.a12;
    // moving frame per component (hamiltonians/BEC_twoComponent.cu:125-129)
    if (p.getBool("trotate")) {
        hparams.vpx[0] = 2 * pi * p.getDouble("omega") / (p.omho1 * p.aho1);
        hparams.vpx[1] = 2 * pi * p.getDouble("omega") / (p.omho2 * p.aho2);
    }
    if (p.getBool("treal")) {
        std::cout << "# Running real time evolution" << std::endl;
        runRealTimeEvol();
    } else {
        std::cout << "# Running imaginary time evolution" << std::endl;
        runImagTimeEvol();
    }
}

void BECtwoComponent::InitPsi()
{
    if (deviceInitPsi(p)) {
        devicePsi(ctx, p, 0, 1, p.getBool("tpi_vortex"), this->psi1);
        devicePsi(ctx, p, 1, 2, p.getBool("tpi_vortex"), this->psi2);
        return;
    }
    Psi psi_1(p, 1);
    Psi psi_2(p, 2);
    if (p.getBool("tpi_vortex")) {
        psi_1.imprintVortexTube(p, p.getDouble("api_prc"));
        psi_2.imprintVortexTube(p, p.getDouble("api_prc"));
    }
    std::memcpy(this->psi1, psi_1.getPsi(), p.Npoints * sizeof(complex));
    std::memcpy(this->psi2, psi_2.getPsi(), p.Npoints * sizeof(complex));
}

void BECtwoComponent::InitVext()
{
    if (deviceInitVext(p)) {
        const bool seed = p.getBool("tvseed");
        for (int c = 1; c <= 2; ++c)
            deviceVext(ctx, p, c - 1, c, seed, seed ? Vext::weightedProtocolAmplitude(p, c) : 0.0,
                       c == 1 ? this->vext1 : this->vext2);
        output.WriteVext(this->vext1, p.Npoints, "1");
        output.WriteVext(this->vext2, p.Npoints, "2");
        return;
    }
    Vext vext_1(p, 1);
    Vext vext_2(p, 2);
    if (p.getBool("tvseed")) {
        vext_1.addWeightedProtocolPotential(p, 1);
        vext_2.addWeightedProtocolPotential(p, 2);
    }
    output.WriteVext(vext_1.getVext(), p.Npoints, "1");
    output.WriteVext(vext_2.getVext(), p.Npoints, "2");
    std::memcpy(this->vext1, vext_1.getVext(), p.Npoints * sizeof(complex));
    std::memcpy(this->vext2, vext_2.getVext(), p.Npoints * sizeof(complex));
}

void BECtwoComponent::call_IM_init()
{
    static char status_buffer[256];
    sprintf(status_buffer, "%-7s %-10s %-10s %-15s %-15s %-15s %-10s %-15s %-15s %-15s %-10s %-10s\n", "Iter.",
            "Time", "Norm1", "mu1", "Energy1", "Px1", "Norm2", "mu2", "Energy2", "Px2", "Ediff1", "Ediff2");
    output.CreateStatusFile(status_buffer);
    output.WritePsiInit(this->psi1, p.Npoints, "1");
    output.WritePsiInit(this->psi2, p.Npoints, "2");
}

void BECtwoComponent::call_IM_loop_before_step() {}

void BECtwoComponent::call_IM_loop_after_step()
{
    static char status_buffer[256];
    int iteration = (int)PARS[ITER];
    if (iteration % p.getInt("itStatMod") == 0) {
        this->alg_calcPx(&PARS[PX1], this->d_psi, this->d_aux);
        this->alg_calcPx(&PARS[PX2], this->d_psi2, this->d_aux2);
        PARS[PX1] /= PARS[NORM1] / p.aho1;
        PARS[PX2] /= PARS[NORM2] / p.aho1;
        sprintf(status_buffer,
                "%-7d %-10.4f %-10.1f %-15.12f %-15.12f %-15.12f %-10.1f %-15.12f %-15.12f %-15.12f %-10.3E %-10.3E\n",
                (int)PARS[ITER], PARS[ITER_TIME], PARS[NORM1], PARS[MU1], PARS[EN1], PARS[PX1] / hbar, PARS[NORM2],
                PARS[MU2], PARS[EN2], PARS[PX2] / hbar, PARS[EN1_PREV] - PARS[EN1], PARS[EN2_PREV] - PARS[EN2]);
        output.WriteStatus(status_buffer);
    }
    if ((iteration + 1) % p.itmod == 0) {
        downloadPsi();
        output.WritePsi(this->psi1, p.Npoints);
        output.WritePsi(this->psi2, p.Npoints);
        PARS[NCYCLES]++;
    }
}

void BECtwoComponent::call_IM_loop_convergence()
{
    static char status_buffer[256];
    this->alg_calcPx(&PARS[PX1], this->d_psi, this->d_aux);
    this->alg_calcPx(&PARS[PX2], this->d_psi2, this->d_aux2);
    PARS[PX1] /= PARS[NORM1] / p.aho1;
    PARS[PX2] /= PARS[NORM2] / p.aho1;
    sprintf(status_buffer,
            "%-7d %-10.4f %-10.1f %-15.12f %-15.12f %-15.12f %-10.1f %-15.12f %-15.12f %-15.12f %-10.3E %-10.3E\n",
            (int)PARS[ITER], PARS[ITER_TIME], PARS[NORM1], PARS[MU1], PARS[EN1], PARS[PX1] / hbar, PARS[NORM2],
            PARS[MU2], PARS[EN2], PARS[PX2] / hbar, PARS[EN1_PREV] - PARS[EN1], PARS[EN2_PREV] - PARS[EN2]);
    output.WriteStatus(status_buffer);
    std::cout << "Convergence reached: dE1 = " << PARS[EN1_PREV] - PARS[EN1] << std::endl;
    std::cout << "                     dE2 = " << PARS[EN2_PREV] - PARS[EN2] << std::endl;
}

void BECtwoComponent::call_IM_end()
{
    output.WritePsiFinal(this->psi1, p.Npoints, "1");
    output.WritePsiFinal(this->psi2, p.Npoints, "2");
    output.CloseStatusFile();
    output.WriteWtxt_2c(p, PARS[NCYCLES]);
    output.WriteVariable2WTXT("en", PARS[EN1], "code");
    output.WriteVariable2WTXT("en", PARS[EN2], "code");
    output.WriteVariable2WTXT("mu", PARS[MU1], "code");
    output.WriteVariable2WTXT("mu", PARS[MU2], "code");
    output.WriteVariable2WTXT("px", PARS[PX1], "hbar");
    output.WriteVariable2WTXT("px", PARS[PX2], "hbar");
    output.WriteVariable2WTXT("convergence1", PARS[EN1_PREV] - PARS[EN1], "code");
    output.WriteVariable2WTXT("convergence2", PARS[EN2_PREV] - PARS[EN2], "code");
    std::cout << "Average iteration time: " << PARS[AVG_TIME] / PARS[ITER] << std::endl;
    std::cout << "Files saved under prefix: " << p.outpath << std::endl;
}

void BECtwoComponent::call_RE_init()
{
    static char status_buffer[256];
    sprintf(status_buffer, "%-15s %-10s %-10s %-15s %-15s %-15s %-10s %-15s %-15s  %-15s %-15s %-15s %-10s %-10s\n",
            "Time", "ItTime", "Norm1", "mu1", "Energy1", "Px1", "Norm2", "mu2", "Energy2", "Px2", "<cos_1>",
            "<cos_2>", "Ediff1", "Ediff2");
    output.CreateRealStatusFile(status_buffer);
    output.WriteWtxt_2c(p, 0);
    output.WritePsiInit(this->psi1, p.Npoints, "1");
    output.WritePsiInit(this->psi2, p.Npoints, "2");
    PARS[NORM1] = p.npart1;
    PARS[NORM2] = p.npart2;
    this->alg_calcHpsi();
    this->alg_calc2Observables(PARS[NORM1], &PARS[MU1_0], &PARS[EN1_0], this->d_psi, this->d_hpsi, this->d_hpsi_en);
    this->alg_calc2Observables(PARS[NORM2], &PARS[MU2_0], &PARS[EN2_0], this->d_psi2, this->d_hpsi2,
                               this->d_hpsi2_en);
    this->alg_calcNorm(&PARS[NORM1], this->d_psi);
    this->alg_calcNorm(&PARS[NORM2], this->d_psi2);
    this->alg_calcCos(&PARS[AUX1_1], this->d_psi);
    PARS[AUX1_1] /= PARS[NORM1];
    this->alg_calcCos(&PARS[AUX1_2], this->d_psi2);
    PARS[AUX1_2] /= PARS[NORM2];
    this->alg_calcPx(&PARS[PX1], this->d_psi, this->d_aux);
    this->alg_calcPx(&PARS[PX2], this->d_psi2, this->d_aux2);
    PARS[PX1] /= (PARS[NORM1] * hbar) / p.aho1;
    PARS[PX2] /= (PARS[NORM2] * hbar) / p.aho1;
    sprintf(status_buffer,
            "%-15.5f %-10.4f %-10.1f %-15.12f %-15.12f %-15.12f %-10.1f %-15.12f %-15.12f %-15.12f %-15.12f %-15.12f "
            "%-10.3E %-10.3E\n",
            0.00, 0.00, PARS[NORM1], PARS[MU1_0], PARS[EN1_0], PARS[PX1], PARS[NORM2], PARS[MU2_0], PARS[EN2_0],
            PARS[PX2], PARS[AUX1_1], PARS[AUX1_2], PARS[EN1] - PARS[EN1_0], PARS[EN2] - PARS[EN2_0]);
    output.WriteStatusReal(status_buffer);
}

void BECtwoComponent::call_RE_loop_before_step() {}

void BECtwoComponent::call_RE_loop_after_step()
{
    static char status_buffer[256];
    if (((int)PARS[ITER]) % p.getInt("itStatMod") == 0) {
        this->alg_calcNorm(&PARS[NORM1], this->d_psi);
        this->alg_calcNorm(&PARS[NORM2], this->d_psi2);
        this->alg_calc2Observables(PARS[NORM1], &PARS[MU1], &PARS[EN1], this->d_psi, this->d_hpsi, this->d_hpsi_en);
        this->alg_calc2Observables(PARS[NORM2], &PARS[MU2], &PARS[EN2], this->d_psi2, this->d_hpsi2,
                                   this->d_hpsi2_en);
        this->alg_calcCos(&PARS[AUX1_1], this->d_psi);
        PARS[AUX1_1] /= PARS[NORM1];
        this->alg_calcCos(&PARS[AUX1_2], this->d_psi2);
        PARS[AUX1_2] /= PARS[NORM2];
        this->alg_calcPx(&PARS[PX1], this->d_psi, this->d_aux);
        this->alg_calcPx(&PARS[PX2], this->d_psi2, this->d_aux);
        PARS[PX1] /= (PARS[NORM1] * hbar) / p.aho1;
        PARS[PX2] /= (PARS[NORM2] * hbar) / p.aho1;
        double time = p.time0 + PARS[ITER] * p.dt;
        sprintf(status_buffer,
                "%-15.5f %-10.4f %-10.1f %-15.12f %-15.12f %-15.12f %-10.1f %-15.12f %-15.12f %-15.12f %-15.12f "
                "%-15.12f %-10.3E %-10.3E\n",
                time, PARS[ITER_TIME], PARS[NORM1], PARS[MU1], PARS[EN1], PARS[PX1], PARS[NORM2], PARS[MU2],
                PARS[EN2], PARS[PX2], PARS[AUX1_1], PARS[AUX1_2], PARS[EN1] - PARS[EN1_0], PARS[EN2] - PARS[EN2_0]);
        output.WriteStatusReal(status_buffer);
    }
    if (((int)PARS[ITER] + 1) % p.itmod == 0) {
        downloadPsi();
        output.WritePsi(this->psi1, p.Npoints, "1");
        output.WritePsi(this->psi2, p.Npoints, "2");
        PARS[NCYCLES]++;
    }
}

void BECtwoComponent::call_RE_end()
{
    output.WritePsiFinal(this->psi1, p.Npoints, "1");
    output.WritePsiFinal(this->psi2, p.Npoints, "2");
    output.CloseRealStatusFile();
    output.WriteWtxt_2c(p, PARS[NCYCLES]);
    output.WriteVariable2WTXT("en10", PARS[EN1_0], "code");
    output.WriteVariable2WTXT("en20", PARS[EN2_0], "code");
    output.WriteVariable2WTXT("mu10", PARS[MU1_0], "code");
    output.WriteVariable2WTXT("mu20", PARS[MU2_0], "code");
    output.WriteVariable2WTXT("en1", PARS[EN1], "code");
    output.WriteVariable2WTXT("en1", PARS[EN2], "code");
    output.WriteVariable2WTXT("mu1", PARS[MU1_0], "code");
    output.WriteVariable2WTXT("mu2", PARS[MU2_0], "code");
    output.WriteVariable2WTXT("px1", PARS[PX1], "hbar");
    output.WriteVariable2WTXT("px2", PARS[PX2], "hbar");
    std::cout << "Average iteration time: " << PARS[AVG_TIME] / PARS[ITER] << std::endl;
    std::cout << "Files saved under prefix: " << p.outpath << std::endl;
}
