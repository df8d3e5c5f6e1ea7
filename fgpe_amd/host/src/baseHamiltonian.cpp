// baseHamiltonian.cpp -- context set-up (Cuda::Cuda, src/parallel.cu:14-113),
// k and x axes (src/baseHamiltonian.cu:25-82) and the reductions through the
// C ABI.
#include "../hpp/baseHamiltonian.hpp"

void FCE(int rc, const char* what)
{
    if (rc != FGPE_OK) {
        std::cerr << "fgpe error " << rc << " in " << what << ": " << fgpe_last_error() << std::endl;
        exit(4);
    }
}

BaseHamiltonian::BaseHamiltonian(ParamsBase* par, int prec) : pars(par), output(par->outpath), precision(prec)
{
    int nz = par->NX[2], ny = par->NX[1];
    FCE(fgpe_ctx_create(par->NX[0], ny, nz, precision, &ctx), "fgpe_ctx_create");
    InitializeKspace();
    InitializeSpace();
}

BaseHamiltonian::~BaseHamiltonian()
{
    fgpe_ctx_destroy(ctx);
    free(x); free(y); free(z);
    free(kx); free(ky); free(kz);
}

void BaseHamiltonian::InitializeSpace()
{
    const int* NX = pars->NX;
    x = (double*)malloc(NX[0] * sizeof(double));
    y = (double*)malloc(NX[1] * sizeof(double));
    z = (double*)malloc(NX[2] * sizeof(double));
    for (int i = 0; i < NX[0]; ++i) x[i] = -pars->XMAX[0] + i * pars->DX[0] + 0.5 * pars->DX[0];
    for (int i = 0; i < NX[1]; ++i) y[i] = pars->DIM >= 2 ? -pars->XMAX[1] + i * pars->DX[1] + 0.5 * pars->DX[1] : 0;
    for (int i = 0; i < NX[2]; ++i) z[i] = pars->DIM == 3 ? -pars->XMAX[2] + i * pars->DX[2] + 0.5 * pars->DX[2] : 0;
}

void BaseHamiltonian::InitializeKspace()
{
    const int* NX = pars->NX;
    double** k[3] = {&kx, &ky, &kz};
    for (int a = 0; a < 3; ++a) {
        *k[a] = (double*)malloc(NX[a] * sizeof(double));
        double* kk = *k[a];
        for (int i = 0; i <= NX[a] / 2 && i < NX[a]; ++i) kk[i] = i * pars->DKX[a];
        for (int i = NX[a] / 2 + 1; i < NX[a]; ++i) kk[i] = -pars->KXMAX[a] + (i - (NX[a] / 2)) * pars->DKX[a];
    }
    if (pars->DIM == 1) {
        for (int i = 0; i < NX[1]; ++i) ky[i] = 0;
        for (int i = 0; i < NX[2]; ++i) kz[i] = 0;
    } else if (pars->DIM == 2) {
        for (int i = 0; i < NX[2]; ++i) kz[i] = 0;
    }
    FCE(fgpe_set_kspace(ctx, kx, ky, kz, pars->DV), "fgpe_set_kspace");
}

void BaseHamiltonian::currentObservables()
{
    syncState();
    if (!obsValid) {
        FCE(fgpe_observables(ctx, obsTime, obs), "fgpe_observables");
        obsValid = true;
    }
}

void BaseHamiltonian::alg_calcNorm(double* norm, DeviceArray psi)
{
    currentObservables();
    *norm = obs[3 * psi.comp];
}

// o = sum Re(psi* O psi) DV / norm (src/baseHamiltonian.cu:230-236).  The
// operators are H and Hen of the current state at obsTime.
void BaseHamiltonian::alg_calc2Observables(double norm, double* o1, double* o2, DeviceArray psi, DeviceArray,
                                           DeviceArray)
{
    currentObservables();
    const int c = psi.comp;
    *o1 = obs[3 * c + 1] * obs[3 * c] / norm;
    *o2 = obs[3 * c + 2] * obs[3 * c] / norm;
}

void BaseHamiltonian::alg_calcPx(double* px, DeviceArray psi, DeviceArray)
{
    syncState();
    FCE(fgpe_calc_px(ctx, psi.comp, px), "fgpe_calc_px");
}

void BaseHamiltonian::alg_calcCosine(double* avgcos, DeviceArray psi)
{
    syncState();
    FCE(fgpe_calc_cos(ctx, psi.comp, x, 2 * pars->XMAX[0], avgcos), "fgpe_calc_cos");
}
