// solvers.cpp -- the time loops of src/oneComponentGPSolver.cu:26-217 and
// src/twoComponentGPSolver.cu:15-209 over the C ABI.
//
// Real time: each loop iteration queues one RK4 step and calls the hooks; the
// queued steps run as one fgpe_rk4_steps launch sequence the moment a hook
// (or the end of the loop) needs device state, so hooks see exactly the
// state of their iteration and the loop itself never waits on the device.
// Imaginary time: heavy ball on the device in chunks that end where the
// hooks write status rows (itStatMod) or dump psi (itmod); per-iteration
// norm, mu and E come back from the device history and are handed to the
// hooks in order, and the convergence test is the device's.
#include "../hpp/oneComponentGPSolver.hpp"
#include "../hpp/twoComponentGPSolver.hpp"

#include <vector>

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// iterations of the next imaginary-time chunk starting at `it`: it ends on
// the first iteration that a status row (it % stat == 0) or a psi dump
// ((it + 1) % itmod == 0) falls on
int chunk_len(int it, int niter, int stat, int itmod)
{
    int end = niter - 1;
    for (int j = it; j < niter; ++j) {
        if ((stat > 0 && j % stat == 0) || (itmod > 0 && (j + 1) % itmod == 0)) {
            end = j;
            break;
        }
    }
    return end - it + 1;
}

int opt_int(const ParamsBase& p, const char* key, int dflt)
{
    return p.getIV().getVariable(key) != "None" ? p.getInt(key) : dflt;
}

}  // namespace

// ---------------------------------------------------------------------------
OneComponentGPSolver::OneComponentGPSolver(Params& par, int precision) : BaseHamiltonian(&par, precision)
{
    p = par;
    pars = &p;
    psi = (complex*)malloc(p.Npoints * sizeof(complex));
    vext = (complex*)malloc(p.Npoints * sizeof(complex));
}

OneComponentGPSolver::~OneComponentGPSolver()
{
    free(psi);
    free(vext);
}

void OneComponentGPSolver::upload()
{
    FCE(fgpe_set_hamiltonian(ctx, &hparams), "fgpe_set_hamiltonian");
    FCE(fgpe_upload_vext(ctx, 0, (const double*)vext), "fgpe_upload_vext");
    FCE(fgpe_upload_psi(ctx, 0, (const double*)psi), "fgpe_upload_psi");
    invalidate();
}

void OneComponentGPSolver::syncState()
{
    if (pending == 0) return;
    const double t0 = now_ms();
    FCE(fgpe_rk4_steps(ctx, dtCode, pendingT0, pending), "fgpe_rk4_steps");
    FCE(fgpe_ctx_sync(ctx), "fgpe_ctx_sync");
    const double el = now_ms() - t0;
    PARS[ITER_TIME] = el / pending;
    PARS[AVG_TIME] += el;
    pending = 0;
    invalidate();
}

void OneComponentGPSolver::downloadPsi()
{
    syncState();
    FCE(fgpe_download_psi(ctx, 0, (double*)psi), "fgpe_download_psi");
}

void OneComponentGPSolver::runImagTimeEvol()
{
    call_IM_init();
    upload();
    const double dt = p.dt * p.omho / 1000;
    const double beta = p.getDouble("beta");
    const int stat = opt_int(p, "itStatMod", 0);
    std::vector<double> hist;
    int it = 0;
    bool done_all = false;
    while (it < p.niter && !done_all) {
        const int n = chunk_len(it, p.niter, stat, p.itmod);
        hist.assign(3 * (size_t)n, 0.0);
        int done = 0, conv = 0;
        const double t0 = now_ms();
        FCE(fgpe_imag_steps(ctx, FGPE_IMAG_HEAVY_BALL, dt, beta, &p.npart, p.epsilon_e, n, hist.data(), &done),
            "fgpe_imag_steps");
        FCE(fgpe_imag_converged(ctx, &conv), "fgpe_imag_converged");
        const double per = done > 0 ? (now_ms() - t0) / done : 0.0;
        invalidate();
        for (int j = 0; j < done; ++j) {
            PARS[EN_PREV] = PARS[EN];
            PARS[T] = (it + j) * dt;
            obsTime = PARS[T];
            call_IM_loop_before_step();
            PARS[NORM] = p.npart;
            PARS[MU] = hist[3 * j + 1];
            PARS[EN] = hist[3 * j + 2];
            PARS[AVG_TIME] += per;
            PARS[ITER_TIME] = per;
            if (conv && j == done - 1) {
                call_IM_loop_convergence();
                done_all = true;
                break;
            }
            call_IM_loop_after_step();
            PARS[ITER]++;
        }
        it += done;
        if (done < n) done_all = true;
    }
    FCE(fgpe_download_psi(ctx, 0, (double*)psi), "fgpe_download_psi");
    call_IM_end();
}

void OneComponentGPSolver::runRealTimeEvol()
{
    upload();
    dtCode = p.dt * p.omho / 1000;
    call_RE_init();
    PARS[T] = 0;
    PARS[AVG_TIME] = 0;
    for (int iteration = 0; iteration < p.niter; ++iteration) {
        PARS[ITER]++;
        call_RE_loop_before_step();
        if (pending == 0) pendingT0 = iteration * dtCode;
        pending++;
        PARS[T] = (iteration + 1) * dtCode;
        obsTime = PARS[T];
        invalidate();
        call_RE_loop_after_step();
    }
    downloadPsi();
    call_RE_end();
}

// ---------------------------------------------------------------------------
TwoComponentGPSolver::TwoComponentGPSolver(Params2& par, int precision) : BaseHamiltonian(&par, precision)
{
    p = par;
    pars = &p;
    psi1 = (complex*)malloc(p.Npoints * sizeof(complex));
    psi2 = (complex*)malloc(p.Npoints * sizeof(complex));
    vext1 = (complex*)malloc(p.Npoints * sizeof(complex));
    vext2 = (complex*)malloc(p.Npoints * sizeof(complex));
}

TwoComponentGPSolver::~TwoComponentGPSolver()
{
    free(psi1);
    free(psi2);
    free(vext1);
    free(vext2);
}

void TwoComponentGPSolver::upload()
{
    FCE(fgpe_set_hamiltonian(ctx, &hparams), "fgpe_set_hamiltonian");
    FCE(fgpe_upload_vext(ctx, 0, (const double*)vext1), "fgpe_upload_vext");
    FCE(fgpe_upload_vext(ctx, 1, (const double*)vext2), "fgpe_upload_vext");
    FCE(fgpe_upload_psi(ctx, 0, (const double*)psi1), "fgpe_upload_psi");
    FCE(fgpe_upload_psi(ctx, 1, (const double*)psi2), "fgpe_upload_psi");
    invalidate();
}

void TwoComponentGPSolver::syncState()
{
    if (pending == 0) return;
    const double t0 = now_ms();
    FCE(fgpe_rk4_steps(ctx, dtCode, pendingT0, pending), "fgpe_rk4_steps");
    FCE(fgpe_ctx_sync(ctx), "fgpe_ctx_sync");
    const double el = now_ms() - t0;
    PARS[ITER_TIME] = el / pending;
    PARS[AVG_TIME] += el;
    pending = 0;
    invalidate();
}

void TwoComponentGPSolver::downloadPsi()
{
    syncState();
    FCE(fgpe_download_psi(ctx, 0, (double*)psi1), "fgpe_download_psi");
    FCE(fgpe_download_psi(ctx, 1, (double*)psi2), "fgpe_download_psi");
}

void TwoComponentGPSolver::runImagTimeEvol()
{
    call_IM_init();
    upload();
    const double dt = p.dt * p.omho1 / 1000;
    const double beta = p.getDouble("beta");
    const int stat = opt_int(p, "itStatMod", 0);
    std::vector<double> hist;
    int it = 0;
    bool done_all = false;
    while (it < p.niter && !done_all) {
        const int n = chunk_len(it, p.niter, stat, p.itmod);
        hist.assign(6 * (size_t)n, 0.0);
        int done = 0, conv = 0;
        const double t0 = now_ms();
        const double np2[2] = {p.npart1, p.npart2};
        FCE(fgpe_imag_steps(ctx, FGPE_IMAG_HEAVY_BALL, dt, beta, np2, p.epsilon_e, n, hist.data(), &done),
            "fgpe_imag_steps");
        FCE(fgpe_imag_converged(ctx, &conv), "fgpe_imag_converged");
        const double per = done > 0 ? (now_ms() - t0) / done : 0.0;
        invalidate();
        for (int j = 0; j < done; ++j) {
            PARS[EN1_PREV] = PARS[EN1];
            PARS[EN2_PREV] = PARS[EN2];
            call_IM_loop_before_step();
            PARS[NORM1] = hist[6 * j + 0];
            PARS[MU1] = hist[6 * j + 1];
            PARS[EN1] = hist[6 * j + 2];
            PARS[NORM2] = hist[6 * j + 3];
            PARS[MU2] = hist[6 * j + 4];
            PARS[EN2] = hist[6 * j + 5];
            PARS[AVG_TIME] += per;
            PARS[ITER_TIME] = per;
            if (conv && j == done - 1) {
                call_IM_loop_convergence();
                done_all = true;
                break;
            }
            call_IM_loop_after_step();
            PARS[ITER]++;
        }
        it += done;
        if (done < n) done_all = true;
    }
    FCE(fgpe_download_psi(ctx, 0, (double*)psi1), "fgpe_download_psi");
    FCE(fgpe_download_psi(ctx, 1, (double*)psi2), "fgpe_download_psi");
    call_IM_end();
}

// The reference's two-component real-time loop starts at iteration 1 and so
// runs niter - 1 steps (src/twoComponentGPSolver.cu:142).
void TwoComponentGPSolver::runRealTimeEvol()
{
    upload();
    dtCode = p.dt * p.omho1 / 1000;
    call_RE_init();
    for (int iteration = 1; iteration < p.niter; ++iteration) {
        call_RE_loop_before_step();
        if (pending == 0) pendingT0 = (iteration - 1) * dtCode;
        pending++;
        obsTime = iteration * dtCode;
        invalidate();
        call_RE_loop_after_step();
        PARS[ITER]++;
    }
    downloadPsi();
    call_RE_end();
}
