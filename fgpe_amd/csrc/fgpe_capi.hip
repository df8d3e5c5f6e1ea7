// fgpe_capi.hip -- the C ABI of include/fgpe_c.h: context set-up and the
// time-stepping drivers (real-time RK4, imaginary-time RK4 with
// renormalisation, heavy-ball imaginary time, H application, observables).
// The pass kernels are launched through fgpe::Launch<T> (launch_f64.hip,
// launch_f32.hip).  gfx950 only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace fgpe;

namespace fgpe {

__global__ void __launch_bounds__(kFinThreads) k_finalize(FinArgs f0, const FinArgs* __restrict__ tab)
{
    const FinArgs f = args_of(f0, tab, blockIdx.x);
    if (f.stop && *glob(f.stop)) return;
    const gp<const double> partials = glob(f.partials);
    const gp<double> out = glob(f.out), scale = glob(f.scale), mu_out = glob(f.mu), eprev = glob(f.eprev);
    const int t = threadIdx.x;
    const int nv = 3 * f.nc;
    // thread t sums the partials of blocks t, t + n, t + 2n, ... in that order
    // (n = kFinThreads); four of them are loaded per round, so that the loads
    // are in flight together rather than one L2 round trip after another
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    constexpr int n = kFinThreads;
    for (int b = t; b < f.nblocks; b += 4 * n) {
        double x[4][6];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 6; ++q)
                x[u][q] = (q < nv && b + u * n < f.nblocks) ? partials[(long long)(b + u * n) * nv + q] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 6; ++q) s[q] += x[u][q];
    }
    // the block's sum in a fixed order: wave64 shuffles, then the waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < 6; ++q) s[q] += __shfl_down(s[q], off, 64);
    __shared__ double red[6][kFinThreads / 64];
    if ((t & 63) == 0)
#pragma unroll
        for (int q = 0; q < 6; ++q) red[q][t >> 6] = s[q];
    __syncthreads();
    if (t != 0) return;
    double S[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double a = 0.0;
        for (int w = 0; w < kFinThreads / 64; ++w) a += red[q][w];
        S[q] = a;
    }
    bool conv = true;
    // (unrolled over the two possible components, so that every index is a
    // constant: a runtime one would put S and the argument struct in scratch)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (c >= f.nc) break;
        const double norm = S[3 * c] * f.dv;
        if (f.kind == FIN_OBS) {
            out[3 * c] = norm;
            out[3 * c + 1] = S[3 * c + 1] * f.dv / norm;
            out[3 * c + 2] = S[3 * c + 2] * f.dv / norm;
        } else if (f.kind == FIN_RENORM) {
            // the RK4 scheme evaluates no energy operator: only the norm is
            // recorded, the mu and E slots say so
            out[3 * c] = norm;
            out[3 * c + 1] = __builtin_nan("");
            out[3 * c + 2] = __builtin_nan("");
            scale[c] = sqrt(f.npart[c] / norm);
        } else {
            const double sc = sqrt(f.npart[c] / norm);
            const double narg = f.nc == 1 ? f.npart[c] : norm;
            const double mu = sc * S[3 * c + 1] * f.dv / narg;
            const double en = sc * S[3 * c + 2] * f.dv / narg;
            out[3 * c] = norm;
            out[3 * c + 1] = mu;
            out[3 * c + 2] = en;
            scale[c] = sc;
            mu_out[c] = mu;
            const double ep = eprev[c];
            if (f.nc == 1) conv = conv && (fabs(en - ep) / ep < f.eps);
            else conv = conv && (fabs(en - ep) < f.eps);
            eprev[c] = en;
        }
    }
    if (f.kind == FIN_HB) {
        glob(f.done)[0] += 1;
        if (conv) glob(f.stop)[0] = 1;
    }
}


thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

long t_begin(fgpe_ctx* c)
{
    if (!c->timing || c->evused + 2 > c->evpool.size()) return -1;
    const size_t k = c->evused;
    c->evused += 2;
    (void)hipEventRecord(c->evpool[k], c->stream);
    return (long)k;
}

void t_end(fgpe_ctx* c, long k, int cls)
{
    if (k < 0) return;
    (void)hipEventRecord(c->evpool[k + 1], c->stream);
    c->evrec.emplace_back(cls, (size_t)k);
}

int dalloc(fgpe_ctx* c, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(FGPE_EHIP, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    c->bytes += bytes;
    return FGPE_OK;
}

void dfree(fgpe_ctx* c, void** p, size_t bytes)
{
    if (*p) {
        (void)hipFree(*p);
        c->bytes -= bytes;
        *p = nullptr;
    }
}

}  // namespace fgpe

namespace {

// ---- the fixed scratch of fgpe_imag_steps_noalloc --------------------------
// The per-iteration history (norm, mu, E per component) of at most
// FGPE_NOALLOC_MAX_STEPS iterations.  Each device holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host-side
// stand-in, not a device address).
__device__ __align__(256) double scratch[3 * 2 * FGPE_NOALLOC_MAX_STEPS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Looks up this device's scratch address once (with the device current).
int scratch_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(FGPE_EUNSUP, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev]) return FGPE_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(FGPE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = (double*)p;
    return FGPE_OK;
}

double* scratch_of(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (dev >= 0 && dev < kMaxDevices) ? g_scratch_dev[dev] : nullptr;
}

bool supported_n(int n)
{
    switch (n) {
#define CASE(N_) case N_:
        FGPE_FOR_EACH_N(CASE)
#undef CASE
        return true;
    default:
        return false;
    }
}

// W_n^m = exp(-2 pi i m / n), evaluated in long double
template <typename T> std::vector<c2_t<T>> twiddle_table(int n)
{
    std::vector<c2_t<T>> w(n);
    const long double pil = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n; ++m) {
        long double th = 2.0L * pil * (long double)m / (long double)n;
        w[m] = c2_t<T>{(T)(double)cosl(th), (T)(-(double)sinl(th))};
    }
    return w;
}

// the last axis of a 3D (2) or 2D (1) grid; 0 for 1D grids, which have only x
int last_axis(fgpe_ctx* c) { return c->n[2] > 1 ? 2 : (c->n[1] > 1 ? 1 : 0); }
// The axis whose pass carries the k-space multiply in a context's own
// (unbatched) passes: z, or y when the autotuner found that faster on a 3D
// grid; the other of y and z runs the strided forward and inverse passes.
int kmul_axis(const fgpe_ctx* c) { return c->n[2] > 1 ? (c->ylast ? 1 : 2) : (c->n[1] > 1 ? 1 : 0); }
bool has_filter(const fgpe_hparams& h) { return h.a_dd != 0.0 || h.sc_U != 0.0; }

// d_quench_as, hamiltonians/BEC_oneComponent.cu:167-178
double a_at(const fgpe_hparams& h, double t)
{
    if (!h.quench_on) return h.a[0];
    const double a0 = h.quench[0], a1 = h.quench[1], a2 = h.quench[2];
    const double t1 = h.quench[3], t2 = h.quench[4], t3 = h.quench[5];
    if (0 < t && t <= t1) return a0 + (a1 - a0) / (t1) * t;
    if (t1 < t && t <= t2) return a1;
    if (t2 < t && t <= t3) return a1 + (a2 - a1) / (t3 - t2) * (t - t2);
    return a2;
}

// Plain 3D transform on a device array (standalone passes)
template <typename T> int fft3d(fgpe_ctx* c, const void* in, void* out, int dir)
{
    FGPE_RET(Launch<T>::fft_x(c, in, out, dir));
    if (c->n[1] > 1) FGPE_RET(Launch<T>::strided(c, out, 1, dir, 1, 1, nullptr, nullptr, false));
    if (c->n[2] > 1) FGPE_RET(Launch<T>::strided(c, out, 2, dir, 1, 1, nullptr, nullptr, false));
    return FGPE_OK;
}

// y-forward / last axis / y-inverse on the spectral channels.  `scale` is
// applied by the first forward pass after a prologue or a final stage.
// half: the real-field channels are in the half-spectrum layout.
// vpx: the moving-frame term of the Hamiltonian's full form is on.
// 1D grids: the k-space multiply on the x stage's spectra
template <typename T> int kmul_1d(fgpe_ctx* c, void* spec, int nch, int nc, const double* scale, bool filter,
                                  const int* stop, bool half, bool vpx)
{
    LastArgs a = last_args(c, spec, 1, nch, nc, scale, filter, stop, half, vpx);
    a.s.nx = c->n[0];
    // the real-field channels' half spectra: kx = 0..nx/2 (one row each, hchan apart)
    if (!half || nch == nc) a.s.hcols = 0;
    if (nch > nc && !half) return fail(FGPE_EUNSUP, "1D grids keep the density as a half spectrum");
    const int total = nc * c->n[0] + (nch > nc ? (nch - nc) * (c->n[0] / 2 + 1) : 0);
    TScope ts(c, FGPE_KCLASS_LAST);
    hipLaunchKernelGGL(k_kmul_1d<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, a);
    FGPE_LAUNCHCHK("k_kmul_1d");
    return FGPE_OK;
}

template <typename T> int spectral_passes_on(fgpe_ctx* c, void* spec, int nch, int nc, const double* scale,
                                             const int* stop, bool half, bool vpx = false)
{
    const bool filter = has_filter(c->hp);
    const int la = kmul_axis(c);
    if (la == 0) return kmul_1d<T>(c, spec, nch, nc, scale, filter, stop, half, vpx);
    if (c->n[2] > 1) {
        // slab by slab of columns: a slab's three passes run while it is still in the memory-side cache
        const int ns = half ? c->nslab : 1, sa = 3 - la;
        for (int sl = 0; sl < ns; ++sl) {
            FGPE_RET(Launch<T>::strided(c, spec, sa, -1, nch, nc, scale, stop, half, nullptr, 1, sl, ns));
            FGPE_RET(Launch<T>::last(c, spec, la, nch, nc, nullptr, filter, stop, half, nullptr, 1, sl, ns, vpx));
            FGPE_RET(Launch<T>::strided(c, spec, sa, +1, nch, nc, nullptr, stop, half, nullptr, 1, sl, ns));
        }
    } else {
        FGPE_RET(Launch<T>::last(c, spec, 1, nch, nc, scale, filter, stop, half, nullptr, 1, 0, 1, vpx));
    }
    return FGPE_OK;
}

// The solver's channels: the components, then the total density as a half
// spectrum when a k-space kernel is on.
template <typename T> int spectral_passes(fgpe_ctx* c, const double* scale, const int* stop, bool vpx = false)
{
    return spectral_passes_on<T>(c, c->spec, c->nch, c->hp.ncomp, scale, stop, true, vpx);
}

template <typename T> XArgs<T> base_xargs(fgpe_ctx* c, double t)
{
    XArgs<T> a{};
    a.spec = (c2_t<T>*)c->spec;
    a.ntot = c->ntot;
    a.specd = (c2_t<T>*)c->spec + (long long)c->hp.ncomp * c->ntot;
    a.dpitch = (int)dens_pitch(c);
    for (int i = 0; i < 2; ++i) {
        a.P0[i] = c->P0[i];
        a.PS[i] = (c2_t<T>*)c->PS[i];
        a.K[i] = (c2_t<T>*)c->K[i];
        a.Vr[i] = (const T*)c->Vr[i];
        a.Vi[i] = c->vcomplex[i] ? (const T*)c->Vi[i] : nullptr;
    }
    const fgpe_hparams& h = c->hp;
    if (h.ncomp == 1) {
        const double av = a_at(h, t);
        a.g4[0][0] = 4 * kPi * av;
        a.g2[0][0] = 2 * kPi * av;
    } else {
        const double aa[2][2] = {{h.a[0], h.a[2]}, {h.a[2], h.a[1]}};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                a.g4[i][j] = 4 * kPi * aa[i][j];
                a.g2[i][j] = 2 * kPi * aa[i][j];
            }
    }
    a.partials = c->partials;
    a.ny = c->n[1];
    a.t = t;
    for (int i = 0; i < 8; ++i) a.user[i] = h.user[i];
    return a;
}

int ensure_ready(fgpe_ctx* c)
{
    if (!c->kset) return fail(FGPE_ESTATE, "fgpe_set_kspace has not been called");
    if (!c->hset) return fail(FGPE_ESTATE, "fgpe_set_hamiltonian has not been called");
    for (int i = 0; i < c->hp.ncomp; ++i) {
        if (!c->Vr[i]) return fail(FGPE_ESTATE, "no external potential uploaded for a component");
        if (!c->P0[i]) return fail(FGPE_ESTATE, "no state uploaded for a component");
    }
    return FGPE_OK;
}

template <typename T> int prologue(fgpe_ctx* c)
{
    XArgs<T> a = base_xargs<T>(c, 0.0);
    FGPE_RET(Launch<T>::xstage(c, a, c->hp.ncomp, c->nch > c->hp.ncomp, MODE_PROLOGUE));
    c->spec_valid = true;
    return FGPE_OK;
}

FinArgs fin_args(fgpe_ctx* c, long long nblocks, int kind, const double* npart, double* out, double eps);

int finalize(fgpe_ctx* c, long long nblocks, int kind, const double* npart, double* out, double eps)
{
    const FinArgs f = fin_args(c, nblocks, kind, npart, out, eps);
    TScope ts(c, FGPE_KCLASS_OTHER);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinThreads), 0, c->stream, f, (const FinArgs*)nullptr);
    FGPE_LAUNCHCHK("k_finalize");
    return FGPE_OK;
}

int apply_pending_scale(fgpe_ctx* c);

// The imaginary-time renormalisation is normally left pending and applied
// by the next forward pass: a component by its own scale, the density by
// scale[0]^2.  With two components and a k-space kernel the density is
// n1 + n2 under two different scales, which no single factor rescales, so
// there the states are rescaled at once and the spectra rebuilt.
bool eager_renorm(const fgpe_ctx* c) { return c->hp.ncomp == 2 && c->nch > c->hp.ncomp; }

template <typename T> int renorm_now(fgpe_ctx* c)
{
    FGPE_RET(apply_pending_scale(c));
    return prologue<T>(c);
}

// RK4 steps; imaginary: c = -dt and renormalisation to npart after each step
// keep (one step only): a buffer of 3 x ncomp bulk states that receives the
// step's stage states y1, y2, y3 (stage s, component i at ((s - 1) ncomp + i)
// states in), which the stages then read from there instead of PS / PSb.
template <typename T> int rk4_run(fgpe_ctx* c, double dt, double t0, int nsteps, bool imaginary,
                                  const double* npart, double* hist_dev, void* keep = nullptr)
{
    FGPE_RET(ensure_ready(c));
    if (keep && nsteps != 1) return fail(FGPE_EINVAL, "stage states are kept for one step at a time");
    const int nc = c->hp.ncomp;
    const bool dens = c->nch > nc;
    if (!c->spec_valid) FGPE_RET(prologue<T>(c));
    const double* scale = imaginary ? c->scal + SC_SCALE : nullptr;
    const long long nb = Launch<T>::xblocks(c);
    // The mixed form keeps the stage-1 state until stage 3, which rebuilds
    // psi_n from it instead of reading the FP64 state: stage 2 writes its
    // state to a second buffer (allocated with the other per-component ones).
    constexpr bool keep1 = sizeof(T) == 4;
    if (keep1)
        for (int i = 0; i < nc; ++i)
            if (!c->PSb[i]) return fail(FGPE_ESTATE, "mixed-form stage buffer missing");
    for (int it = 0; it < nsteps; ++it) {
        for (int s = 1; s <= 4; ++s) {
            // PARS[T] at each stage (src/oneComponentGPSolver.cu:175,184)
            const double frac = s == 1 ? 0.0 : (s == 4 ? 1.0 : 0.5);
            const double t = s == 1 ? t0 + it * dt : t0 + (it + frac) * dt;
            // the moving-frame term rides with alg_calcHpsi, i.e. stage 4 of the
            // one-component step (src/oneComponentGPSolver.cu:170-187)
            FGPE_RET(spectral_passes<T>(c, s == 1 ? scale : nullptr, nullptr, s == 4 && nc == 1));
            XArgs<T> a = base_xargs<T>(c, t);
            const double cf = (s <= 2) ? -0.5 * dt : -1.0 * dt;
            if (imaginary) { a.cr = cf; a.ci = 0.0; }
            else { a.cr = 0.0; a.ci = cf; }
            a.stage = s;
            a.scale = scale;
            a.want_norm = (imaginary && s == 4) ? 1 : 0;
            if (keep) {
                auto y = [&](int st, int i) {
                    return (c2_t<T>*)keep + ((size_t)(st - 1) * nc + i) * (size_t)c->ntot;
                };
                for (int i = 0; i < nc; ++i) {
                    if (s > 1) a.PS[i] = y(s - 1, i);   // stage 1 reads the state
                    if (s < 4) a.PSo[i] = y(s, i);
                    if (keep1 && s == 3) a.PS1[i] = y(1, i);
                }
            } else if (keep1) {
                for (int i = 0; i < nc; ++i) {
                    c2_t<T>* const ps1 = (c2_t<T>*)c->PS[i];
                    c2_t<T>* const ps2 = (c2_t<T>*)c->PSb[i];
                    if (s == 2) a.PSo[i] = ps2;
                    if (s == 3) {
                        a.PS[i] = ps2;
                        a.PS1[i] = ps1;
                        a.PSo[i] = ps1;
                    }
                }
            }
            FGPE_RET(Launch<T>::xstage(c, a, nc, dens, MODE_RK));
            if (imaginary && s == 4) {
                FGPE_RET(finalize(c, nb, FIN_RENORM, npart, hist_dev ? hist_dev + 3 * nc * it : c->scal + SC_TMP,
                                  0.0));
                if (eager_renorm(c)) FGPE_RET(renorm_now<T>(c));
            }
        }
    }
    return FGPE_OK;
}

__global__ void k_hb_continue(int* flags) { flags[FL_DONE] = 0; }

__global__ void k_hb_init(double* scal, int* flags)
{
    for (int i = 0; i < 2; ++i) {
        scal[SC_SCALE + i] = 1.0;
        scal[SC_MU + i] = 0.0;
        scal[SC_EPREV + i] = 0.0;
    }
    flags[FL_STOP] = 0;
    flags[FL_DONE] = 0;
}

// Heavy-ball set-up: on a fresh run resets the iteration scalars and sets
// old = psi (src/oneComponentGPSolver.cu:33), on a continuation only clears
// the iteration counter; then the x prologue of the current state.
template <typename T> int hb_prepare(fgpe_ctx* c)
{
    const int nc = c->hp.ncomp;
    if (c->hb_active) {
        hipLaunchKernelGGL(k_hb_continue, dim3(1), dim3(1), 0, c->stream, c->flags);
        FGPE_LAUNCHCHK("k_hb_continue");
    } else {
        hipLaunchKernelGGL(k_hb_init, dim3(1), dim3(1), 0, c->stream, c->scal, c->flags);
        FGPE_LAUNCHCHK("k_hb_init");
        c->imag_iter = 0;
        for (int i = 0; i < nc; ++i) {
            if (sizeof(T) == 8) {
                FGPE_HIPCHK(hipMemcpyAsync(c->PS[i], c->P0[i], sizeof(double2) * c->ntot, hipMemcpyDeviceToDevice,
                                           c->stream));
            } else {
                const long long n = 2 * c->ntot;
                hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                                   c->stream, (float*)c->PS[i], (const double*)c->P0[i], n);
                FGPE_LAUNCHCHK("k_convert");
            }
        }
        c->hb_active = true;
    }
    return prologue<T>(c);
}

template <typename T> XArgs<T> hb_xargs(fgpe_ctx* c, double t, double dt, double beta)
{
    XArgs<T> a = base_xargs<T>(c, t);
    a.scale = c->scal + SC_SCALE;
    a.dt = dt;
    a.beta = beta;
    a.mu = c->hp.ncomp == 2 ? c->scal + SC_MU : nullptr;
    a.stop = c->flags + FL_STOP;
    return a;
}

FinArgs fin_args(fgpe_ctx* c, long long nblocks, int kind, const double* npart, double* out, double eps)
{
    FinArgs f{};
    f.partials = c->partials;
    f.nblocks = (int)nblocks;
    f.nc = c->hp.ncomp;
    f.kind = kind;
    f.dv = c->dv;
    f.npart[0] = npart[0];
    f.npart[1] = c->hp.ncomp == 2 ? npart[1] : npart[0];
    f.out = out;
    f.scale = c->scal + SC_SCALE;
    f.mu = c->scal + SC_MU;
    f.eprev = c->scal + SC_EPREV;
    f.stop = kind == FIN_HB ? c->flags + FL_STOP : nullptr;
    f.done = c->flags + FL_DONE;
    f.eps = eps;
    return f;
}

// Heavy-ball imaginary time (the reference's scheme), fully on the device: the
// convergence test runs in k_finalize and stops every later launch.
template <typename T> int hb_run(fgpe_ctx* c, double dt, double beta, const double* npart, double eps,
                                 int nsteps, double* hist_dev)
{
    FGPE_RET(ensure_ready(c));
    const int nc = c->hp.ncomp;
    const bool dens = c->nch > nc;
    FGPE_RET(hb_prepare<T>(c));
    const long long nb = Launch<T>::xblocks(c);
    const int* stop = c->flags + FL_STOP;
    for (int it = 0; it < nsteps; ++it) {
        FGPE_RET(spectral_passes<T>(c, c->scal + SC_SCALE, stop, true));
        // PARS[T] = iteration * dt over the whole run (src/oneComponentGPSolver.cu:50)
        FGPE_RET(Launch<T>::xstage(c, hb_xargs<T>(c, (double)(c->imag_iter + it) * dt, dt, beta), nc, dens, MODE_HB));
        FGPE_RET(finalize(c, nb, FIN_HB, npart, hist_dev ? hist_dev + 3 * nc * it : c->scal + SC_TMP, eps));
        if (eager_renorm(c)) FGPE_RET(renorm_now<T>(c));
    }
    return FGPE_OK;
}

template <typename T> int apply_h(fgpe_ctx* c, double t, void* hmu, void* hen, double* obs_dev)
{
    FGPE_RET(ensure_ready(c));
    FGPE_RET(prologue<T>(c));
    FGPE_RET(spectral_passes<T>(c, c->scal + SC_SCALE, nullptr, true));
    XArgs<T> a = base_xargs<T>(c, t);
    a.scale = c->scal + SC_SCALE;
    a.hmu = (double2*)hmu;
    a.hen = (double2*)hen;
    FGPE_RET(Launch<T>::xstage(c, a, c->hp.ncomp, c->nch > c->hp.ncomp, MODE_APPLY));
    c->spec_valid = false;
    const double one[2] = {1.0, 1.0};
    if (obs_dev) FGPE_RET(finalize(c, Launch<T>::xblocks(c), FIN_OBS, one, obs_dev, 0.0));
    return FGPE_OK;
}

__global__ void k_apply_scales(double2* __restrict__ p0, double2* __restrict__ p1, long long n, double* scale)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double s0 = scale[0];
        p0[i] = double2{s0 * p0[i].x, s0 * p0[i].y};
        if (p1) {
            const double s1 = scale[1];
            p1[i] = double2{s1 * p1[i].x, s1 * p1[i].y};
        }
    }
}
__global__ void k_reset_scales(double* scale) { scale[0] = scale[1] = 1.0; }

int apply_pending_scale(fgpe_ctx* c)
{
    const int th = 256;
    hipLaunchKernelGGL(k_apply_scales, dim3((unsigned)((c->ntot + th - 1) / th)), dim3(th), 0, c->stream,
                       c->P0[0], c->hp.ncomp == 2 ? c->P0[1] : nullptr, c->ntot, c->scal + SC_SCALE);
    FGPE_LAUNCHCHK("k_apply_scales");
    hipLaunchKernelGGL(k_reset_scales, dim3(1), dim3(1), 0, c->stream, c->scal + SC_SCALE);
    FGPE_LAUNCHCHK("k_reset_scales");
    c->spec_valid = false;
    return FGPE_OK;
}

// per-block sums of w(i) |a_i|^2 with w = kx[ix] (kind 0) or
// cos(2 pi x[ix] / L) (kind 1); deterministic block order
__global__ void __launch_bounds__(256) k_weighted_sum(const double2* __restrict__ a, const double* __restrict__ w,
                                                      int nx, long long n, int kind, double L, double* partials)
{
    double s[1] = {0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ix = (int)(i % nx);
        const double wt = kind == 0 ? w[ix] : cos(2 * kPi * w[ix] / L);
        s[0] += wt * (a[i].x * a[i].x + a[i].y * a[i].y);
    }
    block_sum_store<1>(s, glob(partials) + blockIdx.x);
}

__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ partials, int nb, double scale,
                                                      double* out)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += partials[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] * scale;
}

int weighted_sum(fgpe_ctx* c, const double2* a, const double* w, int kind, double L, double* host_out)
{
    const int nb = 1024;
    hipLaunchKernelGGL(k_weighted_sum, dim3(nb), dim3(256), 0, c->stream, a, w, c->n[0], c->ntot, kind, L,
                       c->partials);
    FGPE_LAUNCHCHK("k_weighted_sum");
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c->stream, c->partials, nb, c->dv, c->scal + SC_TMP);
    FGPE_LAUNCHCHK("k_sum_partials");
    FGPE_HIPCHK(hipMemcpyAsync(host_out, c->scal + SC_TMP, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    return FGPE_OK;
}

__global__ void k_accumulate_impl(double* __restrict__ acc, const double* __restrict__ add, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] += add[i];
}

int adj_check(fgpe_ctx* c)
{
    FGPE_RET(ensure_ready(c));
    for (int i = 0; i < c->hp.ncomp; ++i)
        if (c->vcomplex[i]) return fail(FGPE_EUNSUP, "the adjoint needs a real trap");
    if (c->hp.vpx[0] != 0.0 || c->hp.vpx[1] != 0.0) return fail(FGPE_EUNSUP, "the adjoint has no moving-frame term");
    if (c->hp.user_id != 0) return fail(FGPE_EUNSUP, "the adjoint does not know the user term's derivative");
    return FGPE_OK;
}

// The adjoint's buffers, sized for nc components: checkpoints and segment
// states (FP64, nc states each), per-step stage states (allocated by the
// backward sweep), the VJP's spectral channels (nc complex ones and up to
// three half spectra), the adjoint vectors and the parameter accumulators.
// A context that switches to more components drops the smaller ones.
int adj_reserve(fgpe_ctx* c, int nc)
{
    if (c->adj_nc >= nc) return FGPE_OK;
    const size_t N = (size_t)c->ntot;
    for (double2*& p : c->ckpt) dfree(c, (void**)&p, sizeof(double2) * N * c->adj_nc);
    for (double2*& p : c->segbuf) dfree(c, (void**)&p, sizeof(double2) * N * c->adj_nc);
    for (void*& p : c->ystep) dfree(c, &p, 3 * 2 * c->tsz() * N * c->adj_nc);
    c->ckpt.clear();
    c->segbuf.clear();
    c->ystep.clear();
    for (int i = 0; i < 3; ++i) dfree(c, (void**)&c->lam[i], sizeof(double2) * N * c->adj_nc);
    dfree(c, &c->spec3, (size_t)(c->adj_nc + 3) * 2 * c->tsz() * N);
    c->adj_nc = nc;
    for (int i = 0; i < 3; ++i) FGPE_RET(dalloc(c, (void**)&c->lam[i], sizeof(double2) * N * nc));
    FGPE_RET(dalloc(c, &c->spec3, (size_t)(nc + 3) * 2 * c->tsz() * N));
    if (!c->pacc) FGPE_RET(dalloc(c, (void**)&c->pacc, sizeof(double) * 4 * (size_t)(c->ntot / 32 + 1024)));
    c->adj_ready = false;
    return FGPE_OK;
}

// nc FP64 states, component-major: copy between the solver's per-component
// state buffers and one of the adjoint's
int copy_states(fgpe_ctx* c, double2* buf, bool to_buf)
{
    const size_t b64 = sizeof(double2) * (size_t)c->ntot;
    for (int i = 0; i < c->hp.ncomp; ++i) {
        double2* b = buf + (size_t)i * c->ntot;
        FGPE_HIPCHK(hipMemcpyAsync(to_buf ? b : c->P0[i], to_buf ? c->P0[i] : b, b64, hipMemcpyDeviceToDevice,
                                   c->stream));
    }
    return FGPE_OK;
}

template <typename T> int adj_forward(fgpe_ctx* c, double dt, double t0, int nsteps, int every)
{
    FGPE_RET(adj_check(c));
    const int nc = c->hp.ncomp;
    FGPE_RET(adj_reserve(c, nc));
    const int nseg = (nsteps + every - 1) / every;
    const size_t bytes = sizeof(double2) * c->ntot * c->adj_nc;
    while ((int)c->ckpt.size() < nseg) {
        double2* p = nullptr;
        FGPE_RET(dalloc(c, (void**)&p, bytes));
        c->ckpt.push_back(p);
    }
    const double one[2] = {1.0, 1.0};
    for (int seg = 0; seg < nseg; ++seg) {
        const int len = std::min(every, nsteps - seg * every);
        FGPE_RET(copy_states(c, c->ckpt[seg], true));
        FGPE_RET(rk4_run<T>(c, dt, t0 + (double)seg * every * dt, len, false, one, nullptr));
    }
    c->adj_every = every;
    c->adj_nsteps = nsteps;
    c->adj_dt = dt;
    c->adj_t0 = t0;
    c->adj_ready = true;
    return FGPE_OK;
}

// One VJP: prologue (abar_c, n, r channels), spectral passes, epilogue; the
// epilogue's per-block parameter sums are added to the accumulators.  With
// both kernels on, the passes run once more over the third half channel (n
// again) with the dipolar kernel alone.
template <typename T> int vjp(fgpe_ctx* c, AdjArgs a, bool y64, bool dens, double* ptmp, int ns)
{
    const int nc = c->hp.ncomp;
    FGPE_RET(Launch<T>::adj(c, a, y64, dens, false));
    FGPE_RET(spectral_passes_on<T>(c, c->spec3, dens ? nc + 2 : nc, nc, nullptr, nullptr, true));
    if (a.both) {
        c2_t<T>* const third = (c2_t<T>*)c->spec3 + (long long)nc * c->ntot + 2 * a.hchan;
        c->kmask = 1;
        const int rc = spectral_passes_on<T>(c, third, 1, 0, nullptr, nullptr, true);
        c->kmask = 3;
        FGPE_RET(rc);
    }
    a.partials = ptmp;
    FGPE_RET(Launch<T>::adj(c, a, y64, dens, true));
    const long long n = (long long)ns * Launch<T>::xblocks(c);
    hipLaunchKernelGGL(k_accumulate_impl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->pacc,
                       ptmp, n);
    FGPE_LAUNCHCHK("k_accumulate");
    return FGPE_OK;
}

template <typename T> int adj_backward(fgpe_ctx* c, const void* lamT, void* lam0, double* vbar, double* grads)
{
    FGPE_RET(adj_check(c));
    if (!c->adj_ready || c->adj_nc < c->hp.ncomp) return fail(FGPE_ESTATE, "fgpe_adjoint_forward has not been run");
    const int nc = c->hp.ncomp;
    const int ns = nc == 1 ? 3 : 5;   // parameter sums per block
    const size_t N = (size_t)c->ntot;
    const size_t b64 = sizeof(double2) * N * nc;   // nc FP64 states
    const size_t bT = 2 * c->tsz() * N;             // one bulk state
    const bool dens = c->nch > nc;
    const int every = c->adj_every, nsteps = c->adj_nsteps;
    const int nseg = (nsteps + every - 1) / every;
    const double dt = c->adj_dt;
    const long long nb = Launch<T>::xblocks(c);
    while ((int)c->segbuf.size() < every) {
        double2* p = nullptr;
        FGPE_RET(dalloc(c, (void**)&p, sizeof(double2) * N * c->adj_nc));
        c->segbuf.push_back(p);
    }
    while ((int)c->ystep.size() < every) {
        void* p = nullptr;
        FGPE_RET(dalloc(c, &p, 3 * bT * c->adj_nc));
        c->ystep.push_back(p);
    }
    if ((size_t)2 * ns * nb > (size_t)4 * (size_t)(c->ntot / 32 + 1024))
        return fail(FGPE_EUNSUP, "grid too small for the adjoint's partial-sum buffer");
    double* ptmp = c->pacc + ns * nb;
    FGPE_HIPCHK(hipMemsetAsync(c->pacc, 0, sizeof(double) * ns * nb, c->stream));
    FGPE_HIPCHK(hipMemsetAsync(vbar, 0, sizeof(double) * N * nc, c->stream));
    double2 *LIN = c->lam[0], *LAM = c->lam[1], *GB = c->lam[2];
    FGPE_HIPCHK(hipMemcpyAsync(LIN, lamT, b64, hipMemcpyDefault, c->stream));
    const double one[2] = {1.0, 1.0};
    const double2 cc = double2{0.0, dt};   // conj(c), c = -i dt
    auto scl = [](double s, double2 z) { return double2{s * z.x, s * z.y}; };
    auto comp = [&](void* base, int i, size_t stride) { return (void*)((char*)base + (size_t)i * stride); };
    // 4 pi a_cc' at the time each stage evaluated H (the one-component a(t)
    // ramp, when on, is part of the dynamics the adjoint differentiates)
    auto set_g4 = [&](AdjArgs& a, double t) {
        if (nc == 1) {
            a.g4[0][0] = 4 * kPi * a_at(c->hp, t);
            a.g4[0][1] = a.g4[1][0] = a.g4[1][1] = 0.0;
        } else {
            const double* aa = c->hp.a;   // a_11, a_22, a_12
            a.g4[0][0] = 4 * kPi * aa[0];
            a.g4[1][1] = 4 * kPi * aa[1];
            a.g4[0][1] = a.g4[1][0] = 4 * kPi * aa[2];
        }
    };

    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int len = std::min(every, nsteps - seg * every);
        const int step0 = seg * every;
        // recompute the states of this segment, keeping every step's stage
        // states y1..y3 (src/oneComponentGPSolver.cu:165-182,
        // src/twoComponentGPSolver.cu:150-182), stage s evaluating H at PARS[T]
        // as the forward step did
        FGPE_RET(copy_states(c, c->ckpt[seg], false));
        c->spec_valid = false;
        for (int j = 0; j < len; ++j) {
            FGPE_RET(copy_states(c, c->segbuf[j], true));
            FGPE_RET(rk4_run<T>(c, dt, c->adj_t0 + (step0 + j) * dt, 1, false, one, nullptr, c->ystep[j]));
        }
        for (int j = len - 1; j >= 0; --j) {
            double2* Y0 = c->segbuf[j];
            const double tj = c->adj_t0 + (double)(step0 + j) * dt;   // this step's start
            void* const ys = c->ystep[j];
            auto ystage = [&](int s) { return (void*)((char*)ys + (size_t)(s - 1) * nc * bT); };
            AdjArgs a{};
            a.spec = c->spec3;
            a.ntot = c->ntot;
            a.hpitch = (int)dens_pitch(c);
            a.hchan = dens_pitch(c) * c->n[1] * c->n[2];
            a.both = (c->hp.a_dd != 0.0 && c->hp.sc_U != 0.0) ? 1 : 0;
            for (int i = 0; i < nc; ++i) {
                a.Vr[i] = c->Vr[i];
                a.lam_in[i] = LIN + (size_t)i * N;
                a.lam[i] = LAM + (size_t)i * N;
                a.abar[i] = GB + (size_t)i * N;
                a.vbar[i] = vbar + (size_t)i * N;
            }
            auto set_y = [&](void* base, size_t stride) {
                for (int i = 0; i < nc; ++i) a.y[i] = comp(base, i, stride);
            };
            // s = 3: abar = conj(c)/6 lam_in at y3; lam = lam_in + G; next abar = conj(c)/3 lam_in + conj(c) G
            set_y(ystage(3), bT);
            set_g4(a, tj + dt);
            for (int i = 0; i < nc; ++i) {
                a.src[i] = LIN + (size_t)i * N;
                a.gb_out[i] = GB + (size_t)i * N;
            }
            a.alpha = scl(1.0 / 6.0, cc);
            a.init = 1;
            a.ga = scl(1.0 / 3.0, cc);
            a.gg = cc;
            FGPE_RET(vjp<T>(c, a, sizeof(T) == 8, dens, ptmp, ns));
            // s = 2 at y2: next abar = conj(c)/3 lam_in + conj(c)/2 G
            set_y(ystage(2), bT);
            set_g4(a, tj + 0.5 * dt);
            for (int i = 0; i < nc; ++i) a.src[i] = nullptr;
            a.init = 0;
            a.ga = scl(1.0 / 3.0, cc);
            a.gg = scl(0.5, cc);
            FGPE_RET(vjp<T>(c, a, sizeof(T) == 8, dens, ptmp, ns));
            // s = 1 at y1: next abar = conj(c)/6 lam_in + conj(c)/2 G
            set_y(ystage(1), bT);
            set_g4(a, tj + 0.5 * dt);
            a.ga = scl(1.0 / 6.0, cc);
            a.gg = scl(0.5, cc);
            FGPE_RET(vjp<T>(c, a, sizeof(T) == 8, dens, ptmp, ns));
            // s = 0 at y
            set_y(Y0, sizeof(double2) * N);
            set_g4(a, tj);
            for (int i = 0; i < nc; ++i) a.gb_out[i] = nullptr;
            FGPE_RET(vjp<T>(c, a, true, dens, ptmp, ns));
            std::swap(LIN, LAM);
        }
    }
    c->lam[0] = LIN;
    c->lam[1] = LAM;
    if (lam0) FGPE_HIPCHK(hipMemcpyAsync(lam0, LIN, b64, hipMemcpyDefault, c->stream));
    // fixed-order sum of the per-block accumulators
    std::vector<double> h((size_t)ns * nb);
    FGPE_HIPCHK(hipMemcpyAsync(h.data(), c->pacc, sizeof(double) * ns * nb, hipMemcpyDeviceToHost, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long b = 0; b < nb; ++b)
        for (int k = 0; k < ns; ++k) sum[k] += h[(size_t)ns * b + k];
    // the contact sums: dJ/da (or, with the a(t) ramp on, the derivative along a
    // uniform shift of a(t)); the kernel sum <r, Phi[n]>: each kernel's strength
    // (a_dd or U) scales its part of Phi, the dipolar part <r, Phi_dd[n]> summed
    // on its own when both are on
    const double sk = nc == 1 ? sum[1] : sum[3], sdd = nc == 1 ? sum[2] : sum[4];
    const bool both = c->hp.a_dd != 0.0 && c->hp.sc_U != 0.0;
    grads[0] = 4 * kPi * sum[0];
    grads[1] = c->hp.a_dd != 0.0 ? (both ? sdd : sk) / c->hp.a_dd : 0.0;
    grads[2] = c->hp.sc_U != 0.0 ? (both ? sk - sdd : sk) / c->hp.sc_U : 0.0;
    grads[3] = nc == 2 ? 4 * kPi * sum[1] : 0.0;
    grads[4] = nc == 2 ? 4 * kPi * sum[2] : 0.0;
    c->spec_valid = false;
    return FGPE_OK;
}

// Times both plan variants of each kernel class on the context's own buffers
// (stage-2 x passes write only PS, K and the spectral channels, never the
// state, so it survives; the spectrum is recomputed before the next step)
// and keeps the faster one.
template <typename T> int autotune(fgpe_ctx* c, int reps, double* out)
{
    FGPE_RET(ensure_ready(c));
    const int nc = c->hp.ncomp;
    const bool dens = c->nch > nc;
    const int la0 = last_axis(c);
    hipEvent_t e0, e1;
    FGPE_HIPCHK(hipEventCreate(&e0));
    FGPE_HIPCHK(hipEventCreate(&e1));
    const bool timing = c->timing;
    c->timing = false;
    auto run = [&](int cls) -> int {
        const int la = kmul_axis(c);
        if (cls == FGPE_KCLASS_STRIDED) {
            FGPE_RET(Launch<T>::strided(c, c->spec, 3 - la, -1, c->nch, nc, nullptr, nullptr, true));
            return Launch<T>::strided(c, c->spec, 3 - la, +1, c->nch, nc, nullptr, nullptr, true);
        }
        if (cls == FGPE_KCLASS_LAST)
            return Launch<T>::last(c, c->spec, la, c->nch, nc, nullptr, has_filter(c->hp), nullptr, true);
        XArgs<T> a = base_xargs<T>(c, 0.0);
        a.cr = 0.0;
        a.ci = -1e-30;
        a.stage = 2;
        return Launch<T>::xstage(c, a, nc, dens, MODE_RK);
    };
    // average time of n calls of f after one untimed call
    auto timed_n = [&](auto&& f, int n, double* per) -> int {
        FGPE_RET(f());
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < n; ++r) FGPE_RET(f());
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) return fail(FGPE_EHIP, "autotune timing failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *per = ms / n;
        return FGPE_OK;
    };
    auto timed = [&](auto&& f, double* per) -> int { return timed_n(f, reps, per); };
    // one RK stage: the spectral passes and an x stage
    auto stage = [&]() -> int {
        FGPE_RET(spectral_passes_on<T>(c, c->spec, c->nch, nc, nullptr, nullptr, true));
        return run(FGPE_KCLASS_XSTAGE);
    };
    // the plan variants of one kernel class on one axis (the x class has a
    // third where its length has one); keeps the fastest.  The x stage is
    // timed alone; the passes in whole RK stages (twice the repetitions),
    // since a pass runs on what the pass before it left in the memory-side
    // cache: at 256^3 FP64 the two last-axis plans tie alone and differ by 7%
    // in a stage, and the y passes' split plan loses alone and wins by 3% in
    // a stage.
    double tt[7] = {};
    auto tune = [&](int cls, int axis, int ci) -> int {
        const int nv = cls == FGPE_KCLASS_XSTAGE && x_variant2(c->n[0]) ? 3 : 2;
        double best = 1e300;
        int bestv = 0;
        for (int v = 0; v < nv; ++v) {
            c->var[cls][axis] = v;
            double per = 0.0;
            if (cls == FGPE_KCLASS_XSTAGE) FGPE_RET(timed([&] { return run(cls); }, &per));
            else FGPE_RET(timed_n(stage, 2 * reps, &per));
            tt[v < 2 ? 2 * ci + v : 6] = per;
            if (per < best) {
                best = per;
                bestv = v;
            }
        }
        c->var[cls][axis] = bestv;
        return FGPE_OK;
    };
    // block order of the four passes of a stage (reversed or not per pass),
    // timed as whole stages; keeps the fastest of the 16 combinations
    auto tune_order = [&](double* best_ms) -> int {
        double best = 1e300;
        int besto = 0;
        for (int o = 0; o < 16; ++o) {
            if (la0 != 2 && (o & 5)) continue;   // 2D grids have no strided passes
            if (la0 == 0 && (o & 2)) continue;   // 1D grids no last-axis pass
            c->order = o;
            double per = 0.0;
            FGPE_RET(timed(stage, &per));
            if (per < best) {
                best = per;
                besto = o;
            }
        }
        c->order = besto;
        *best_ms = best;
        return FGPE_OK;
    };
    auto body = [&]() -> int {
        c->nslab = 1;
        FGPE_RET(tune(FGPE_KCLASS_XSTAGE, 0, 2));
        if (la0 == 1) FGPE_RET(tune(FGPE_KCLASS_LAST, 1, 1));
        double best = 0.0;
        if (la0 == 2) {
            // 3D: the k-space multiply on the z pass or on the y pass.  Each
            // choice gets its pass variants and block order tuned, then the
            // faster whole stage wins (y by at least 1%).
            double keep[7] = {};
            int keep_order = 0;
            bool besty = false;
            best = 1e300;
            for (int yl = 0; yl < 2; ++yl) {
                c->ylast = yl != 0;
                const int la = kmul_axis(c);
                FGPE_RET(tune(FGPE_KCLASS_STRIDED, 3 - la, 0));
                FGPE_RET(tune(FGPE_KCLASS_LAST, la, 1));
                double per = 0.0;
                FGPE_RET(tune_order(&per));
                // the pass plans again in the chosen block order (what a pass
                // finds in the memory-side cache depends on it: at 256^3 the
                // FP64 y passes' split plan loses in the forward order and wins
                // by 3% in the chosen one), and the order again if one changed
                const int vs = c->var[FGPE_KCLASS_STRIDED][3 - la], vl = c->var[FGPE_KCLASS_LAST][la];
                FGPE_RET(tune(FGPE_KCLASS_STRIDED, 3 - la, 0));
                FGPE_RET(tune(FGPE_KCLASS_LAST, la, 1));
                if (c->var[FGPE_KCLASS_STRIDED][3 - la] != vs || c->var[FGPE_KCLASS_LAST][la] != vl)
                    FGPE_RET(tune_order(&per));
                if (per < best * (yl ? 0.99 : 1.0)) {
                    best = per;
                    besty = yl != 0;
                    keep_order = c->order;
                    std::memcpy(keep, tt, sizeof(tt));
                }
            }
            c->ylast = besty;
            c->order = keep_order;
            std::memcpy(tt, keep, sizeof(tt));
        } else {
            FGPE_RET(tune_order(&best));
            if (la0 == 1) {   // 2D: the last-axis plan again in the chosen order
                const int vl = c->var[FGPE_KCLASS_LAST][1];
                FGPE_RET(tune(FGPE_KCLASS_LAST, 1, 1));
                if (c->var[FGPE_KCLASS_LAST][1] != vl) FGPE_RET(tune_order(&best));
            }
        }
        if (out) {
            for (int k = 0; k < 6; ++k) out[k] = tt[k];
            out[9] = tt[6];
        }
        // column slabs of the strided / last / strided sequence (3D only): 1, 2
        // or 4, timed as whole RK stages (passes + x stage), since the x stage
        // between two pass sequences decides what the memory-side cache holds
        if (la0 == 2) {
            const int cand[3] = {1, 2, 4};
            int bests = 1;
            if (out) out[6] = best;
            for (int k = 1; k < 3; ++k) {
                c->nslab = cand[k];
                double per = 0.0;
                const int r = timed(stage, &per);
                if (r == FGPE_EUNSUP) continue;   // slab width not a multiple of a tile width
                FGPE_RET(r);
                if (out) out[6 + k] = per;
                // More slabs have to win clearly: at 256^3 the timing of two
                // slabs has tied with one slab here (0.05% apart) and then run
                // 10% slower in a step, whose x stages differ from the timing
                // mode's.
                if (per < best * 0.95) {
                    best = per;
                    bests = cand[k];
                }
            }
            c->nslab = bests;
        }
        return FGPE_OK;
    };
    const int rc = body();
    c->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->spec_valid = false;
    return rc;
}

// The soft-core Vt(k) table from the context's k axes and Hamiltonian (built
// by fgpe_set_hamiltonian, rebuilt by fgpe_set_kspace).
int build_softcore(fgpe_ctx* c)
{
    const size_t N = (size_t)c->ntot;
    const fgpe_hparams* hp = &c->hp;
    if (!c->wtab) FGPE_RET(dalloc(c, &c->wtab, c->tsz() * N));
    const int th = 256;
    dim3 grid((unsigned)((N + th - 1) / th));
    if (c->mixed())
        hipLaunchKernelGGL(k_softcore_table<float>, grid, dim3(th), 0, c->stream, (float*)c->wtab, c->k[0], c->k[1],
                           c->k[2], c->n[0], c->n[1], c->ntot, hp->sc_dim, hp->sc_R, hp->sc_U);
    else
        hipLaunchKernelGGL(k_softcore_table<double>, grid, dim3(th), 0, c->stream, (double*)c->wtab, c->k[0],
                           c->k[1], c->k[2], c->n[0], c->n[1], c->ntot, hp->sc_dim, hp->sc_R, hp->sc_U);
    FGPE_LAUNCHCHK("k_softcore_table");
    return FGPE_OK;
}

int ensure_hist(fgpe_ctx* c, size_t n)
{
    if (n > c->hist_n) {
        dfree(c, (void**)&c->hist, sizeof(double) * c->hist_n);
        c->hist_n = 0;
        FGPE_RET(dalloc(c, (void**)&c->hist, sizeof(double) * n));
        c->hist_n = n;
    }
    return FGPE_OK;
}

}  // namespace

// ---- batched groups and captured graphs ---------------------------------
// A group's per-instance argument tables live in one device buffer: the
// y-forward, last-axis, y-inverse, x-stage and finalize arguments of every
// instance, each table 256-byte aligned.
struct fgpe_group {
    std::vector<fgpe_ctx*> c;
    void* tab = nullptr;   // device argument tables of the last batched run
    size_t cap = 0;
};

struct fgpe_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    void* tab = nullptr;   // the argument tables the captured launches read
    // one capture stream per group: the groups are independent, so each is a
    // branch of the graph and their kernels may run side by side
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;   // the fork event, then one join event per group
};

namespace {

struct GroupTabs {
    size_t o_sf = 0, o_l = 0, o_si = 0, o_x = 0, o_f = 0, bytes = 0;
    std::vector<unsigned char> host;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// timing: no scale or stop in the passes and an RK stage-2 x-stage (reads P0,
// writes PS and K, so the state survives); else one heavy-ball iteration.
template <typename T> GroupTabs group_tabs(fgpe_group* g, double dt, double beta, const double* npart, double eps,
                                           bool timing)
{
    const size_t n = g->c.size();
    GroupTabs t;
    t.o_sf = 0;
    t.o_l = align256(t.o_sf + n * sizeof(StridedArgs));
    t.o_si = align256(t.o_l + n * sizeof(LastArgs));
    t.o_x = align256(t.o_si + n * sizeof(StridedArgs));
    t.o_f = align256(t.o_x + n * sizeof(XArgs<T>));
    t.bytes = align256(t.o_f + n * sizeof(FinArgs));
    t.host.assign(t.bytes, 0);
    // The group's block order (bit k: pass k, y fwd / last / y inv / x, runs
    // reversed): a reversed pass takes its tiles in reverse (the entries' rev)
    // and its instances too (its entries stored in reverse instance order, since
    // blockIdx.y indexes the table), so that it starts on what the pass before it
    // wrote last.
    const int ord = g->c[0]->order;
    auto slot = [&](int bit, size_t i) { return ((ord >> bit) & 1) ? n - 1 - i : i; };
    for (size_t i = 0; i < n; ++i) {
        fgpe_ctx* c = g->c[i];
        const int la = last_axis(c);
        const double* scale = timing ? nullptr : c->scal + SC_SCALE;
        const int* stop = timing ? nullptr : c->flags + FL_STOP;
        StridedArgs sf = strided_args(c, c->spec, 1, scale, c->nch, 1, true, stop);
        LastArgs l = last_args(c, c->spec, la, c->nch, 1, la == 2 ? nullptr : scale, has_filter(c->hp), stop,
                               true, !timing);
        StridedArgs si = strided_args(c, c->spec, 1, nullptr, c->nch, 1, true, stop);
        sf.rev = ord & 1;
        l.s.rev = (ord >> 1) & 1;
        si.rev = (ord >> 2) & 1;
        XArgs<T> x;
        if (timing) {
            x = base_xargs<T>(c, 0.0);
            x.cr = 0.0;
            x.ci = -1e-30;
            x.stage = 2;
        } else {
            x = hb_xargs<T>(c, 0.0, dt, beta);
        }
        x.rev = (ord >> 3) & 1;
        const FinArgs f = fin_args(c, Launch<T>::xblocks(c), FIN_HB, npart, c->scal + SC_TMP, eps);
        std::memcpy(&t.host[t.o_sf + slot(0, i) * sizeof(StridedArgs)], &sf, sizeof(sf));
        std::memcpy(&t.host[t.o_l + slot(1, i) * sizeof(LastArgs)], &l, sizeof(l));
        std::memcpy(&t.host[t.o_si + slot(2, i) * sizeof(StridedArgs)], &si, sizeof(si));
        std::memcpy(&t.host[t.o_x + slot(3, i) * sizeof(XArgs<T>)], &x, sizeof(x));
        std::memcpy(&t.host[t.o_f + i * sizeof(FinArgs)], &f, sizeof(f));
    }
    return t;
}

// Kernel classes of one batched iteration (cls = -1: all, then the finalize;
// -2: all but the finalize).
template <typename T> int group_launch(fgpe_group* g, const GroupTabs& t, const unsigned char* dtab, int cls,
                                       int mode)
{
    fgpe_ctx* c0 = g->c[0];
    const int n = (int)g->c.size();
    const int nc = 1, nch = c0->nch;
    const bool dens = nch > nc, filter = has_filter(c0->hp);
    const int la = last_axis(c0);
    if ((cls < 0 || cls == FGPE_KCLASS_STRIDED) && la == 2)
        FGPE_RET(Launch<T>::strided(c0, c0->spec, 1, -1, nch, nc, nullptr, nullptr, true, dtab + t.o_sf, n));
    if (cls < 0 || cls == FGPE_KCLASS_LAST)
        FGPE_RET(Launch<T>::last(c0, c0->spec, la, nch, nc, nullptr, filter, nullptr, true, dtab + t.o_l, n));
    if ((cls < 0 || cls == FGPE_KCLASS_STRIDED) && la == 2)
        FGPE_RET(Launch<T>::strided(c0, c0->spec, 1, +1, nch, nc, nullptr, nullptr, true, dtab + t.o_si, n));
    if (cls < 0 || cls == FGPE_KCLASS_XSTAGE)
        FGPE_RET(Launch<T>::xstage(c0, XArgs<T>{}, nc, dens, mode, dtab + t.o_x, n));
    if (cls == -1) {
        TScope ts(c0, FGPE_KCLASS_OTHER);
        hipLaunchKernelGGL(k_finalize, dim3((unsigned)n), dim3(kFinThreads), 0, c0->stream, FinArgs{},
                           (const FinArgs*)(dtab + t.o_f));
        FGPE_LAUNCHCHK("k_finalize");
    }
    return FGPE_OK;
}

int group_check(fgpe_group* g)
{
    fgpe_ctx* c0 = g->c[0];
    for (fgpe_ctx* c : g->c) {
        FGPE_RET(ensure_ready(c));
        if (c->hp.ncomp != 1) return fail(FGPE_EUNSUP, "group instances must have one component");
        if (c->hp.quench_on) return fail(FGPE_EUNSUP, "group instances cannot have an a(t) ramp");
        if (c->n[0] != c0->n[0] || c->n[1] != c0->n[1] || c->n[2] != c0->n[2])
            return fail(FGPE_EINVAL, "group instances must share the grid shape");
        if (c->prec != c0->prec) return fail(FGPE_EINVAL, "group instances must share the precision");
        if (c->stream != c0->stream) return fail(FGPE_EINVAL, "group instances must share the stream");
        if (c->nch != c0->nch) return fail(FGPE_EINVAL, "group instances must all have or all lack a k-space kernel");
        if (c->vcomplex[0]) return fail(FGPE_EUNSUP, "group instances need a real trap");
        if (c->hp.user_id != 0) return fail(FGPE_EUNSUP, "group instances run without a user term");
        if (last_axis(c) == 0) return fail(FGPE_EUNSUP, "groups of 1D grids are not supported");
    }
    return FGPE_OK;
}

// Uploads a table set after the stream has drained (earlier launches may
// still read the previous contents).
int group_upload(fgpe_ctx* c0, void** dtab, size_t* cap, const GroupTabs& t)
{
    FGPE_HIPCHK(hipStreamSynchronize(c0->stream));
    if (*cap < t.bytes) {
        if (*dtab) (void)hipFree(*dtab);
        *dtab = nullptr;
        *cap = 0;
        FGPE_HIPCHK(hipMalloc(dtab, t.bytes));
        *cap = t.bytes;
    }
    FGPE_HIPCHK(hipMemcpy(*dtab, t.host.data(), t.bytes, hipMemcpyHostToDevice));
    return FGPE_OK;
}

template <typename T> int group_imag(fgpe_group* g, double dt, double beta, const double* npart, double eps,
                                     int nsteps)
{
    FGPE_RET(group_check(g));
    for (fgpe_ctx* c : g->c) FGPE_RET(hb_prepare<T>(c));
    const GroupTabs t = group_tabs<T>(g, dt, beta, npart, eps, false);
    FGPE_RET(group_upload(g->c[0], &g->tab, &g->cap, t));
    for (int it = 0; it < nsteps; ++it)
        FGPE_RET(group_launch<T>(g, t, (const unsigned char*)g->tab, -1, MODE_HB));
    return FGPE_OK;
}

template <typename T> int group_autotune(fgpe_group* g, int reps, double* out)
{
    FGPE_RET(group_check(g));
    fgpe_ctx* c0 = g->c[0];
    // The timing x stages overwrite PS, which holds a heavy-ball run's old
    // state, so a run in progress ends here: its pending renormalisation is
    // applied, and the next run starts afresh from the normalised state.
    for (fgpe_ctx* c : g->c) {
        FGPE_RET(apply_pending_scale(c));
        c->hb_active = false;
    }
    const double one[2] = {1.0, 1.0};
    const GroupTabs t = group_tabs<T>(g, 0.0, 0.0, one, 0.0, true);
    FGPE_RET(group_upload(c0, &g->tab, &g->cap, t));
    const unsigned char* dt = (const unsigned char*)g->tab;
    const int la = last_axis(c0);
    hipEvent_t e0, e1;
    FGPE_HIPCHK(hipEventCreate(&e0));
    FGPE_HIPCHK(hipEventCreate(&e1));
    int rc = FGPE_OK;
    const int classes[3] = {FGPE_KCLASS_STRIDED, FGPE_KCLASS_LAST, FGPE_KCLASS_XSTAGE};
    for (int ci = 0; ci < 3 && rc == FGPE_OK; ++ci) {
        const int cls = classes[ci];
        if (cls == FGPE_KCLASS_STRIDED && la != 2) continue;
        const int axis = cls == FGPE_KCLASS_STRIDED ? 1 : (cls == FGPE_KCLASS_LAST ? la : 0);
        double best = 1e300;
        int bestv = 0;
        const int nv = cls == FGPE_KCLASS_XSTAGE && x_variant2(c0->n[0]) ? 3 : 2;
        for (int v = 0; v < nv && rc == FGPE_OK; ++v) {
            for (fgpe_ctx* c : g->c) c->var[cls][axis] = v;
            rc = group_launch<T>(g, t, dt, cls, MODE_RK);   // warm-up
            if (rc != FGPE_OK) break;
            (void)hipEventRecord(e0, c0->stream);
            for (int r = 0; r < reps && rc == FGPE_OK; ++r) rc = group_launch<T>(g, t, dt, cls, MODE_RK);
            (void)hipEventRecord(e1, c0->stream);
            if (hipEventSynchronize(e1) != hipSuccess) rc = fail(FGPE_EHIP, "autotune timing failed");
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double per = ms / reps;
            if (out) out[v < 2 ? 2 * ci + v : 9] = per;
            if (per < best) {
                best = per;
                bestv = v;
            }
        }
        for (fgpe_ctx* c : g->c) c->var[cls][axis] = bestv;
    }
    // Block order: the 16 combinations of passes run reversed (tiles and
    // instances), timed as whole batched iterations (at least 10 each: one
    // iteration of a small group takes about 0.1 ms), so that each pass starts
    // on what the pass before it wrote last; a reversal has to beat the
    // forward order by 1%.
    if (rc == FGPE_OK) {
        const int oreps = reps > 10 ? reps : 10;
        double t_fwd = 0.0, best = 1e300;
        int besto = 0;
        for (int o = 0; o < 16 && rc == FGPE_OK; ++o) {
            if (la != 2 && (o & 5)) continue;   // no y passes: bits 0 and 2 unused
            for (fgpe_ctx* c : g->c) c->order = o;
            const GroupTabs to = group_tabs<T>(g, 0.0, 0.0, one, 0.0, true);
            rc = group_upload(c0, &g->tab, &g->cap, to);
            if (rc != FGPE_OK) break;
            const unsigned char* dto = (const unsigned char*)g->tab;
            rc = group_launch<T>(g, to, dto, -2, MODE_RK);   // warm-up
            if (rc != FGPE_OK) break;
            (void)hipEventRecord(e0, c0->stream);
            for (int r = 0; r < oreps && rc == FGPE_OK; ++r) rc = group_launch<T>(g, to, dto, -2, MODE_RK);
            (void)hipEventRecord(e1, c0->stream);
            if (hipEventSynchronize(e1) != hipSuccess) rc = fail(FGPE_EHIP, "autotune timing failed");
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double per = ms / oreps;
            if (o == 0) t_fwd = per;
            if (per < best) {
                best = per;
                besto = o;
            }
        }
        if (best > 0.99 * t_fwd) besto = 0;
        for (fgpe_ctx* c : g->c) c->order = besto;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (fgpe_ctx* c : g->c) c->spec_valid = false;
    return rc;
}

}  // namespace

// =========================================================================
extern "C" {

int fgpe_user_term(int* enabled)
{
    if (!enabled) return fail(FGPE_EINVAL, "NULL argument");
    *enabled = user_term_count() > 0 ? 1 : 0;
    return FGPE_OK;
}

int fgpe_user_term_ids(int* ids, int cap, int* count)
{
    if (!count || (cap > 0 && !ids)) return fail(FGPE_EINVAL, "NULL argument");
    int n = 0;
#define FGPE_USER_TERM_ID(ID, S) \
    if (n < cap) ids[n] = ID;    \
    ++n;
    FGPE_USER_TERMS(FGPE_USER_TERM_ID)
#undef FGPE_USER_TERM_ID
    *count = n;
    return FGPE_OK;
}

const char* fgpe_last_error(void) { return g_err.c_str(); }

int fgpe_ctx_create(int nx, int ny, int nz, int precision, fgpe_ctx** out)
{
    if (!out) return fail(FGPE_EINVAL, "out is NULL");
    *out = nullptr;
    if (precision != FGPE_PREC_F64 && precision != FGPE_PREC_MIXED)
        return fail(FGPE_EINVAL, "precision must be FGPE_PREC_F64 or FGPE_PREC_MIXED");
    if (nx <= 1 || ny < 1 || nz < 1) return fail(FGPE_EINVAL, "grid extents must be positive (nx > 1)");
    if (!supported_n(nx) || (ny > 1 && !supported_n(ny)) || (nz > 1 && !supported_n(nz))) {
        std::string list;
#define FGPE_LIST_N(N_) list += (list.empty() ? "" : ", ") + std::to_string(N_);
        FGPE_FOR_EACH_N(FGPE_LIST_N)
#undef FGPE_LIST_N
        return fail(FGPE_EUNSUP, "grid extents must be 1 or one of " + list + " (nx, and ny, nz when above 1); got " +
                                     std::to_string(nx) + " x " + std::to_string(ny) + " x " + std::to_string(nz));
    }
    if (ny == 1 && nz > 1) return fail(FGPE_EUNSUP, "a grid with ny = 1 must have nz = 1");
    int dev = 0;
    {
        const hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return fail(FGPE_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    }
    // once per device: the address of the no-allocation entry point's scratch
    FGPE_RET(scratch_setup(dev));
    fgpe_ctx* c = new fgpe_ctx();
    c->dev = dev;
    c->n[0] = nx;
    c->n[1] = ny;
    c->n[2] = nz;
    c->ntot = (long long)nx * ny * nz;
    c->prec = precision;
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(FGPE_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->stream = c->own;
    int r = FGPE_OK;
    const size_t N = (size_t)c->ntot;
    for (int a = 0; a < 3 && r == FGPE_OK; ++a) {
        r = dalloc(c, (void**)&c->k[a], sizeof(double) * c->n[a]);
        if (r == FGPE_OK) r = dalloc(c, (void**)&c->k32[a], sizeof(float2) * c->n[a]);
        if (r == FGPE_OK) r = dalloc(c, (void**)&c->k64[a], sizeof(double2) * c->n[a]);
        if (r == FGPE_OK) r = dalloc(c, &c->tw64[a], sizeof(double2) * c->n[a]);
        if (r == FGPE_OK) r = dalloc(c, &c->tw32[a], sizeof(float2) * c->n[a]);
        if (r == FGPE_OK) {
            auto w64 = twiddle_table<double>(c->n[a]);
            auto w32 = twiddle_table<float>(c->n[a]);
            if (hipMemcpy(c->tw64[a], w64.data(), sizeof(double2) * c->n[a], hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(c->tw32[a], w32.data(), sizeof(float2) * c->n[a], hipMemcpyHostToDevice) != hipSuccess)
                r = fail(FGPE_EHIP, "twiddle upload failed");
        }
    }
    if (r == FGPE_OK) r = dalloc(c, (void**)&c->P0[0], sizeof(double2) * N);
    if (r == FGPE_OK) r = dalloc(c, &c->PS[0], 2 * c->tsz() * N);
    if (r == FGPE_OK && c->mixed()) r = dalloc(c, &c->PSb[0], 2 * c->tsz() * N);
    if (r == FGPE_OK) r = dalloc(c, &c->K[0], 2 * c->tsz() * N);
    if (r == FGPE_OK) r = dalloc(c, (void**)&c->scal, sizeof(double) * SC_COUNT);
    if (r == FGPE_OK) r = dalloc(c, (void**)&c->flags, sizeof(int) * FL_COUNT);
    if (r == FGPE_OK) {
        c->partials_n = 6 * (size_t)(c->ntot / 32 + 1024);
        r = dalloc(c, (void**)&c->partials, sizeof(double) * c->partials_n);
    }
    if (r == FGPE_OK) {
        double init[SC_COUNT] = {1.0, 1.0};
        int fl[FL_COUNT] = {0};
        if (hipMemcpy(c->scal, init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->flags, fl, sizeof(fl), hipMemcpyHostToDevice) != hipSuccess)
            r = fail(FGPE_EHIP, "scalar init failed");
    }
    if (r != FGPE_OK) {
        std::string msg = g_err;
        fgpe_ctx_destroy(c);
        return fail(r, msg);
    }
    *out = c;
    return FGPE_OK;
}

int fgpe_ctx_destroy(fgpe_ctx* c)
{
    if (!c) return FGPE_OK;
    // work queued by this context may still read its buffers
    (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->k[0], c->k[1], c->k[2], c->k32[0], c->k32[1], c->k32[2], c->k64[0], c->k64[1], c->k64[2], c->tw64[0], c->tw64[1], c->tw64[2], c->tw32[0], c->tw32[1],
                    c->tw32[2], c->P0[0], c->P0[1], c->PS[0], c->PS[1], c->PSb[0], c->PSb[1], c->K[0], c->K[1], c->Vr[0], c->Vr[1],
                    c->Vi[0], c->Vi[1], c->spec, c->wtab, c->partials, c->scal, c->flags, c->hist,
                    c->aux, c->xbuf, c->vstage[0], c->vstage[1], c->axbuf, c->spec3, c->lam[0],
                    c->lam[1], c->lam[2], c->pacc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (double2* p : c->ckpt) (void)hipFree(p);
    for (double2* p : c->segbuf) (void)hipFree(p);
    for (void* p : c->ystep) (void)hipFree(p);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return FGPE_OK;
}

int fgpe_ctx_set_stream(fgpe_ctx* c, void* s)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    c->stream = (hipStream_t)s;
    return FGPE_OK;
}

int fgpe_ctx_sync(fgpe_ctx* c)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    return FGPE_OK;
}

int fgpe_ctx_device_bytes(fgpe_ctx* c, size_t* bytes)
{
    if (!c || !bytes) return fail(FGPE_EINVAL, "NULL argument");
    *bytes = c->bytes;
    return FGPE_OK;
}

int fgpe_set_kspace(fgpe_ctx* c, const double* kx, const double* ky, const double* kz, double dv)
{
    if (!c || !kx || !ky || !kz) return fail(FGPE_EINVAL, "NULL argument");
    // launches still queued on the context's stream read the k tables
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    const double* src[3] = {kx, ky, kz};
    for (int a = 0; a < 3; ++a) {
        FGPE_HIPCHK(hipMemcpy(c->k[a], src[a], sizeof(double) * c->n[a], hipMemcpyHostToDevice));
        const int n = c->n[a];
        std::vector<float2> f(n);
        std::vector<double2> d(n);
        for (int i = 0; i < n; ++i) {
            f[i] = float2{(float)src[a][i], (float)src[a][(n - i) % n]};
            d[i] = double2{src[a][i], src[a][(n - i) % n]};
        }
        FGPE_HIPCHK(hipMemcpy(c->k32[a], f.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
        FGPE_HIPCHK(hipMemcpy(c->k64[a], d.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    }
    c->dv = dv;
    c->kset = true;
    c->spec_valid = false;
    if (c->hset && c->hp.sc_U != 0.0) FGPE_RET(build_softcore(c));
    return FGPE_OK;
}

int fgpe_set_hamiltonian(fgpe_ctx* c, const fgpe_hparams* hp)
{
    if (!c || !hp) return fail(FGPE_EINVAL, "NULL argument");
    if (hp->ncomp != 1 && hp->ncomp != 2) return fail(FGPE_EINVAL, "ncomp must be 1 or 2");
    if (hp->sc_U != 0.0 && (hp->sc_dim < 1 || hp->sc_dim > 3)) return fail(FGPE_EINVAL, "sc_dim must be 1, 2 or 3");
    if (hp->sc_U != 0.0 && !c->kset) return fail(FGPE_ESTATE, "set the k space before a soft-core term");
    if (hp->ncomp == 2 && hp->quench_on) return fail(FGPE_EUNSUP, "the a(t) ramp is one-component only");
    if (hp->user_id != 0 && !user_term_compiled(hp->user_id))
        return fail(FGPE_EINVAL, "user term " + std::to_string(hp->user_id) + " is not compiled into this library");
    const int nch = hp->ncomp + (has_filter(*hp) ? 1 : 0);
    // the half-spectrum density channel must fit its ntot-element slot
    if (nch > hp->ncomp && dens_pitch(c) * c->n[1] * c->n[2] > c->ntot)
        return fail(FGPE_EUNSUP, "density channel rows wider than the grid's x extent");
    c->hp = *hp;
    const size_t N = (size_t)c->ntot;
    if (nch != c->nch) {
        dfree(c, &c->spec, 2 * c->tsz() * N * c->nch);
        FGPE_RET(dalloc(c, &c->spec, 2 * c->tsz() * N * nch));
        c->nch = nch;
    }
    if (hp->ncomp == 2) {
        if (!c->P0[1]) FGPE_RET(dalloc(c, (void**)&c->P0[1], sizeof(double2) * N));
        if (!c->PS[1]) FGPE_RET(dalloc(c, &c->PS[1], 2 * c->tsz() * N));
        if (c->mixed() && !c->PSb[1]) FGPE_RET(dalloc(c, &c->PSb[1], 2 * c->tsz() * N));
        if (!c->K[1]) FGPE_RET(dalloc(c, &c->K[1], 2 * c->tsz() * N));
    }
    if (hp->sc_U != 0.0) {
        FGPE_RET(build_softcore(c));
    } else {
        dfree(c, &c->wtab, c->tsz() * N);
    }
    c->hset = true;
    c->spec_valid = false;
    c->hb_active = false;
    c->imag_iter = 0;
    return FGPE_OK;
}

int fgpe_upload_psi(fgpe_ctx* c, int comp, const double* host)
{
    if (!c || !host) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1) return fail(FGPE_EINVAL, "comp must be 0 or 1");
    if (!c->P0[comp]) FGPE_RET(dalloc(c, (void**)&c->P0[comp], sizeof(double2) * c->ntot));
    FGPE_HIPCHK(hipMemcpyAsync(c->P0[comp], host, sizeof(double2) * c->ntot, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_reset_scales, dim3(1), dim3(1), 0, c->stream, c->scal + SC_SCALE);
    FGPE_LAUNCHCHK("k_reset_scales");
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    c->spec_valid = false;
    c->hb_active = false;
    c->imag_iter = 0;
    return FGPE_OK;
}

int fgpe_download_psi(fgpe_ctx* c, int comp, double* host)
{
    if (!c || !host) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1 || !c->P0[comp]) return fail(FGPE_EINVAL, "no such component");
    FGPE_HIPCHK(hipMemcpyAsync(host, c->P0[comp], sizeof(double2) * c->ntot, hipMemcpyDeviceToHost, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    return FGPE_OK;
}

int fgpe_upload_vext(fgpe_ctx* c, int comp, const double* host)
{
    if (!c || !host) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1) return fail(FGPE_EINVAL, "comp must be 0 or 1");
    const size_t N = (size_t)c->ntot;
    // launches still queued on the context's stream read the trap
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    bool cplx = false;
    for (size_t i = 0; i < N && !cplx; ++i) cplx = host[2 * i + 1] != 0.0;
    if (!c->Vr[comp]) FGPE_RET(dalloc(c, &c->Vr[comp], c->tsz() * N));
    if (cplx && !c->Vi[comp]) FGPE_RET(dalloc(c, &c->Vi[comp], c->tsz() * N));
    if (c->mixed()) {
        std::vector<float> re(N), im(cplx ? N : 0);
        for (size_t i = 0; i < N; ++i) {
            re[i] = (float)host[2 * i];
            if (cplx) im[i] = (float)host[2 * i + 1];
        }
        FGPE_HIPCHK(hipMemcpy(c->Vr[comp], re.data(), 4 * N, hipMemcpyHostToDevice));
        if (cplx) FGPE_HIPCHK(hipMemcpy(c->Vi[comp], im.data(), 4 * N, hipMemcpyHostToDevice));
    } else {
        std::vector<double> re(N), im(cplx ? N : 0);
        for (size_t i = 0; i < N; ++i) {
            re[i] = host[2 * i];
            if (cplx) im[i] = host[2 * i + 1];
        }
        FGPE_HIPCHK(hipMemcpy(c->Vr[comp], re.data(), 8 * N, hipMemcpyHostToDevice));
        if (cplx) FGPE_HIPCHK(hipMemcpy(c->Vi[comp], im.data(), 8 * N, hipMemcpyHostToDevice));
    }
    c->vcomplex[comp] = cplx;
    return FGPE_OK;
}

int fgpe_upload_psi_dev(fgpe_ctx* c, int comp, const void* dev)
{
    if (!c || !dev) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1) return fail(FGPE_EINVAL, "comp must be 0 or 1");
    if (!c->P0[comp]) FGPE_RET(dalloc(c, (void**)&c->P0[comp], sizeof(double2) * c->ntot));
    FGPE_HIPCHK(hipMemcpyAsync(c->P0[comp], dev, sizeof(double2) * c->ntot, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_reset_scales, dim3(1), dim3(1), 0, c->stream, c->scal + SC_SCALE);
    FGPE_LAUNCHCHK("k_reset_scales");
    c->spec_valid = false;
    c->hb_active = false;
    c->imag_iter = 0;
    return FGPE_OK;
}

int fgpe_download_psi_dev(fgpe_ctx* c, int comp, void* dev)
{
    if (!c || !dev) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1 || !c->P0[comp]) return fail(FGPE_EINVAL, "no such component");
    FGPE_HIPCHK(hipMemcpyAsync(dev, c->P0[comp], sizeof(double2) * c->ntot, hipMemcpyDeviceToDevice, c->stream));
    return FGPE_OK;
}

int fgpe_upload_vext_real_dev(fgpe_ctx* c, int comp, const double* dev)
{
    if (!c || !dev) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1) return fail(FGPE_EINVAL, "comp must be 0 or 1");
    const size_t N = (size_t)c->ntot;
    if (!c->Vr[comp]) FGPE_RET(dalloc(c, &c->Vr[comp], c->tsz() * N));
    if (c->mixed()) {
        const long long n = c->ntot;
        hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           (float*)c->Vr[comp], dev, n);
        FGPE_LAUNCHCHK("k_convert");
    } else {
        FGPE_HIPCHK(hipMemcpyAsync(c->Vr[comp], dev, 8 * N, hipMemcpyDeviceToDevice, c->stream));
    }
    c->vcomplex[comp] = false;
    return FGPE_OK;
}

int fgpe_psi_device_ptr(fgpe_ctx* c, int comp, void** p)
{
    if (!c || !p) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1 || !c->P0[comp]) return fail(FGPE_EINVAL, "no such component");
    *p = c->P0[comp];
    return FGPE_OK;
}

int fgpe_fft3d_c2c_f64(fgpe_ctx* c, const void* in, void* out, int dir)
{
    if (!c || !in || !out) return fail(FGPE_EINVAL, "NULL argument");
    if (dir != -1 && dir != 1) return fail(FGPE_EINVAL, "direction must be -1 or +1");
    return fft3d<double>(c, in, out, dir);
}

int fgpe_fft3d_c2c_f32(fgpe_ctx* c, const void* in, void* out, int dir)
{
    if (!c || !in || !out) return fail(FGPE_EINVAL, "NULL argument");
    if (dir != -1 && dir != 1) return fail(FGPE_EINVAL, "direction must be -1 or +1");
    return fft3d<float>(c, in, out, dir);
}

int fgpe_laplacian_f64(fgpe_ctx* c, const void* psi, void* out)
{
    if (!c || !psi || !out) return fail(FGPE_EINVAL, "NULL argument");
    if (!c->kset) return fail(FGPE_ESTATE, "fgpe_set_kspace has not been called");
    FGPE_RET(Launch<double>::fft_x(c, psi, out, -1));
    const int la = last_axis(c);
    if (la == 0) FGPE_RET(kmul_1d<double>(c, out, 1, 1, nullptr, false, nullptr, false, false));
    if (la == 2) FGPE_RET(Launch<double>::strided(c, out, 1, -1, 1, 1, nullptr, nullptr, false));
    if (la > 0) FGPE_RET(Launch<double>::last(c, out, la, 1, 1, nullptr, false, nullptr, false));
    if (la == 2) FGPE_RET(Launch<double>::strided(c, out, 1, +1, 1, 1, nullptr, nullptr, false));
    FGPE_RET(Launch<double>::fft_x(c, out, out, +1));
    return FGPE_OK;
}

int fgpe_apply_h(fgpe_ctx* c, double t, void* hmu, void* hen)
{
    if (!c || !hmu) return fail(FGPE_EINVAL, "NULL argument");
    return c->mixed() ? apply_h<float>(c, t, hmu, hen, nullptr) : apply_h<double>(c, t, hmu, hen, nullptr);
}

int fgpe_observables(fgpe_ctx* c, double t, double* out)
{
    if (!c || !out) return fail(FGPE_EINVAL, "NULL argument");
    double* obs = c->scal + SC_OBS;
    FGPE_RET(c->mixed() ? apply_h<float>(c, t, nullptr, nullptr, obs) : apply_h<double>(c, t, nullptr, nullptr, obs));
    FGPE_HIPCHK(hipMemcpyAsync(out, obs, 3 * c->hp.ncomp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    return FGPE_OK;
}

int fgpe_rk4_steps(fgpe_ctx* c, double dt, double t0, int nsteps)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    c->hb_active = false;
    c->imag_iter = 0;
    if (nsteps < 0) return fail(FGPE_EINVAL, "nsteps < 0");
    const double one[2] = {1.0, 1.0};
    return c->mixed() ? rk4_run<float>(c, dt, t0, nsteps, false, one, nullptr)
                      : rk4_run<double>(c, dt, t0, nsteps, false, one, nullptr);
}

}  // extern "C"

namespace {

int imag_check(fgpe_ctx* c, int scheme, const double* npart, int nsteps)
{
    if (!c || !npart) return fail(FGPE_EINVAL, "NULL argument");
    if (nsteps < 0) return fail(FGPE_EINVAL, "nsteps < 0");
    if (scheme != FGPE_IMAG_RK4_RENORM && scheme != FGPE_IMAG_HEAVY_BALL)
        return fail(FGPE_EINVAL, "unknown imaginary-time scheme");
    if (!c->hset) return fail(FGPE_ESTATE, "fgpe_set_hamiltonian has not been called");
    return FGPE_OK;
}

// The imaginary-time steps with the history written to the device buffer hd
// (3 * ncomp * nsteps doubles, or null), then copied to host_hist.
int imag_steps_on(fgpe_ctx* c, int scheme, double dt, double beta, const double* npart, double eps, int nsteps,
                  double* hd, double* host_hist, int* steps_done)
{
    const int nc = c->hp.ncomp;
    if (scheme == FGPE_IMAG_RK4_RENORM) {
        c->hb_active = false;
        const double t0 = (double)c->imag_iter * dt;
        FGPE_RET(c->mixed() ? rk4_run<float>(c, dt, t0, nsteps, true, npart, hd)
                            : rk4_run<double>(c, dt, t0, nsteps, true, npart, hd));
    } else {
        FGPE_RET(c->mixed() ? hb_run<float>(c, dt, beta, npart, eps, nsteps, hd)
                            : hb_run<double>(c, dt, beta, npart, eps, nsteps, hd));
    }
    FGPE_RET(apply_pending_scale(c));
    int done = nsteps;
    if (scheme == FGPE_IMAG_HEAVY_BALL)
        FGPE_HIPCHK(hipMemcpyAsync(&done, c->flags + FL_DONE, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    c->imag_iter += done;
    if (hd && done > 0)
        FGPE_HIPCHK(hipMemcpy(host_hist, hd, sizeof(double) * 3 * nc * done, hipMemcpyDeviceToHost));
    if (steps_done) *steps_done = done;
    return FGPE_OK;
}

}  // namespace

extern "C" {

int fgpe_imag_steps(fgpe_ctx* c, int scheme, double dt, double beta, const double* npart, double eps,
                    int nsteps, double* host_hist, int* steps_done)
{
    FGPE_RET(imag_check(c, scheme, npart, nsteps));
    double* hd = nullptr;
    if (host_hist && nsteps > 0) {
        FGPE_RET(ensure_hist(c, (size_t)3 * c->hp.ncomp * nsteps));
        hd = c->hist;
    }
    return imag_steps_on(c, scheme, dt, beta, npart, eps, nsteps, hd, host_hist, steps_done);
}

int fgpe_imag_steps_noalloc(fgpe_ctx* c, int scheme, double dt, double beta, const double* npart, double eps,
                            int nsteps, double* host_hist, int* steps_done)
{
    FGPE_RET(imag_check(c, scheme, npart, nsteps));
    double* hd = nullptr;
    if (host_hist && nsteps > 0) {
        if (nsteps > FGPE_NOALLOC_MAX_STEPS)
            return fail(FGPE_EINVAL, "fgpe_imag_steps_noalloc: nsteps " + std::to_string(nsteps) +
                                         " with a history exceeds FGPE_NOALLOC_MAX_STEPS (" +
                                         std::to_string(FGPE_NOALLOC_MAX_STEPS) + ")");
        hd = scratch_of(c->dev);
        if (!hd) return fail(FGPE_ESTATE, "no scratch address for the context's device");
    }
    return imag_steps_on(c, scheme, dt, beta, npart, eps, nsteps, hd, host_hist, steps_done);
}

int fgpe_calc_px(fgpe_ctx* c, int comp, double* out)
{
    if (!c || !out) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1 || !c->P0[comp]) return fail(FGPE_EINVAL, "no such component");
    if (!c->kset) return fail(FGPE_ESTATE, "fgpe_set_kspace has not been called");
    if (!c->aux) FGPE_RET(dalloc(c, (void**)&c->aux, sizeof(double2) * c->ntot));
    FGPE_RET(fft3d<double>(c, c->P0[comp], c->aux, -1));
    return weighted_sum(c, c->aux, c->k[0], 0, 1.0, out);
}

int fgpe_calc_cos(fgpe_ctx* c, int comp, const double* host_x, double L, double* out)
{
    if (!c || !out || !host_x) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1 || !c->P0[comp]) return fail(FGPE_EINVAL, "no such component");
    if (!c->xbuf) FGPE_RET(dalloc(c, (void**)&c->xbuf, sizeof(double) * c->n[0]));
    FGPE_HIPCHK(hipMemcpy(c->xbuf, host_x, sizeof(double) * c->n[0], hipMemcpyHostToDevice));
    return weighted_sum(c, c->P0[comp], c->xbuf, 1, L, out);
}

int fgpe_adjoint_forward(fgpe_ctx* c, double dt, double t0, int nsteps, int ckpt_every)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (nsteps < 1 || ckpt_every < 1) return fail(FGPE_EINVAL, "nsteps and ckpt_every must be positive");
    c->hb_active = false;
    c->adj_ready = false;
    return c->mixed() ? adj_forward<float>(c, dt, t0, nsteps, ckpt_every)
                      : adj_forward<double>(c, dt, t0, nsteps, ckpt_every);
}

int fgpe_adjoint_backward(fgpe_ctx* c, const void* dev_lamT, void* dev_lam0, double* dev_gradV, double* host_grad)
{
    if (!c || !dev_lamT || !dev_gradV || !host_grad) return fail(FGPE_EINVAL, "NULL argument");
    c->hb_active = false;
    return c->mixed() ? adj_backward<float>(c, dev_lamT, dev_lam0, dev_gradV, host_grad)
                      : adj_backward<double>(c, dev_lamT, dev_lam0, dev_gradV, host_grad);
}

int fgpe_autotune(fgpe_ctx* c, int reps, double* host_ms)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (reps < 1) return fail(FGPE_EINVAL, "reps must be positive");
    if (host_ms)
        for (int i = 0; i < 9; ++i) host_ms[i] = 0.0;
    c->hb_active = false;
    return c->mixed() ? autotune<float>(c, reps, host_ms) : autotune<double>(c, reps, host_ms);
}

int fgpe_set_variant(fgpe_ctx* c, int kclass, int axis, int variant)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (kclass < 0 || kclass >= FGPE_NKCLASS || axis < 0 || axis > 2 || variant < 0 ||
        variant > (kclass == FGPE_KCLASS_XSTAGE ? 2 : 1))
        return fail(FGPE_EINVAL, "kclass, axis or variant out of range");
    c->var[kclass][axis] = variant;
    c->spec_valid = false;
    return FGPE_OK;
}

int fgpe_get_variant(fgpe_ctx* c, int kclass, int axis, int* variant)
{
    if (!c || !variant) return fail(FGPE_EINVAL, "NULL argument");
    if (kclass < 0 || kclass >= FGPE_NKCLASS || axis < 0 || axis > 2)
        return fail(FGPE_EINVAL, "kclass or axis out of range");
    *variant = c->var[kclass][axis];
    return FGPE_OK;
}

int fgpe_set_order(fgpe_ctx* c, int mask)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (mask < 0 || mask > 15) return fail(FGPE_EINVAL, "order mask must be in 0..15");
    c->order = mask;
    return FGPE_OK;
}

int fgpe_get_order(fgpe_ctx* c, int* mask)
{
    if (!c || !mask) return fail(FGPE_EINVAL, "NULL argument");
    *mask = c->order;
    return FGPE_OK;
}

int fgpe_set_kmul_axis(fgpe_ctx* c, int axis)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (c->n[2] > 1) {
        if (axis != 1 && axis != 2) return fail(FGPE_EINVAL, "3D grids take the k-space multiply on axis 1 or 2");
    } else if (axis != last_axis(c)) {
        return fail(FGPE_EINVAL, "a 1D or 2D grid has one k-space multiply axis");
    }
    c->ylast = c->n[2] > 1 && axis == 1;
    c->spec_valid = false;
    return FGPE_OK;
}

int fgpe_get_kmul_axis(fgpe_ctx* c, int* axis)
{
    if (!c || !axis) return fail(FGPE_EINVAL, "NULL argument");
    *axis = kmul_axis(c);
    return FGPE_OK;
}

int fgpe_set_slabs(fgpe_ctx* c, int nslab)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (nslab != 1 && nslab != 2 && nslab != 4) return fail(FGPE_EINVAL, "nslab must be 1, 2 or 4");
    if (c->n[0] % nslab) return fail(FGPE_EUNSUP, "nx not divisible by nslab");
    c->nslab = nslab;
    return FGPE_OK;
}

int fgpe_get_slabs(fgpe_ctx* c, int* nslab)
{
    if (!c || !nslab) return fail(FGPE_EINVAL, "NULL argument");
    *nslab = c->nslab;
    return FGPE_OK;
}

int fgpe_imag_converged(fgpe_ctx* c, int* converged)
{
    if (!c || !converged) return fail(FGPE_EINVAL, "NULL argument");
    int f = 0;
    FGPE_HIPCHK(hipMemcpyAsync(&f, c->flags + FL_STOP, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    *converged = (c->hb_active && f) ? 1 : 0;
    return FGPE_OK;
}

int fgpe_timing_enable(fgpe_ctx* c, int on)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    c->timing = on != 0;
    c->evused = 0;
    c->evrec.clear();
    if (c->timing && c->evpool.empty()) {
        c->evpool.resize(16384);
        for (auto& e : c->evpool) FGPE_HIPCHK(hipEventCreate(&e));
    }
    return FGPE_OK;
}

int fgpe_timing_read(fgpe_ctx* c, double* out)
{
    if (!c || !out) return fail(FGPE_EINVAL, "NULL argument");
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 2 * FGPE_NKCLASS; ++k) out[k] = 0.0;
    for (auto& r : c->evrec) {
        float ms = 0.f;
        FGPE_HIPCHK(hipEventElapsedTime(&ms, c->evpool[r.second], c->evpool[r.second + 1]));
        out[2 * r.first] += ms;
        out[2 * r.first + 1] += 1.0;
    }
    c->evused = 0;
    c->evrec.clear();
    return FGPE_OK;
}

int fgpe_group_create(fgpe_ctx* const* ctxs, int n, fgpe_group** out)
{
    if (!ctxs || !out) return fail(FGPE_EINVAL, "NULL argument");
    *out = nullptr;
    if (n < 1) return fail(FGPE_EINVAL, "a group needs at least one context");
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(FGPE_EINVAL, "NULL context in the group");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(FGPE_EINVAL, "a context appears twice in the group");
    }
    fgpe_group* g = new fgpe_group();
    g->c.assign(ctxs, ctxs + n);
    *out = g;
    return FGPE_OK;
}

int fgpe_group_destroy(fgpe_group* g)
{
    if (!g) return FGPE_OK;
    if (g->tab) {
        (void)hipStreamSynchronize(g->c[0]->stream);
        (void)hipFree(g->tab);
    }
    delete g;
    return FGPE_OK;
}

int fgpe_group_autotune(fgpe_group* g, int reps, double* host_ms)
{
    if (!g) return fail(FGPE_EINVAL, "group is NULL");
    if (reps < 1) return fail(FGPE_EINVAL, "reps < 1");
    if (host_ms)
        for (int i = 0; i < 9; ++i) host_ms[i] = 0.0;
    return g->c[0]->mixed() ? group_autotune<float>(g, reps, host_ms) : group_autotune<double>(g, reps, host_ms);
}

int fgpe_group_imag_steps(fgpe_group* g, double dt, double beta, const double* npart, double eps, int nsteps)
{
    if (!g || !npart) return fail(FGPE_EINVAL, "NULL argument");
    if (nsteps < 0) return fail(FGPE_EINVAL, "nsteps < 0");
    return g->c[0]->mixed() ? group_imag<float>(g, dt, beta, npart, eps, nsteps)
                            : group_imag<double>(g, dt, beta, npart, eps, nsteps);
}

int fgpe_group_imag_status(fgpe_group* g, int* iters, int* converged)
{
    if (!g) return fail(FGPE_EINVAL, "group is NULL");
    const size_t n = g->c.size();
    std::vector<int> fl(FL_COUNT * n);
    for (size_t i = 0; i < n; ++i) {
        fgpe_ctx* c = g->c[i];
        FGPE_HIPCHK(hipMemcpyAsync(&fl[FL_COUNT * i], c->flags, sizeof(int) * FL_COUNT, hipMemcpyDeviceToHost,
                                   c->stream));
    }
    for (size_t i = 0; i < n; ++i) FGPE_HIPCHK(hipStreamSynchronize(g->c[i]->stream));
    for (size_t i = 0; i < n; ++i) {
        if (iters) iters[i] = fl[FL_COUNT * i + FL_DONE];
        if (converged) converged[i] = (g->c[i]->hb_active && fl[FL_COUNT * i + FL_STOP]) ? 1 : 0;
    }
    return FGPE_OK;
}

int fgpe_group_imag_finish(fgpe_group* g, int* iters, int* converged)
{
    if (!g) return fail(FGPE_EINVAL, "group is NULL");
    for (fgpe_ctx* c : g->c) FGPE_RET(apply_pending_scale(c));
    return fgpe_group_imag_status(g, iters, converged);
}

int fgpe_graph_capture_imag(fgpe_group* const* groups, int ngroups, double dt, double beta,
                            const double* npart, double eps, int iters, fgpe_graph** out)
{
    if (!groups || !npart || !out) return fail(FGPE_EINVAL, "NULL argument");
    *out = nullptr;
    if (ngroups < 1 || iters < 1) return fail(FGPE_EINVAL, "need at least one group and one iteration");
    hipStream_t s = groups[0]->c[0]->stream;
    if (!s) return fail(FGPE_EINVAL, "the null stream cannot be captured; set a stream on the contexts");
    std::vector<GroupTabs> tabs(ngroups);
    size_t total = 0;
    std::vector<size_t> base(ngroups);
    for (int k = 0; k < ngroups; ++k) {
        fgpe_group* g = groups[k];
        if (!g) return fail(FGPE_EINVAL, "NULL group");
        FGPE_RET(group_check(g));
        if (g->c[0]->stream != s) return fail(FGPE_EINVAL, "all groups must share one stream");
        for (fgpe_ctx* c : g->c) {
            c->timing = false;
            FGPE_RET(c->mixed() ? hb_prepare<float>(c) : hb_prepare<double>(c));
        }
        tabs[k] = g->c[0]->mixed() ? group_tabs<float>(g, dt, beta, npart, eps, false)
                                   : group_tabs<double>(g, dt, beta, npart, eps, false);
        base[k] = total;
        total += tabs[k].bytes;
    }
    fgpe_graph* gr = new fgpe_graph();
    gr->stream = s;
    int rc = FGPE_OK;
    for (int k = 0; k < ngroups && rc == FGPE_OK; ++k) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            rc = fail(FGPE_EHIP, "hipStreamCreate for a graph branch failed");
        else
            gr->side.push_back(st);
    }
    for (int k = 0; k <= ngroups && rc == FGPE_OK; ++k) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            rc = fail(FGPE_EHIP, "hipEventCreate for a graph branch failed");
        else
            gr->ev.push_back(e);
    }
    if (rc == FGPE_OK && hipMalloc(&gr->tab, total) != hipSuccess)
        rc = fail(FGPE_EHIP, "hipMalloc of the graph tables failed");
    for (int k = 0; k < ngroups && rc == FGPE_OK; ++k)
        if (hipMemcpy((unsigned char*)gr->tab + base[k], tabs[k].host.data(), tabs[k].bytes,
                      hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(FGPE_EHIP, "graph table upload failed");
    if (rc == FGPE_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(FGPE_EHIP, "stream synchronise failed");
    if (rc == FGPE_OK && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
        rc = fail(FGPE_EHIP, "hipStreamBeginCapture failed");
    if (rc == FGPE_OK) {
        // fork: every group's iterations on a branch of its own, joined at the end
        if (hipEventRecord(gr->ev[0], s) != hipSuccess) rc = fail(FGPE_EHIP, "fork event failed");
        for (int k = 0; k < ngroups && rc == FGPE_OK; ++k) {
            fgpe_group* g = groups[k];
            if (hipStreamWaitEvent(gr->side[k], gr->ev[0], 0) != hipSuccess) {
                rc = fail(FGPE_EHIP, "branch wait failed");
                break;
            }
            for (fgpe_ctx* c : g->c) c->stream = gr->side[k];
            const unsigned char* d = (const unsigned char*)gr->tab + base[k];
            for (int it = 0; it < iters && rc == FGPE_OK; ++it)
                rc = g->c[0]->mixed() ? group_launch<float>(g, tabs[k], d, -1, MODE_HB)
                                      : group_launch<double>(g, tabs[k], d, -1, MODE_HB);
            for (fgpe_ctx* c : g->c) c->stream = s;
            if (rc == FGPE_OK && (hipEventRecord(gr->ev[k + 1], gr->side[k]) != hipSuccess ||
                                  hipStreamWaitEvent(s, gr->ev[k + 1], 0) != hipSuccess))
                rc = fail(FGPE_EHIP, "branch join failed");
        }
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc == FGPE_OK && e != hipSuccess)
            rc = fail(FGPE_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        gr->graph = graph;
    }
    if (rc == FGPE_OK && hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = fail(FGPE_EHIP, "hipGraphInstantiate failed");
    if (rc != FGPE_OK) {
        const std::string msg = g_err;
        fgpe_graph_destroy(gr);
        return fail(rc, msg);
    }
    *out = gr;
    return FGPE_OK;
}

int fgpe_graph_launch(fgpe_graph* g)
{
    if (!g || !g->exec) return fail(FGPE_EINVAL, "graph is NULL");
    FGPE_HIPCHK(hipGraphLaunch(g->exec, g->stream));
    return FGPE_OK;
}

int fgpe_graph_destroy(fgpe_graph* g)
{
    if (!g) return FGPE_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->tab) (void)hipFree(g->tab);
    for (hipEvent_t e : g->ev) (void)hipEventDestroy(e);
    for (hipStream_t st : g->side) (void)hipStreamDestroy(st);
    delete g;
    return FGPE_OK;
}

}  // extern "C"
