// fgpe_capi.hip -- the C ABI of include/fgpe_c.h: context (device buffers,
// twiddle tables, stream), launch dispatch over the supported pencil lengths,
// and the time-stepping drivers.  gfx950 only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fgpe_c.h"
#include "gpe_kernels.hpp"

using namespace fgpe;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(FGPE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK(what)                                                                \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess)                                                          \
            return fail(FGPE_EHIP, std::string(what) + ": " + hipGetErrorString(e_));  \
    } while (0)

#define FGPE_FOR_EACH_N(X) X(32) X(64) X(96) X(128) X(192) X(256) X(512)

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

}  // namespace

struct fgpe_ctx {
    int n[3] = {1, 1, 1};
    long long ntot = 0;
    int prec = FGPE_PREC_F64;
    hipStream_t own = nullptr, stream = nullptr;
    size_t bytes = 0;

    double* k[3] = {nullptr, nullptr, nullptr};
    double dv = 0.0;
    bool kset = false;

    fgpe_hparams hp{};
    bool hset = false;
    int nch = 0;

    double2* P0[2] = {nullptr, nullptr};
    void* PS[2] = {nullptr, nullptr};
    void* K[2] = {nullptr, nullptr};
    void* Vr[2] = {nullptr, nullptr};
    void* Vi[2] = {nullptr, nullptr};
    bool vcomplex[2] = {false, false};
    void* spec = nullptr;
    void* wtab = nullptr;
    double* partials = nullptr;
    size_t partials_n = 0;
    double* scal = nullptr;   // [0] scale, [1..15] observables scratch
    double* hist = nullptr;
    size_t hist_n = 0;
    bool spec_valid = false;

    void* tw64[3] = {nullptr, nullptr, nullptr};
    void* tw32[3] = {nullptr, nullptr, nullptr};

    // launch timing (fgpe_timing_enable)
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    std::vector<std::pair<int, size_t>> evrec;

    bool mixed() const { return prec == FGPE_PREC_MIXED; }
    size_t tsz() const { return mixed() ? 4 : 8; }
};

namespace {

template <typename T> void* tw_of(fgpe_ctx* c, int axis) { return sizeof(T) == 8 ? c->tw64[axis] : c->tw32[axis]; }

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

// ---- launch timing --------------------------------------------------------
// Returns the event-pair slot to close with t_end, or -1 when not timing.
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
struct TScope {
    fgpe_ctx* c;
    int cls;
    long k;
    TScope(fgpe_ctx* c_, int cls_) : c(c_), cls(cls_), k(t_begin(c_)) {}
    ~TScope() { t_end(c, k, cls); }
};

// ---- launchers ---------------------------------------------------------
template <typename T, int N> int launch_fft_x(fgpe_ctx* c, const void* in, void* out, int dir)
{
    constexpr int P = px_of<T, N>();
    const long long rows = c->ntot / N;
    if (rows % P) return fail(FGPE_EUNSUP, "x rows not a multiple of the pencils per block");
    dim3 grid((unsigned)(rows / P)), block(P * Plan<N>::T);
    using T2 = c2_t<T>;
    const T2* tw = (const T2*)tw_of<T>(c, 0);
    TScope ts(c, FGPE_KCLASS_FFTX);
    if (dir < 0) hipLaunchKernelGGL((k_fft_x<T, N, -1>), grid, block, 0, c->stream, (const T2*)in, (T2*)out, tw);
    else hipLaunchKernelGGL((k_fft_x<T, N, +1>), grid, block, 0, c->stream, (const T2*)in, (T2*)out, tw);
    LAUNCHCHK("k_fft_x");
    return FGPE_OK;
}

StridedArgs strided_args(fgpe_ctx* c, int axis, const double* scale, int ncomp)
{
    StridedArgs a;
    a.nx = c->n[0];
    a.chan_stride = c->ntot;
    a.scale = scale;
    a.ncomp = ncomp;
    if (axis == 1) {
        a.stride = c->n[0];
        a.outer_stride = (long long)c->n[0] * c->n[1];
    } else {
        a.stride = (long long)c->n[0] * c->n[1];
        a.outer_stride = c->n[0];
    }
    return a;
}

long long n_outer(fgpe_ctx* c, int axis) { return axis == 1 ? c->n[2] : c->n[1]; }

template <typename T, int N> int launch_strided(fgpe_ctx* c, void* data, int axis, int dir, int nch,
                                                const double* scale, int ncomp)
{
    constexpr int P = ps_of<T, N>();
    if (c->n[0] % P) return fail(FGPE_EUNSUP, "nx not a multiple of the strided tile width");
    StridedArgs a = strided_args(c, axis, scale, ncomp);
    dim3 grid((unsigned)((c->n[0] / P) * n_outer(c, axis)), (unsigned)nch), block(P * Plan<N>::T);
    using T2 = c2_t<T>;
    const T2* tw = (const T2*)tw_of<T>(c, axis);
    TScope ts(c, FGPE_KCLASS_STRIDED);
    if (dir < 0) hipLaunchKernelGGL((k_strided<T, N, -1>), grid, block, 0, c->stream, (T2*)data, tw, a);
    else hipLaunchKernelGGL((k_strided<T, N, +1>), grid, block, 0, c->stream, (T2*)data, tw, a);
    LAUNCHCHK("k_strided");
    return FGPE_OK;
}

template <typename T, int N> int launch_last(fgpe_ctx* c, void* data, int axis, int nch, int ncomp,
                                             const double* scale, bool filter)
{
    constexpr int P = ps_of<T, N>();
    if (c->n[0] % P) return fail(FGPE_EUNSUP, "nx not a multiple of the strided tile width");
    LastArgs a;
    a.s = strided_args(c, axis, scale, ncomp);
    const int oaxis = axis == 1 ? 2 : 1;
    a.kA = c->k[axis];
    a.kX = c->k[0];
    a.kO = c->k[oaxis];
    const double* d = c->hp.dip;
    a.dA = d[axis];
    a.dX = d[0];
    a.dO = d[oaxis];
    a.lap_scale = 19.739208802 / (double)c->ntot;
    a.inv_ntot = 1.0 / (double)c->ntot;
    a.c_dd = filter ? 12.0 * kPi * c->hp.a_dd : 0.0;
    a.wtab = filter ? c->wtab : nullptr;
    dim3 grid((unsigned)((c->n[0] / P) * n_outer(c, axis)), (unsigned)nch), block(P * Plan<N>::T);
    using T2 = c2_t<T>;
    TScope ts(c, FGPE_KCLASS_LAST);
    hipLaunchKernelGGL((k_last<T, N>), grid, block, 0, c->stream, (T2*)data, (const T2*)tw_of<T>(c, axis), a);
    LAUNCHCHK("k_last");
    return FGPE_OK;
}

template <typename T, int N, bool DENS, int MODE> int launch_xstage_t(fgpe_ctx* c, const XArgs<T>& a)
{
    constexpr int P = px_of<T, N>();
    const long long rows = c->ntot / N;
    if (rows % P) return fail(FGPE_EUNSUP, "x rows not a multiple of the pencils per block");
    dim3 grid((unsigned)(rows / P)), block(P * Plan<N>::T);
    TScope ts(c, FGPE_KCLASS_XSTAGE);
    hipLaunchKernelGGL((k_xstage<T, N, DENS, MODE>), grid, block, 0, c->stream, a,
                       (const c2_t<T>*)tw_of<T>(c, 0));
    LAUNCHCHK("k_xstage");
    return FGPE_OK;
}

template <typename T, int N> int launch_xstage(fgpe_ctx* c, const XArgs<T>& a, bool dens, int mode)
{
    if (dens) {
        if (mode == MODE_PROLOGUE) return launch_xstage_t<T, N, true, MODE_PROLOGUE>(c, a);
        if (mode == MODE_RK) return launch_xstage_t<T, N, true, MODE_RK>(c, a);
        return launch_xstage_t<T, N, true, MODE_APPLY>(c, a);
    }
    if (mode == MODE_PROLOGUE) return launch_xstage_t<T, N, false, MODE_PROLOGUE>(c, a);
    if (mode == MODE_RK) return launch_xstage_t<T, N, false, MODE_RK>(c, a);
    return launch_xstage_t<T, N, false, MODE_APPLY>(c, a);
}

template <typename T> long long xblocks(fgpe_ctx* c)
{
    switch (c->n[0]) {
#define CASE(N_) \
    case N_: return c->ntot / N_ / px_of<T, N_>();
        FGPE_FOR_EACH_N(CASE)
#undef CASE
    }
    return 0;
}

// ---- runtime dispatch over the axis length ---------------------------------
template <typename T> int fft_x(fgpe_ctx* c, const void* in, void* out, int dir)
{
    switch (c->n[0]) {
#define CASE(N_) \
    case N_: return launch_fft_x<T, N_>(c, in, out, dir);
        FGPE_FOR_EACH_N(CASE)
#undef CASE
    }
    return fail(FGPE_EUNSUP, "unsupported nx");
}
template <typename T> int strided(fgpe_ctx* c, void* data, int axis, int dir, int nch, const double* scale, int ncomp)
{
    switch (c->n[axis]) {
#define CASE(N_) \
    case N_: return launch_strided<T, N_>(c, data, axis, dir, nch, scale, ncomp);
        FGPE_FOR_EACH_N(CASE)
#undef CASE
    }
    return fail(FGPE_EUNSUP, "unsupported strided axis length");
}
template <typename T> int last(fgpe_ctx* c, void* data, int axis, int nch, int ncomp, const double* scale, bool filter)
{
    switch (c->n[axis]) {
#define CASE(N_) \
    case N_: return launch_last<T, N_>(c, data, axis, nch, ncomp, scale, filter);
        FGPE_FOR_EACH_N(CASE)
#undef CASE
    }
    return fail(FGPE_EUNSUP, "unsupported last axis length");
}
template <typename T> int xstage(fgpe_ctx* c, const XArgs<T>& a, bool dens, int mode)
{
    switch (c->n[0]) {
#define CASE(N_) \
    case N_: return launch_xstage<T, N_>(c, a, dens, mode);
        FGPE_FOR_EACH_N(CASE)
#undef CASE
    }
    return fail(FGPE_EUNSUP, "unsupported nx");
}

int last_axis(fgpe_ctx* c) { return c->n[2] > 1 ? 2 : 1; }

#define RET(expr)                 \
    do {                          \
        int r_ = (expr);          \
        if (r_ != FGPE_OK) return r_; \
    } while (0)

// Plain 3D transform on a device array (standalone passes)
template <typename T> int fft3d(fgpe_ctx* c, const void* in, void* out, int dir)
{
    RET(fft_x<T>(c, in, out, dir));
    if (c->n[1] > 1) RET(strided<T>(c, out, 1, dir, 1, nullptr, 1));
    if (c->n[2] > 1) RET(strided<T>(c, out, 2, dir, 1, nullptr, 1));
    return FGPE_OK;
}

// y-forward / last axis / y-inverse on the spectral channels.  `scale` is
// applied by the first forward pass after a prologue or stage 4.
template <typename T> int spectral_passes(fgpe_ctx* c, void* spec, int nch, int ncomp, const double* scale,
                                          bool filter)
{
    const int la = last_axis(c);
    if (la == 2) {
        RET(strided<T>(c, spec, 1, -1, nch, scale, ncomp));
        RET(last<T>(c, spec, 2, nch, ncomp, nullptr, filter));
        RET(strided<T>(c, spec, 1, +1, nch, nullptr, ncomp));
    } else {
        RET(last<T>(c, spec, 1, nch, ncomp, scale, filter));
    }
    return FGPE_OK;
}

bool has_filter(const fgpe_hparams& h) { return h.a_dd != 0.0 || h.sc_U != 0.0; }

double a_at(const fgpe_hparams& h, double t)
{
    if (!h.quench_on) return h.a[0];
    // d_quench_as, hamiltonians/BEC_oneComponent.cu:167-178
    const double a0 = h.quench[0], a1 = h.quench[1], a2 = h.quench[2];
    const double t1 = h.quench[3], t2 = h.quench[4], t3 = h.quench[5];
    if (0 < t && t <= t1) return a0 + (a1 - a0) / (t1) * t;
    if (t1 < t && t <= t2) return a1;
    if (t2 < t && t <= t3) return a1 + (a2 - a1) / (t3 - t2) * (t - t2);
    return a2;
}

template <typename T> XArgs<T> base_xargs(fgpe_ctx* c)
{
    XArgs<T> a{};
    a.spec = (c2_t<T>*)c->spec;
    a.ntot = c->ntot;
    a.P0 = c->P0[0];
    a.PS = (c2_t<T>*)c->PS[0];
    a.K = (c2_t<T>*)c->K[0];
    a.Vr = (const T*)c->Vr[0];
    a.Vi = c->vcomplex[0] ? (const T*)c->Vi[0] : nullptr;
    a.partials = c->partials;
    return a;
}

int ensure_ready(fgpe_ctx* c)
{
    if (!c->kset) return fail(FGPE_ESTATE, "fgpe_set_kspace has not been called");
    if (!c->hset) return fail(FGPE_ESTATE, "fgpe_set_hamiltonian has not been called");
    if (!c->Vr[0]) return fail(FGPE_ESTATE, "no external potential uploaded");
    if (c->hp.ncomp != 1) return fail(FGPE_EUNSUP, "two-component stepping is not available in this build");
    if (c->n[1] == 1) return fail(FGPE_EUNSUP, "1D grids are not supported");
    return FGPE_OK;
}

template <typename T> int prologue(fgpe_ctx* c)
{
    XArgs<T> a = base_xargs<T>(c);
    RET(xstage<T>(c, a, c->nch > c->hp.ncomp, MODE_PROLOGUE));
    c->spec_valid = true;
    return FGPE_OK;
}

// RK4 steps; imaginary: c = -dt and renormalisation to npart after each step
template <typename T> int rk4_run(fgpe_ctx* c, double dt, double t0, int nsteps, bool imaginary, double npart,
                                  double* hist_dev)
{
    RET(ensure_ready(c));
    const bool dens = c->nch > c->hp.ncomp;
    const bool filter = has_filter(c->hp);
    if (!c->spec_valid) RET(prologue<T>(c));
    const double* scale = imaginary ? c->scal : nullptr;
    const long long nb = xblocks<T>(c);
    for (int it = 0; it < nsteps; ++it) {
        for (int s = 1; s <= 4; ++s) {
            const double frac = s == 1 ? 0.0 : (s == 4 ? 1.0 : 0.5);
            // PARS[T] at each stage (src/oneComponentGPSolver.cu:175,184)
            const double t = s == 1 ? t0 + it * dt : t0 + (it + frac) * dt;
            RET(spectral_passes<T>(c, c->spec, c->nch, c->hp.ncomp, s == 1 ? scale : nullptr, filter));
            XArgs<T> a = base_xargs<T>(c);
            const double av = a_at(c->hp, t);
            a.g4 = 4 * kPi * av;
            a.g2 = 2 * kPi * av;
            const double cf = (s <= 2) ? -0.5 * dt : -1.0 * dt;
            if (imaginary) { a.cr = cf; a.ci = 0.0; }
            else { a.cr = 0.0; a.ci = cf; }
            a.stage = s;
            a.scale = scale;
            a.want_norm = (imaginary && s == 4) ? 1 : 0;
            RET(xstage<T>(c, a, dens, MODE_RK));
            if (imaginary && s == 4) {
                hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->partials, (int)nb, 1,
                                   (int)FIN_RENORM, c->dv, npart, hist_dev ? hist_dev + 3 * it : c->scal + 8,
                                   c->scal);
                LAUNCHCHK("k_finalize");
            }
        }
    }
    return FGPE_OK;
}

template <typename T> int apply_h(fgpe_ctx* c, double t, void* hmu, void* hen, double* obs_dev)
{
    RET(ensure_ready(c));
    RET(prologue<T>(c));
    RET(spectral_passes<T>(c, c->spec, c->nch, c->hp.ncomp, c->scal, has_filter(c->hp)));
    XArgs<T> a = base_xargs<T>(c);
    const double av = a_at(c->hp, t);
    a.g4 = 4 * kPi * av;
    a.g2 = 2 * kPi * av;
    a.scale = c->scal;
    a.hmu = (double2*)hmu;
    a.hen = (double2*)hen;
    RET(xstage<T>(c, a, c->nch > c->hp.ncomp, MODE_APPLY));
    c->spec_valid = false;
    if (obs_dev) {
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->partials, (int)xblocks<T>(c), 3,
                           (int)FIN_OBS, c->dv, 1.0, obs_dev, c->scal);
        LAUNCHCHK("k_finalize");
    }
    return FGPE_OK;
}

int apply_pending_scale(fgpe_ctx* c)
{
    const int th = 256;
    for (int comp = 0; comp < c->hp.ncomp; ++comp) {
        hipLaunchKernelGGL(k_apply_scale, dim3((unsigned)((c->ntot + th - 1) / th)), dim3(th), 0, c->stream,
                           c->P0[comp], c->ntot, c->scal);
        LAUNCHCHK("k_apply_scale");
    }
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c->stream, c->scal, 1.0);
    LAUNCHCHK("k_set_scalar");
    c->spec_valid = false;
    return FGPE_OK;
}

}  // namespace

// =========================================================================
extern "C" {

const char* fgpe_last_error(void) { return g_err.c_str(); }

int fgpe_ctx_create(int nx, int ny, int nz, int precision, fgpe_ctx** out)
{
    if (!out) return fail(FGPE_EINVAL, "out is NULL");
    *out = nullptr;
    if (precision != FGPE_PREC_F64 && precision != FGPE_PREC_MIXED)
        return fail(FGPE_EINVAL, "precision must be FGPE_PREC_F64 or FGPE_PREC_MIXED");
    if (nx <= 1 || ny < 1 || nz < 1) return fail(FGPE_EINVAL, "grid extents must be positive (nx > 1)");
    if (!supported_n(nx) || (ny > 1 && !supported_n(ny)) || (nz > 1 && !supported_n(nz)))
        return fail(FGPE_EUNSUP, "grid extents must be 1 or one of 32, 64, 96, 128, 192, 256, 512");
    fgpe_ctx* c = new fgpe_ctx();
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
    if (r == FGPE_OK) r = dalloc(c, &c->K[0], 2 * c->tsz() * N);
    if (r == FGPE_OK) r = dalloc(c, (void**)&c->scal, sizeof(double) * 16);
    if (r == FGPE_OK) {
        c->partials_n = 3 * (size_t)(c->ntot / 2 + 1024);
        r = dalloc(c, (void**)&c->partials, sizeof(double) * c->partials_n);
    }
    if (r == FGPE_OK) {
        double one[16] = {1.0};
        if (hipMemcpy(c->scal, one, sizeof(one), hipMemcpyHostToDevice) != hipSuccess)
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
    (void)hipDeviceSynchronize();
    void* ptrs[] = {c->k[0], c->k[1], c->k[2], c->tw64[0], c->tw64[1], c->tw64[2], c->tw32[0], c->tw32[1],
                    c->tw32[2], c->P0[0], c->P0[1], c->PS[0], c->PS[1], c->K[0], c->K[1], c->Vr[0], c->Vr[1],
                    c->Vi[0], c->Vi[1], c->spec, c->wtab, c->partials, c->scal, c->hist};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
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
    HIPCHK(hipStreamSynchronize(c->stream));
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
    const double* src[3] = {kx, ky, kz};
    for (int a = 0; a < 3; ++a)
        HIPCHK(hipMemcpy(c->k[a], src[a], sizeof(double) * c->n[a], hipMemcpyHostToDevice));
    c->dv = dv;
    c->kset = true;
    return FGPE_OK;
}

int fgpe_set_hamiltonian(fgpe_ctx* c, const fgpe_hparams* hp)
{
    if (!c || !hp) return fail(FGPE_EINVAL, "NULL argument");
    if (hp->ncomp != 1 && hp->ncomp != 2) return fail(FGPE_EINVAL, "ncomp must be 1 or 2");
    if (hp->sc_U != 0.0 && (hp->sc_dim < 1 || hp->sc_dim > 3)) return fail(FGPE_EINVAL, "sc_dim must be 1, 2 or 3");
    if (hp->sc_U != 0.0 && !c->kset) return fail(FGPE_ESTATE, "set the k space before a soft-core term");
    c->hp = *hp;
    const int nch = hp->ncomp + (has_filter(*hp) ? 1 : 0);
    const size_t N = (size_t)c->ntot;
    if (nch != c->nch) {
        if (c->spec) { (void)hipFree(c->spec); c->bytes -= 2 * c->tsz() * N * c->nch; c->spec = nullptr; }
        RET(dalloc(c, &c->spec, 2 * c->tsz() * N * nch));
        c->nch = nch;
    }
    if (hp->ncomp == 2) {
        if (!c->P0[1]) RET(dalloc(c, (void**)&c->P0[1], sizeof(double2) * N));
        if (!c->PS[1]) RET(dalloc(c, &c->PS[1], 2 * c->tsz() * N));
        if (!c->K[1]) RET(dalloc(c, &c->K[1], 2 * c->tsz() * N));
    }
    if (hp->sc_U != 0.0) {
        if (!c->wtab) RET(dalloc(c, &c->wtab, c->tsz() * N));
        const int th = 256;
        dim3 grid((unsigned)((N + th - 1) / th));
        if (c->mixed())
            hipLaunchKernelGGL(k_softcore_table<float>, grid, dim3(th), 0, c->stream, (float*)c->wtab, c->k[0],
                               c->k[1], c->k[2], c->n[0], c->n[1], c->ntot, hp->sc_dim, hp->sc_R, hp->sc_U);
        else
            hipLaunchKernelGGL(k_softcore_table<double>, grid, dim3(th), 0, c->stream, (double*)c->wtab, c->k[0],
                               c->k[1], c->k[2], c->n[0], c->n[1], c->ntot, hp->sc_dim, hp->sc_R, hp->sc_U);
        LAUNCHCHK("k_softcore_table");
    } else if (c->wtab) {
        (void)hipFree(c->wtab);
        c->bytes -= c->tsz() * N;
        c->wtab = nullptr;
    }
    c->hset = true;
    c->spec_valid = false;
    return FGPE_OK;
}

int fgpe_upload_psi(fgpe_ctx* c, int comp, const double* host)
{
    if (!c || !host) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1) return fail(FGPE_EINVAL, "comp must be 0 or 1");
    if (!c->P0[comp]) RET(dalloc(c, (void**)&c->P0[comp], sizeof(double2) * c->ntot));
    HIPCHK(hipMemcpyAsync(c->P0[comp], host, sizeof(double2) * c->ntot, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c->stream, c->scal, 1.0);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->spec_valid = false;
    return FGPE_OK;
}

int fgpe_download_psi(fgpe_ctx* c, int comp, double* host)
{
    if (!c || !host) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1 || !c->P0[comp]) return fail(FGPE_EINVAL, "no such component");
    HIPCHK(hipMemcpyAsync(host, c->P0[comp], sizeof(double2) * c->ntot, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FGPE_OK;
}

int fgpe_upload_vext(fgpe_ctx* c, int comp, const double* host)
{
    if (!c || !host) return fail(FGPE_EINVAL, "NULL argument");
    if (comp < 0 || comp > 1) return fail(FGPE_EINVAL, "comp must be 0 or 1");
    const size_t N = (size_t)c->ntot;
    bool cplx = false;
    for (size_t i = 0; i < N && !cplx; ++i) cplx = host[2 * i + 1] != 0.0;
    if (!c->Vr[comp]) RET(dalloc(c, &c->Vr[comp], c->tsz() * N));
    if (cplx && !c->Vi[comp]) RET(dalloc(c, &c->Vi[comp], c->tsz() * N));
    if (c->mixed()) {
        std::vector<float> re(N), im(cplx ? N : 0);
        for (size_t i = 0; i < N; ++i) {
            re[i] = (float)host[2 * i];
            if (cplx) im[i] = (float)host[2 * i + 1];
        }
        HIPCHK(hipMemcpy(c->Vr[comp], re.data(), 4 * N, hipMemcpyHostToDevice));
        if (cplx) HIPCHK(hipMemcpy(c->Vi[comp], im.data(), 4 * N, hipMemcpyHostToDevice));
    } else {
        std::vector<double> re(N), im(cplx ? N : 0);
        for (size_t i = 0; i < N; ++i) {
            re[i] = host[2 * i];
            if (cplx) im[i] = host[2 * i + 1];
        }
        HIPCHK(hipMemcpy(c->Vr[comp], re.data(), 8 * N, hipMemcpyHostToDevice));
        if (cplx) HIPCHK(hipMemcpy(c->Vi[comp], im.data(), 8 * N, hipMemcpyHostToDevice));
    }
    c->vcomplex[comp] = cplx;
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
    if (c->n[1] == 1) return fail(FGPE_EUNSUP, "1D grids are not supported");
    RET(fft_x<double>(c, psi, out, -1));
    const int la = last_axis(c);
    if (la == 2) RET(strided<double>(c, out, 1, -1, 1, nullptr, 1));
    RET(last<double>(c, out, la, 1, 1, nullptr, false));
    if (la == 2) RET(strided<double>(c, out, 1, +1, 1, nullptr, 1));
    RET(fft_x<double>(c, out, out, +1));
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
    double* obs = c->scal + 4;
    RET(c->mixed() ? apply_h<float>(c, t, nullptr, nullptr, obs) : apply_h<double>(c, t, nullptr, nullptr, obs));
    HIPCHK(hipMemcpyAsync(out, obs, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FGPE_OK;
}

int fgpe_rk4_steps(fgpe_ctx* c, double dt, double t0, int nsteps)
{
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (nsteps < 0) return fail(FGPE_EINVAL, "nsteps < 0");
    return c->mixed() ? rk4_run<float>(c, dt, t0, nsteps, false, 1.0, nullptr)
                      : rk4_run<double>(c, dt, t0, nsteps, false, 1.0, nullptr);
}

int fgpe_imag_steps(fgpe_ctx* c, int scheme, double dt, double beta, double npart, double eps, int nsteps,
                    double* host_hist, int* steps_done)
{
    (void)beta;
    (void)eps;
    if (!c) return fail(FGPE_EINVAL, "ctx is NULL");
    if (nsteps < 0) return fail(FGPE_EINVAL, "nsteps < 0");
    if (scheme != FGPE_IMAG_RK4_RENORM) return fail(FGPE_EUNSUP, "only FGPE_IMAG_RK4_RENORM is available");
    double* hd = nullptr;
    if (host_hist && nsteps > 0) {
        if ((size_t)nsteps * 3 > c->hist_n) {
            if (c->hist) { (void)hipFree(c->hist); c->bytes -= sizeof(double) * c->hist_n; c->hist = nullptr; }
            RET(dalloc(c, (void**)&c->hist, sizeof(double) * 3 * nsteps));
            c->hist_n = 3 * (size_t)nsteps;
        }
        hd = c->hist;
    }
    RET(c->mixed() ? rk4_run<float>(c, dt, 0.0, nsteps, true, npart, hd)
                   : rk4_run<double>(c, dt, 0.0, nsteps, true, npart, hd));
    RET(apply_pending_scale(c));
    if (hd) HIPCHK(hipMemcpyAsync(host_hist, hd, sizeof(double) * 3 * nsteps, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (steps_done) *steps_done = nsteps;
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
        for (auto& e : c->evpool) HIPCHK(hipEventCreate(&e));
    }
    return FGPE_OK;
}

int fgpe_timing_read(fgpe_ctx* c, double* out)
{
    if (!c || !out) return fail(FGPE_EINVAL, "NULL argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 2 * FGPE_NKCLASS; ++k) out[k] = 0.0;
    for (auto& r : c->evrec) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->evpool[r.second], c->evpool[r.second + 1]));
        out[2 * r.first] += ms;
        out[2 * r.first + 1] += 1.0;
    }
    c->evused = 0;
    c->evrec.clear();
    return FGPE_OK;
}

}  // extern "C"
