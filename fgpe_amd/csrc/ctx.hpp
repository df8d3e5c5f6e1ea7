// ctx.hpp -- the context behind fgpe_ctx* (device buffers, twiddle tables,
// stream, launch timing) and the per-precision launcher interface shared by
// the translation units of libfgpe.so.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/fgpe_c.h"
#include "gpe_kernels.hpp"

namespace fgpe {

// Capacity of the MI355X memory-side cache (Infinity Cache).
constexpr size_t kMallBytes = 256ull << 20;

int fail(int code, const std::string& msg);

#define FGPE_HIPCHK(expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::fgpe::fail(FGPE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define FGPE_LAUNCHCHK(what)                                                                  \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return ::fgpe::fail(FGPE_EHIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define FGPE_RET(expr)                    \
    do {                                  \
        int r_ = (expr);                  \
        if (r_ != FGPE_OK) return r_;     \
    } while (0)

// Pencil lengths with a compiled plan.  A build for the user-term tests
// (FGPE_TEST_SIZES) compiles the lengths those tests use only.
#ifdef FGPE_TEST_SIZES
#define FGPE_FOR_EACH_N(X) X(32) X(64)
#else
#define FGPE_FOR_EACH_N(X) X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(1024)
#endif

// Lengths whose x kernels have a distinct third plan (Plan<N, 2>).
#define FGPE_FOR_EACH_N_X2(X) X(128) X(256)

// Whether the x-stage class of an nx-long grid has a distinct variant 2.
constexpr bool x_variant2(int nx) { return nx == 128 || nx == 256; }

// Scalar slots of fgpe_ctx::scal (doubles) and ::flags (ints) on the device.
enum {
    SC_SCALE = 0,   // [2] pending renormalisation per component
    SC_MU = 2,      // [2] heavy-ball mu of the previous iteration
    SC_EPREV = 4,   // [2] heavy-ball E of the previous iteration
    SC_OBS = 6,     // [6] observables (norm, mu, E) per component
    SC_TMP = 12,    // [3] scratch output
    SC_COUNT = 16
};
enum { FL_STOP = 0, FL_DONE = 1, FL_COUNT = 4 };

}  // namespace fgpe

struct fgpe_ctx {
    int n[3] = {1, 1, 1};
    long long ntot = 0;
    int dev = 0;   // HIP device the context was created on
    int prec = FGPE_PREC_F64;
    hipStream_t own = nullptr, stream = nullptr;
    size_t bytes = 0;

    double* k[3] = {nullptr, nullptr, nullptr};
    float2* k32[3] = {nullptr, nullptr, nullptr};  // the same axes in FP32 as (k[i], k[(n - i) % n])
    double2* k64[3] = {nullptr, nullptr, nullptr}; // and in FP64 as (k[i], k[(n - i) % n])
    double dv = 0.0;
    bool kset = false;

    fgpe_hparams hp{};
    bool hset = false;
    int nch = 0;

    double2* P0[2] = {nullptr, nullptr};
    void* PS[2] = {nullptr, nullptr};
    void* PSb[2] = {nullptr, nullptr};   // mixed form: RK stage-2 state (stage 3 reads stage 1's in PS)
    void* K[2] = {nullptr, nullptr};
    void* Vr[2] = {nullptr, nullptr};
    void* Vi[2] = {nullptr, nullptr};
    bool vcomplex[2] = {false, false};
    void* spec = nullptr;
    void* wtab = nullptr;
    double* partials = nullptr;
    size_t partials_n = 0;
    double* scal = nullptr;
    int* flags = nullptr;
    double* hist = nullptr;
    size_t hist_n = 0;
    bool spec_valid = false;
    bool hb_active = false;   // a heavy-ball run can be continued
    // imaginary-time iterations run since the state or the Hamiltonian was
    // last set (or a heavy-ball run started afresh): the time of the next
    // iteration is imag_iter * dt, as the reference's PARS[T] = iteration * dt
    long long imag_iter = 0;
    double2* aux = nullptr;   // FP64 scratch of the diagnostics
    // adjoint: checkpoints, segment states, stage states, VJP buffers
    std::vector<double2*> ckpt;
    std::vector<double2*> segbuf;
    std::vector<void*> ystep;   // per step of a segment: its stage states y1..y3 (3 x adj_nc bulk states)
    void* spec3 = nullptr;
    double2* lam[3] = {nullptr, nullptr, nullptr};   // LIN, LAM, GB
    double* pacc = nullptr;
    int adj_nc = 0;   // components the adjoint buffers hold (0: none allocated)
    // the k-space kernels the next spectral multiplies apply (bit 0: dipolar,
    // bit 1: the soft-core table); the adjoint takes the dipolar one alone
    // for one channel when both are on
    int kmask = 3;
    int adj_every = 0, adj_nsteps = 0;
    double adj_dt = 0.0, adj_t0 = 0.0;
    bool adj_ready = false;
    double* xbuf = nullptr;   // x cell centres of fgpe_calc_cos
    double* vstage[2] = {nullptr, nullptr};   // FP64 trap being assembled (fgpe_vext_*)
    double* axbuf = nullptr;                   // per-axis factor tables of the initialisers

    void* tw64[3] = {nullptr, nullptr, nullptr};
    void* tw32[3] = {nullptr, nullptr, nullptr};
    int var[FGPE_NKCLASS][3] = {};   // plan variant per kernel class and axis
    int nslab = 1;                   // column slabs of the y / last / y pass sequence
    int order = 0;                   // bit k set: pass k (y fwd, last, y inv, x) runs its blocks reversed
    // 3D grids: the k-space multiply rides on the y pass (z pencils run the
    // strided forward and inverse passes) instead of the z pass
    bool ylast = false;

    // launch timing (fgpe_timing_enable)
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    std::vector<std::pair<int, size_t>> evrec;

    bool mixed() const { return prec == FGPE_PREC_MIXED; }
    size_t tsz() const { return mixed() ? 4 : 8; }
    // Bytes one time step streams through: per component the FP64 state, the
    // stage state and accumulator (bulk complex) and the trap (bulk real),
    // plus the spectral channels.
    size_t working_set() const
    {
        const size_t comp = 16 + 2 * (2 * tsz()) + tsz() + (mixed() ? 2 * tsz() : 0);
        return (size_t)ntot * (hp.ncomp * comp + (size_t)nch * 2 * tsz());
    }
};

namespace fgpe {

// ---- launch timing: HIP events around each launch when enabled -------------
long t_begin(fgpe_ctx* c);
void t_end(fgpe_ctx* c, long k, int cls);
// Device allocation charged to the context (fgpe_ctx_device_bytes).
int dalloc(fgpe_ctx* c, void** p, size_t bytes);
void dfree(fgpe_ctx* c, void** p, size_t bytes);
struct TScope {
    fgpe_ctx* c;
    int cls;
    long k;
    TScope(fgpe_ctx* c_, int cls_) : c(c_), cls(cls_), k(t_begin(c_)) {}
    ~TScope() { t_end(c, k, cls); }
};

// Row pitch (elements) of a half-spectrum real channel: kx = 0..nx/2 stored,
// rows padded to nx/2 + 16 elements (128-byte row starts for nx >= 32 in
// both precisions), but never wider than nx: the channel lives in a slot of
// ntot = nx ny nz elements of the spectral buffer (at nx = 16 the padded
// pitch, 24, would overrun it).
inline long long dens_pitch(const fgpe_ctx* c)
{
    const long long padded = c->n[0] / 2 + 16;
    return padded < c->n[0] ? padded : c->n[0];
}

// Slab s of nslab of the columns of a strided-pass argument set: complex
// columns in equal parts (multiples of the tile width P), half-layout columns
// in near-equal parts.  nslab = 1 leaves the arguments unchanged.
inline StridedArgs slab_of(StridedArgs a, int s, int nslab)
{
    if (nslab <= 1) return a;
    const int w = a.nx / nslab;
    a.x0 = s * w;
    a.ncols = w;
    if (a.hcols) {
        const int hw = (a.hcols + nslab - 1) / nslab;
        a.hx0 = s * hw;
        a.hcols = s * hw >= a.hcols ? 0 : (a.hcols - s * hw < hw ? a.hcols - s * hw : hw);
    }
    return a;
}

inline long long n_outer(const fgpe_ctx* c, int axis) { return axis == 1 ? c->n[2] : c->n[1]; }

// Arguments of the strided passes and of the k-space multiply for one context.
inline StridedArgs strided_args(const fgpe_ctx* c, void* data, int axis, const double* scale, int nch, int ncomp,
                                bool half, const int* stop)
{
    StridedArgs a{};
    a.data = data;
    a.nx = c->n[0];
    a.ncols = c->n[0];
    a.chan_stride = c->ntot;
    a.scale = scale;
    a.ncomp = ncomp;
    a.nchan = nch;
    a.stop = stop;
    a.nouter = (int)n_outer(c, axis);
    const long long n0 = c->n[0], n1 = c->n[1];
    if (axis == 1) {
        a.stride = n0;
        a.outer_stride = n0 * n1;
    } else {
        a.stride = n0 * n1;
        a.outer_stride = n0;
    }
    if (half && nch > ncomp) {
        const long long hp = dens_pitch(c);
        a.hcols = c->n[0] / 2 + 1;
        a.hbase = (long long)ncomp * c->ntot;
        a.hchan = hp * n1 * c->n[2];
        if (axis == 1) {
            a.hstride = hp;
            a.houter = hp * n1;
        } else {
            a.hstride = hp * n1;
            a.houter = hp;
        }
    }
    return a;
}

// axis: the last axis; vpx: include the moving-frame term of the
// Hamiltonian's full form (fgpe_hparams::vpx)
inline LastArgs last_args(const fgpe_ctx* c, void* data, int axis, int nch, int ncomp, const double* scale,
                          bool filter, const int* stop, bool half, bool vpx = false)
{
    LastArgs a;
    a.s = strided_args(c, data, axis, scale, nch, ncomp, half, stop);
    a.s.gather = 1;
    const int oaxis = axis == 1 ? 2 : 1;
    a.kA32 = c->k32[axis];
    a.kX32 = c->k32[0];
    a.kO32 = c->k32[oaxis];
    a.kA64 = c->k64[axis];
    a.kX64 = c->k64[0];
    a.kO64 = c->k64[oaxis];
    const double* d = c->hp.dip;
    a.dA = d[axis];
    a.dX = d[0];
    a.dO = d[oaxis];
    a.lap_scale = 19.739208802 / (double)c->ntot;
    a.inv_ntot = 1.0 / (double)c->ntot;
    a.c_dd = (filter && (c->kmask & 1)) ? 12.0 * kPi * c->hp.a_dd : 0.0;
    a.wtab = (filter && (c->kmask & 2)) ? c->wtab : nullptr;
    for (int i = 0; i < 2; ++i) a.vpx[i] = vpx ? c->hp.vpx[i] / (double)c->ntot : 0.0;
    return a;
}

// ---- per-precision launchers (launch_f64.hip / launch_f32.hip) --------------
// axis: 1 = y, 2 = z.  nch channels at ctx->spec (or `data`): ncomp complex
// channels of ntot points, then real-field channels, in the half-spectrum
// layout when `half` is set (rows of dens_pitch elements), else like the
// complex ones.  tab / ninst: a device table of ninst per-instance argument
// structs for a batched launch over same-shape contexts (null: this context).
template <typename T> struct Launch {
    static long long xblocks(fgpe_ctx* c);
    static int fft_x(fgpe_ctx* c, const void* in, void* out, int dir);
    static int strided(fgpe_ctx* c, void* data, int axis, int dir, int nch, int ncomp, const double* scale,
                       const int* stop, bool half, const void* tab = nullptr, int ninst = 1, int slab = 0,
                       int nslab = 1);
    static int last(fgpe_ctx* c, void* data, int axis, int nch, int ncomp, const double* scale, bool filter,
                    const int* stop, bool half, const void* tab = nullptr, int ninst = 1, int slab = 0,
                    int nslab = 1, bool vpx = false);
    static int xstage(fgpe_ctx* c, const XArgs<T>& a, int nc, bool dens, int mode, const void* tab = nullptr,
                      int ninst = 1);
    // adjoint VJP prologue / epilogue; y64: the state is FP64 (else T)
    static int adj(fgpe_ctx* c, const AdjArgs& a, bool y64, bool dens, bool epilogue);
};

}  // namespace fgpe
