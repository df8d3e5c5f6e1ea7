// gpe_kernels.hpp -- HIP kernels of the fused GPE time step for gfx950.
//
// A Hamiltonian application on a 3D grid is five kernels instead of the
// reference's cuFFT-forward / Laplace / cuFFT-inverse / pointwise sequence
// (src/baseHamiltonian.cu:86-96, hamiltonians/BEC_oneComponent.cu:92-111):
//
//   k_xstage   inverse x-FFT of every spectral channel, the pointwise H,
//              the RK4 stage combination, then the forward x-FFT of the next
//              stage's state (and of its density)             [x pencils]
//   k_strided  forward y-FFT                                   [y pencils]
//   k_last     forward z-FFT, k-space multiply (kinetic 19.739208802/N k^2,
//              dipolar / soft-core Vt(k)/N), inverse z-FFT     [z pencils]
//   k_strided  inverse y-FFT                                   [y pencils]
//
// Spectral channels live channel-major in one buffer: channel c < ncomp is
// psi_c, channel ncomp (if present) is the total density.  All arrays use the
// reference layout ID(x,y,z) = z*NX*NY + y*NX + x.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fft_core.hpp"
#include "user_term.hpp"

namespace fgpe {

constexpr double kPi = 3.141592653589793;  // hpp/header.hpp:28

// pencils per block of the x kernels and tile width of the strided kernels
// Variant 0: 128-thread x blocks, 128-byte row segments in the strided
// passes; variant 1 (with its plan of more threads per pencil): 256-thread x
// blocks and 256-byte row segments; variant 2 (x kernels only): 128-thread
// x blocks (256-thread blocks of four one-wave pencils at nx = 256, which
// halves the blocks that stage a twiddle table: x stage 337 -> 334 us FP64,
// 179 -> 174 us FP32 at 256^3).  FP64 256-point pencils in variant 0 (16 elements per thread,
// one exchange per transform) take 16-column tiles, 256-byte rows, as in
// variant 1: the y and last-axis passes fit three such tiles in a CU by
// splitting their exchanges (strided_split, last_split below).
template <typename T, int N, int V = 0> __host__ __device__ constexpr int px_of()
{
    if (V == 2 && N == 256) return 4;
    constexpr int th = V == 1 ? 256 : 128;
    return Plan<N, V>::T >= th ? 1 : th / Plan<N, V>::T;
}
template <typename T, int N, int V = 0> __host__ __device__ constexpr int ps_of()
{
    constexpr int base = (sizeof(T) == 8) ? (N >= 512 ? 4 : 8) : (N >= 512 ? 8 : 16);
    // a pencil of 1024 takes 17 KB (FP64) of exchange space per column: a
    // wider tile would not fit the last-axis pass's LDS
    if (sizeof(T) == 8 && N == 256 && V == 0) return 16;
    return (V == 0 || N >= 1024) ? base : 2 * base;
}

// The arguments of this block: the by-value kernel argument, or this
// instance's entry of a device table (batched launches, blockIdx.y), read
// through the constant address space so that the loads are scalar.
template <typename A> __device__ __forceinline__ A args_of(const A& a0, const A* tab, unsigned inst)
{
    if (!tab) return a0;
    static_assert(sizeof(A) % 4 == 0, "argument structs are dword-sized");
    struct Raw {
        unsigned w[sizeof(A) / 4];
    } r;
#if defined(__HIP_DEVICE_COMPILE__)
    using cu = const __attribute__((address_space(4))) unsigned;
#else
    using cu = const unsigned;
#endif
    cu* src = (cu*)(tab + inst);
#pragma unroll
    for (unsigned i = 0; i < sizeof(A) / 4; ++i) r.w[i] = src[i];
    return __builtin_bit_cast(A, r);
}

template <typename T> __device__ __forceinline__ double2 to_d2(c2_t<T> a) { return double2{(double)a.x, (double)a.y}; }
template <typename T> __device__ __forceinline__ c2_t<T> from_d2(double2 a) { return c2_t<T>{(T)a.x, (T)a.y}; }

// ---------------------------------------------------------------------------
// Block reduction of NV doubles per thread into out[0..NV): wave64 shuffles,
// then one LDS slot per wave, summed in wave order (deterministic).
template <int NV> __device__ __forceinline__ void block_sum_store(double (&s)[NV], gp<double> out)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] += __shfl_down(s[q], off, 64);
    __shared__ double red[NV][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q][w] = s[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double a = 0.0;
            for (int i = 0; i < nw; ++i) a += red[q][i];
            out[q] = a;
        }
    }
}

// ---------------------------------------------------------------------------
// Standalone x-axis transform of rows of length N (the building-block API).
// The last block may hold pencils past the last row: they transform row 0
// again and store nothing.
template <typename T, int N, int V, int DIR>
__global__ void __launch_bounds__((px_of<T, N, V>() * Plan<N, V>::T))
k_fft_x(const c2_t<T>* in, c2_t<T>* out, const c2_t<T>* __restrict__ twg, long long nrows)
{
    using T2 = c2_t<T>;
    const gp<const T2> tw = glob(twg);
    constexpr int TT = Plan<N, V>::T, E = Plan<N, V>::E, P = px_of<T, N, V>();
    __shared__ T2 lds[P * lds_stride<N>()];
    const int t = threadIdx.x % TT, p = threadIdx.x / TT;
    const long long r = (long long)blockIdx.x * P + p;
    const bool pv = r < nrows;
    const long long row = pv ? r : 0;
    const T2* src = in + row * N;
    T2* dst = out + row * N;
    T2* my = lds + p * lds_stride<N>();
    T2 v[E];
    if (DIR < 0) {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = src[elemA<N, V>(t, q)];
        fft_forward<T2, N, V, true>(v, t, my, tw);
        if (pv)
#pragma unroll
            for (int q = 0; q < E; ++q) dst[elemB<N, V>(t, q)] = v[q];
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = src[elemB<N, V>(t, q)];
        fft_inverse<T2, N, V, true>(v, t, my, tw);
        if (pv)
#pragma unroll
            for (int q = 0; q < E; ++q) dst[elemA<N, V>(t, q)] = v[q];
    }
}

// ---------------------------------------------------------------------------
// Strided axis (y or z) transform, in place.  Block = tile of P consecutive x
// columns at one `outer` index of one channel.  Thread (p, t) with p fastest,
// so a row segment of P columns is one coalesced access.
//
// Channels [0, ncomp) hold complex fields (psi).  The channels after them hold
// real fields (the density), either in the same layout or, with hcols > 0, as
// a half spectrum: after the x transform of a real row only kx = 0..nx/2 are
// stored (the rest are conjugates), hcols = nx/2 + 1 columns in rows of
// pitch hpitch.  Those channels take ceil(hcols/P) tiles per outer index; the
// lanes of a last partial tile compute on zeros and store nothing.
//
// A launch may cover a slab of columns only: complex columns [x0, x0 + ncols)
// and half-layout columns [hx0, hx0 + hcols), so that the y-forward, last-axis
// and y-inverse passes of one slab run back to back while the slab is still
// in the memory-side cache.
//
// scale (forward only, may be null): channel c < ncomp is multiplied by
// scale[c], a real-field channel by scale[0]^2 (the pending renormalisation
// of the imaginary-time schemes).
//
// Batched launches: the pass kernels take a table of per-instance argument
// structs (device memory, one per blockIdx.y) for several contexts of one
// shape; with a null table they use the by-value argument.
struct StridedArgs {
    void* data;             // the channels (c2_t<T>)
    int nx;                 // x extent of the grid
    int ncols, x0;          // complex columns of this launch
    int hx0;                // first half-layout column of this launch
    long long stride;       // element stride along the axis (complex layout)
    long long outer_stride; // stride of the outer index (complex layout)
    long long chan_stride;  // elements per complex-layout channel
    int nouter;             // extent of the outer index
    int ncomp;              // complex channels
    int nchan;              // channels in all
    int hcols;              // > 0: real channels in half-spectrum layout, columns of this launch
    long long hstride;      // half layout: element stride along the axis
    long long houter;       //              stride of the outer index
    long long hbase;        //              offset of the first real channel
    long long hchan;        //              elements per real channel
    const double* scale;    // per-component pending scale (forward only) or null
    const int* stop;        // nonzero: skip (a converged imaginary-time loop)
    int rev;                // 1: blocks take the tiles in reverse order
    int gather;             // 1: the half layout's leftover columns go to remainder tiles (below)
};

// Blocks of one strided launch: per outer index the complex-layout tiles,
// then the half-layout ones.  A half-layout row has hcols = nx/2 + 1
// columns, so hcols % P of them (the Nyquist column kx = nx/2, for P
// dividing nx/2) are left over per outer index.  Without `gather` each outer
// index spends a partial tile on them; with it, the per-outer-index
// half-layout tiles are the full ones only, and remainder tiles after all
// outer indices gather P leftover columns of consecutive outer indices (and
// channels), one per lane.  The last-axis pass gathers (its partial tiles run
// two transforms for one column; measured at 256^3, FP32: 72 -> 69 us); the
// y passes do not (their remainder pencils stride whole planes apart, and
// the slow tiles at the end of the launch cost more than they save: 128^3
// batched, 131 -> 151 us).  A partial tile's spare lanes compute on zeros and
// store nothing.
template <int P> __host__ __device__ inline long long strided_blocks(const StridedArgs& a)
{
    const long long ntf = (a.ncols + P - 1) / P;
    if (a.hcols == 0) return (long long)a.nchan * ntf * a.nouter;
    const long long nreal = a.nchan - a.ncomp;
    if (!a.gather) return ((long long)a.ncomp * ntf + nreal * ((a.hcols + P - 1) / P)) * a.nouter;
    const long long nhf = a.hcols / P, hr = a.hcols % P;
    return ((long long)a.ncomp * ntf + nreal * nhf) * a.nouter + (nreal * a.nouter * hr + P - 1) / P;
}

struct TileAt {
    long long off;      // element offset of this thread's column at axis index 0
    long long st;       // element stride along the axis
    long long sp_off;   // complex-layout offset of the same (x, outer) position, channel-relative
    long long outer;
    int ch, ix;
    bool valid;         // this lane's column is stored
};

template <int P> __device__ __forceinline__ TileAt strided_tile(const StridedArgs& a, int p, long long tile,
                                                               long long ntiles)
{
    // Tiles are enumerated outer index by outer index, all channels of one
    // outer index together (complex-layout tiles, then half-layout ones):
    // the x stage writes both channels of a row together, so a pass that
    // starts where it ended finds both in the memory-side cache; the
    // remainder tiles come last.  Selects, not branches: every value but those
    // of a remainder tile is block-uniform.
    TileAt r;
    const int ntf = (a.ncols + P - 1) / P;
    // half-layout tiles per outer index and channel, and (gather) the columns
    // per outer index and channel left to the remainder tiles
    const int nhf = a.gather ? a.hcols / P : (a.hcols + P - 1) / P;
    const int hr = a.gather ? a.hcols - nhf * P : 0;
    const int nfc = a.hcols ? a.ncomp : a.nchan;          // complex-layout channels
    const long long nfo = (long long)nfc * ntf;           // their tiles per outer index
    const long long per = nfo + (a.hcols ? (long long)(a.nchan - a.ncomp) * nhf : 0);
    // the tile this block takes: in launch order, or reversed so that the pass
    // starts on what the previous pass wrote last (still in the memory-side cache)
    const long long bx = a.rev ? ntiles - 1 - tile : tile;
    const long long nmain = per * a.nouter;
    if (bx >= nmain) {
        // remainder tile: lane p takes remainder pencil g (channel-major, then
        // outer index, then column)
        const long long npen = (long long)(a.nchan - a.ncomp) * a.nouter * hr;
        const long long g0 = (bx - nmain) * P + p;
        r.valid = g0 < npen;
        const long long g = r.valid ? g0 : 0;
        const long long perch = (long long)a.nouter * hr;
        const int cl = (int)(g / perch);
        const long long w = g - cl * perch;
        r.outer = w / hr;
        r.ix = a.hx0 + nhf * P + (int)(w - r.outer * hr);
        r.ch = a.ncomp + cl;
        r.sp_off = r.outer * a.outer_stride + r.ix;
        r.off = a.hbase + cl * a.hchan + r.outer * a.houter + r.ix;
        r.st = a.hstride;
        return r;
    }
    r.outer = bx / per;
    const long long w = bx - r.outer * per;
    const bool half = w >= nfo;
    const long long wl = half ? w - nfo : w;
    const int nt = half ? nhf : ntf;
    const int cl = (int)(wl / nt);            // channel within its layout
    const int xt = (int)(wl - (long long)cl * nt);
    r.ch = half ? a.ncomp + cl : cl;
    const int loc = xt * P + p;   // column within this launch
    r.valid = loc < (half ? a.hcols : a.ncols);
    r.ix = (half ? a.hx0 : a.x0) + loc;
    if (!r.valid) r.ix = 0;   // keeps table lookups in range; nothing is stored
    r.sp_off = r.outer * a.outer_stride + r.ix;
    // blends by mask: a select between two struct fields would be folded into
    // a load through a selected address, which keeps the struct in scratch
    const long long m = half ? -1LL : 0LL;
    const long long off_h = a.hbase + cl * a.hchan + r.outer * a.houter + r.ix;
    const long long off_c = cl * a.chan_stride + r.sp_off;
    r.off = (off_h & m) | (off_c & ~m);
    r.st = (a.hstride & m) | (a.stride & ~m);
    return r;
}

// The FP64 radix-16 y pass (plan 0 at length 256, 16 pencils, 256 threads)
// splits its exchange like the last-axis pass below: 39 KB of LDS instead of
// 74, so three blocks share a CU at its register budget of 3 waves per SIMD
// (132 VGPRs, no spill); in a running step at 256^3 it then takes 127 us
// against 131 unsplit and 133 for plan 1.
template <typename T, int N, int V> constexpr bool strided_split()
{
    return sizeof(T) == 8 && N == 256 && V == 0;
}
#ifndef FGPE_STRIDED_WAVES
#define FGPE_STRIDED_WAVES 3
#endif
template <typename T, int N, int V> constexpr int strided_waves()
{
    return strided_split<T, N, V>() ? FGPE_STRIDED_WAVES : 1;
}

template <typename T, int N, int V, int DIR>
__global__ void __launch_bounds__((ps_of<T, N, V>() * Plan<N, V>::T))
__attribute__((amdgpu_waves_per_eu(strided_waves<T, N, V>())))
k_strided(const c2_t<T>* __restrict__ twg, StridedArgs a0, const StridedArgs* __restrict__ tab)
{
    using T2 = c2_t<T>;
    constexpr int TT = Plan<N, V>::T, E = Plan<N, V>::E, P = ps_of<T, N, V>();
    using LT = std::conditional_t<strided_split<T, N, V>(), T, T2>;   // exchange buffer element
    __shared__ LT lds[P * lds_stride<N>()];
    const StridedArgs a = args_of(a0, tab, blockIdx.y);
    if (a.stop && *glob(a.stop)) return;
    __shared__ T2 twl[N];
    const lds_table<T2> tw = stage_twiddles<T2, N>(twl, twg);
    const gp<T2> data = glob((T2*)a.data);
    const int p = threadIdx.x % P, t = threadIdx.x / P;
    const TileAt at = strided_tile<P>(a, p, blockIdx.x, gridDim.x);
    const gp<T2> base = data + at.off;
    LT* my = lds + p * lds_stride<N>();
    T2 v[E];
    if (DIR < 0) {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = at.valid ? base[(long long)elemA<N, V>(t, q) * at.st] : T2{0, 0};
        if (a.scale) {
            double s = glob(a.scale)[at.ch < a.ncomp ? at.ch : 0];
            if (at.ch >= a.ncomp) s = s * s;
            const T st = (T)s;
#pragma unroll
            for (int q = 0; q < E; ++q) v[q] = T2{v[q].x * st, v[q].y * st};
        }
        fft_forward<T2, N, V>(v, t, my, tw);
        if (at.valid)
#pragma unroll
            for (int q = 0; q < E; ++q) base[(long long)elemB<N, V>(t, q) * at.st] = v[q];
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = at.valid ? base[(long long)elemB<N, V>(t, q) * at.st] : T2{0, 0};
        fft_inverse<T2, N, V>(v, t, my, tw);
        if (at.valid)
#pragma unroll
            for (int q = 0; q < E; ++q) base[(long long)elemA<N, V>(t, q) * at.st] = v[q];
    }
}

// ---------------------------------------------------------------------------
// Last axis: forward FFT, k-space multiply, inverse FFT, in place.
// Channel c < ncomp (psi):  v *= (19.739208802/Ntot) * |k|^2   (src/parallel.cu:194-205)
// Real-field channel (density): v *= Vt(k) / Ntot, Vt = dipolar (inline,
//   src/parallel.cu:326-352) + table (soft-core, may be null).
// The density result is used through its real part, Re F^-1[Vt F(n)], which
// for a real n equals F^-1[Vs F(n)] with Vs(k) = (Vt(k) + Vt(k'))/2, k' the
// index-negated wave vector.  Vs is what the half-spectrum layout needs (its
// product with F(n) stays Hermitian).  Off the Nyquist planes k' = -k exactly
// and Vt(k') = Vt(k) bit for bit, so only Nyquist-plane entries change.
struct LastArgs {
    StridedArgs s;
    double dA, dX, dO;  // dipole direction components along A (this axis), X, O (outer axis)
    double lap_scale;   // 19.739208802 / Ntot
    double inv_ntot;    // 1 / Ntot
    double c_dd;        // 12 pi a_dd (0: no dipolar term)
    const void* wtab;   // T table of Vt (complex-layout spatial index) or null
    double vpx[2];      // moving-frame factor vpx_c / Ntot per component (0: off)
    const double2* kA64;   // k along A, X, O as pairs (k[i], k[(n - i) % n]): k and the
    const double2* kX64;   // index-negated partner in one load
    const double2* kO64;
    const float2* kA32;    // the same pairs rounded to FP32 (the mixed form)
    const float2* kX32;
    const float2* kO32;
};

// 12 pi a_dd ((k.d)^2/|k|^2 - 1/3), 0 for |k| < 1e-6 (src/parallel.cu:336-349)
__device__ __forceinline__ double vdd_of(double c_dd, double num, double k2)
{
    const double den = sqrt(k2);
    if (den < 1E-6) return 0.0;
    return c_dd * (((num / den) * (num / den)) - 1.0 / 3.0);
}

// The moving-frame factor of complex channel ch (0 or 1).  Blended
// arithmetically: an indexed or selected field would be turned into a load
// through a computed address, which keeps the whole argument struct in
// scratch memory.
__device__ __forceinline__ double vpx_of(const LastArgs& a, int ch)
{
    const double v0 = a.vpx[0], v1 = a.vpx[1];
    return v0 + (double)(ch != 0) * (v1 - v0);
}

// The k-space multiply of the FP64 form.
template <int N, int V>
__device__ __forceinline__ void last_multiply_f64(double2* v, int t, const LastArgs& a, const TileAt& at,
                                                  lds_table<double2> kA, double2 kxp2, double2 kop2)
{
    constexpr int E = Plan<N, V>::E;
    // k and its index-negated partner k' come as one 16-byte pair
    const double kx = kxp2.x, ko = kop2.x;
    const double kxo2 = kx * kx + ko * ko;
    if (at.ch < a.s.ncomp) {
        const double w = vpx_of(a, at.ch) * kx;   // moving frame: - vpx k_x / N
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const double2 kk = kA[elemB<N, V>(t, q)];
            const double ka = kk.x;
            const double k2 = kxo2 + ka * ka;
            // w = 0 leaves (lap_scale v) k2 unchanged (a zero product is exact)
            v[q] = double2{(a.lap_scale * v[q].x) * k2 - w * v[q].x, (a.lap_scale * v[q].y) * k2 - w * v[q].y};
        }
        return;
    }
    // Vt rows are addressed from one base: every in-grid offset fits 32 bits.
    const gp<const double> wt = (a.wtab && at.valid) ? glob((const double*)a.wtab) + at.sp_off : nullptr;
    const double kxp = kxp2.y, kop = kop2.y;
    const double num0 = kx * a.dX + ko * a.dO;
    const double num0p = kxp * a.dX + kop * a.dO;
    const double kxo2p = kxp * kxp + kop * kop;
    const bool xo_neg = (kxp == -kx) && (kop == -ko);
    const int st = (int)a.s.stride;
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int ea = elemB<N, V>(t, q);
        double vt = 0.0;
        if (a.c_dd != 0.0) {
            const double2 kk = kA[ea];
            const double ka = kk.x, kap = kk.y;
            vt = vdd_of(a.c_dd, num0 + ka * a.dA, kxo2 + ka * ka);
            if (!(xo_neg && kap == -ka))
                vt = 0.5 * (vt + vdd_of(a.c_dd, num0p + kap * a.dA, kxo2p + kap * kap));
        }
        if (wt) vt = vt + wt[ea * st];
        const double f = vt * a.inv_ntot;
        v[q] = double2{v[q].x * f, v[q].y * f};
    }
}

// The k-space multiply of the mixed-precision form: factors in FP32 from
// FP32 copies of the k axes (the data is FP32), same formulas and the same
// Nyquist symmetrisation as the FP64 form, the partner test k' = -k made
// on the FP32 values.
// |k| < 1e-6 tested as k^2 < 1e-12; the quotient through the hardware
// reciprocal (1 ulp), well inside the FP32 bulk's own rounding.
// s / 3 in FP32 without a bias: q = s (1/3), corrected by the exact residual
// s - 3 q; error about half an ulp, against the 3e-8 relative bias of the
// product with the rounded 1/3 alone.
__device__ __forceinline__ float div3(float s)
{
    constexpr float third = 1.f / 3.f;
    const float q = s * third;
    return __builtin_fmaf(__builtin_fmaf(-3.f, q, s), third, q);
}

__device__ __forceinline__ float vdd_of_f32(float c_dd, float num, float k2)
{
    if (k2 < 1E-12f) return 0.f;
    return c_dd * ((num * num) * __builtin_amdgcn_rcpf(k2) - (1.f / 3.f));
}

template <int N, int V>
__device__ __forceinline__ void last_multiply_f32(float2* v, int t, const LastArgs& a, const TileAt& at,
                                                  lds_table<float2> kA, float2 kxp2, float2 kop2)
{
    constexpr int E = Plan<N, V>::E;
    const float kx = kxp2.x, ko = kop2.x;
    const float kxo2 = kx * kx + ko * ko;
    if (at.ch < a.s.ncomp) {
        const float ls = (float)a.lap_scale;
        const float w = (float)vpx_of(a, at.ch) * kx;   // moving frame: - vpx k_x / N
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const float2 kk = kA[elemB<N, V>(t, q)];
            const float ka = kk.x;
            const float f = ls * (kxo2 + ka * ka) - w;
            v[q] = float2{v[q].x * f, v[q].y * f};
        }
        return;
    }
    // Vt rows are addressed from one base: every in-grid offset fits 32 bits.
    const gp<const float> wt = (a.wtab && at.valid) ? glob((const float*)a.wtab) + at.sp_off : nullptr;
    const float kxp = kxp2.y, kop = kop2.y;
    const bool xo_neg = (kxp == -kx) && (kop == -ko);
    const float dA = (float)a.dA, dX = (float)a.dX, dO = (float)a.dO, c_dd = (float)a.c_dd;
    const float num0 = kx * dX + ko * dO, num0p = kxp * dX + kop * dO;
    const float kxo2p = kxp * kxp + kop * kop;
    const float inv = (float)a.inv_ntot;
    const int st = (int)a.s.stride;
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int ea = elemB<N, V>(t, q);
        float vt = 0.f;
        if (c_dd != 0.f) {
            const float2 kk = kA[ea];
            const float ka = kk.x, kap = kk.y;
            vt = vdd_of_f32(c_dd, num0 + ka * dA, kxo2 + ka * ka);
            if (!(xo_neg && kap == -ka))
                vt = 0.5f * (vt + vdd_of_f32(c_dd, num0p + kap * dA, kxo2p + kap * kap));
        }
        if (wt) vt = vt + wt[ea * st];
        const float f = vt * inv;
        v[q] = float2{v[q].x * f, v[q].y * f};
    }
}

// The FP64 radix-16 last-axis pass (16 pencils of 256, 256 threads) exchanges
// the real and the imaginary parts in turn through a scalar LDS buffer: 43 KB
// per block instead of 78, so three blocks share a CU and one block's
// transforms overlap the others' HBM traffic (the pass went from 166 to
// 154 us at 256^3; its access pattern alone, chained as in the step, takes
// about 122 us).
template <typename T, int N, int V> constexpr bool last_split()
{
    return sizeof(T) == 8 && N == 256 && V == 0;
}

// Waves per SIMD the radix-16 last-axis passes are compiled for: four for
// FP32 (at most 128 VGPRs) keeps four blocks per CU in flight while the
// register-resident pencil waits on HBM; three for the split FP64 pass (at
// most 168 VGPRs, three 4-wave blocks).
template <typename T, int N, int V> constexpr int last_waves()
{
    if constexpr (last_split<T, N, V>()) return 3;
    return (sizeof(T) == 4 && N == 256 && V == 0) ? 4 : 1;
}

template <typename T, int N, int V>
__global__ void __launch_bounds__((ps_of<T, N, V>() * Plan<N, V>::T))
__attribute__((amdgpu_waves_per_eu(last_waves<T, N, V>())))
k_last(const c2_t<T>* __restrict__ twg, LastArgs a0, const LastArgs* __restrict__ tab)
{
    using T2 = c2_t<T>;
    constexpr int TT = Plan<N, V>::T, E = Plan<N, V>::E, P = ps_of<T, N, V>();
    using LT = std::conditional_t<last_split<T, N, V>(), T, T2>;   // exchange buffer element
    __shared__ LT lds[P * lds_stride<N>()];
    const LastArgs a = args_of(a0, tab, blockIdx.y);
    if (a.s.stop && *glob(a.s.stop)) return;
    // the last axis's k pairs, read once per element by the multiply, are
    // staged in LDS with the twiddle table (one barrier for both)
    __shared__ T2 twl[N];
    __shared__ T2 kal[N];
    {
        const gp<const T2> kag = glob((const T2*)(sizeof(T) == 8 ? (const void*)a.kA64 : (const void*)a.kA32));
        for (int i = threadIdx.x; i < N; i += blockDim.x) kal[i] = kag[i];
    }
    const lds_table<T2> tw = stage_twiddles<T2, N>(twl, twg);
    const lds_table<T2> kA = (lds_table<T2>)kal;
    const gp<T2> data = glob((T2*)a.s.data);
    const int p = threadIdx.x % P, t = threadIdx.x / P;
    const TileAt at = strided_tile<P>(a.s, p, blockIdx.x, gridDim.x);
    const gp<T2> base = data + at.off;
    LT* my = lds + p * lds_stride<N>();
    // this column's k_x and outer-axis k pairs, loaded before the pencil so
    // that their latency hides behind it rather than after the transform
    const T2 kxp2 = glob((const T2*)(sizeof(T) == 8 ? (const void*)a.kX64 : (const void*)a.kX32))[at.ix];
    const T2 kop2 = glob((const T2*)(sizeof(T) == 8 ? (const void*)a.kO64 : (const void*)a.kO32))[at.outer];
    T2 v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = at.valid ? base[(long long)elemA<N, V>(t, q) * at.st] : T2{0, 0};
    if (a.s.scale) {
        double s = glob(a.s.scale)[at.ch < a.s.ncomp ? at.ch : 0];
        if (at.ch >= a.s.ncomp) s = s * s;
        const T st = (T)s;
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = T2{v[q].x * st, v[q].y * st};
    }
    fft_forward<T2, N, V>(v, t, my, tw);

    if constexpr (sizeof(T) == 4) last_multiply_f32<N, V>(v, t, a, at, kA, kxp2, kop2);
    else last_multiply_f64<N, V>(v, t, a, at, kA, kxp2, kop2);
    fft_inverse<T2, N, V>(v, t, my, tw);
    if (at.valid)
#pragma unroll
        for (int q = 0; q < E; ++q) base[(long long)elemA<N, V>(t, q) * at.st] = v[q];
}

// ---------------------------------------------------------------------------
// 1D grids: the k-space multiply between the x stage's forward and inverse
// transforms (there is no pass along another axis to carry it), the same
// factors as the last-axis pass with k = k_x: (19.739208802/N) k^2 - vpx_c k
// on component c (src/parallel.cu:194-205) and Vs(k)/N on the real-field
// channels' half spectra (kx = 0..N/2).  One thread per spectral element,
// complex channels first (natural order), factors in FP64.  A pending scale
// multiplies component c by scale[c] and the density by scale[0]^2.
template <typename T> __global__ void __launch_bounds__(256) k_kmul_1d(LastArgs a)
{
    using T2 = c2_t<T>;
    if (a.s.stop && *glob(a.s.stop)) return;
    const int nx = a.s.nx, nc = a.s.ncomp;
    const int nh = a.s.nchan > nc ? (a.s.nchan - nc) * a.s.hcols : 0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc * nx + nh) return;
    const gp<T2> data = glob((T2*)a.s.data);
    const gp<const double2> kx = glob(a.kX64);
    const gp<const double> scale = glob(a.s.scale);
    if (i < nc * nx) {
        const int c = i / nx, e = i - c * nx;
        const double k = kx[e].x;
        const double w = vpx_of(a, c) * k;
        const long long at = (long long)c * a.s.chan_stride + e;
        double2 z = to_d2<T>(data[at]);
        if (a.s.scale) z = double2{scale[c] * z.x, scale[c] * z.y};
        data[at] = from_d2<T>(double2{(a.lap_scale * z.x) * (k * k) - w * z.x, (a.lap_scale * z.y) * (k * k) - w * z.y});
        return;
    }
    const int j = i - nc * nx, ch = j / a.s.hcols;   // real-field channel ch, element e
    const int e = j - ch * a.s.hcols;
    const double k = kx[e].x, kp = kx[e].y;
    double vt = 0.0;
    if (a.c_dd != 0.0) {
        vt = vdd_of(a.c_dd, k * a.dX, k * k);
        if (kp != -k) vt = 0.5 * (vt + vdd_of(a.c_dd, kp * a.dX, kp * kp));
    }
    if (a.wtab) vt = vt + (sizeof(T) == 8 ? (double)glob((const double*)a.wtab)[e] : (double)glob((const float*)a.wtab)[e]);
    double f = vt * a.inv_ntot;
    if (a.s.scale) f = f * (scale[0] * scale[0]);
    const long long at = a.s.hbase + ch * a.s.hchan + e;
    const T2 v = data[at];
    data[at] = T2{(T)((double)v.x * f), (T)((double)v.y * f)};
}

// ---------------------------------------------------------------------------
// x-axis stage kernel, NC = 1 or 2 components.
//   MODE_PROLOGUE: spec <- Fx(psi_c) [, Fx(sum_c |psi_c|^2)] of the stored
//                  (raw) state; a pending scale is applied by the next forward
//                  pass.
//   MODE_RK:       inverse x of spec -> KE_c [, Phi]; H_c at every point; RK4
//                  stage `stage` (src/oneComponentGPSolver.cu:165-187 and
//                  src/twoComponentGPSolver.cu:150-176: UpdateRKStep /
//                  FinalRKStep, src/parallel.cu:230-271); spec <- Fx(next
//                  state) [, Fx(its density)].
//   MODE_APPLY:    inverse x -> H_c and Hen_c (BECHamiltonian,
//                  hamiltonians/BEC_oneComponent.cu:128-150; BECHamiltonian_2,
//                  hamiltonians/BEC_twoComponent.cu:168-212) to hmu/hen (if
//                  non-null) and per-block sums |psi|^2, Re(psi* H), Re(psi* Hen).
//   MODE_HB:       heavy-ball imaginary-time update (BECUpdatePsi,
//                  src/parallel.cu:294-302): new = (1 + mu dt) psi - dt H
//                  + beta (psi - old); P0 <- new (raw), PS <- psi (the next
//                  old), sums of new; spec <- Fx(new) [, Fx(|new|^2)].
// H_c = KE_c + (V_c.x + g4[c][c] n_c + g4[c][c'] n_c' + Phi) psi_c + i V_c.y psi_c,
// Hen_c with g2 and Phi/2.
enum { MODE_PROLOGUE = 0, MODE_RK = 1, MODE_APPLY = 2, MODE_HB = 3 };

template <typename T> struct XArgs {
    c2_t<T>* spec;          // complex channels, component c at c*ntot
    long long ntot;
    c2_t<T>* specd;         // density channel, half spectrum: row r at r*dpitch, kx = 0..N/2
    int dpitch;
    double2* P0[2];         // FP64 state psi_n (RK: psi_old of the step)
    c2_t<T>* PS[2];         // RK: stage state; HB: previous psi
    c2_t<T>* PSo[2];        // RK stages 1-3: where the next stage state goes (null: PS)
    c2_t<T>* PS1[2];        // mixed RK stage 3: the stage-1 state (null: the FP64 state is read)
    c2_t<T>* K[2];          // RK accumulator
    const T* Vr[2];         // Re V_c
    const T* Vi[2];         // Im V_c or null
    double g4[2][2];        // 4 pi a_cc'
    double g2[2][2];        // 2 pi a_cc'
    double cr, ci;          // RK stage coefficient c
    int stage;              // RK stage 1..4
    const double* scale;    // pending renormalisation per component (null: 1)
    double* partials;       // 3*NC doubles per block
    int want_norm;          // RK stage 4: per-block sums of |psi_new|^2
    double2* hmu;           // APPLY outputs, component c at c*ntot
    double2* hen;
    double dt, beta;        // HB
    const double* mu;       // HB: mu_c of the previous iteration (null: 0)
    const int* stop;        // nonzero: the loop has converged, do nothing
    int rev;                // 1: blocks take the row groups in reverse order
    int ny;                 // grid rows per z plane (the user term's grid index)
    double t;               // time of this H application (the user term)
    double user[8];         // the user term's parameters
};

// Minimum waves per SIMD the register allocator targets for an x-stage
// kernel: 4 for the FP32 RK stages of the variant-1 plans up to N = 256
// (256: 138 -> 128 VGPRs with 3 spilled, 9% faster; shorter pencils fit
// without spills); for the one-component RK stages of the variant-2 plans
// (4 elements per thread) 7 in FP32 (N = 256: 81 -> 70 VGPRs) and 5 in FP64
// (108 -> 96 VGPRs), none spilled (at 256^3 the x stage went from 345 to
// 335 us FP64 and from 191 to 181 us FP32); 1 (no target) elsewhere, where
// forcing it spills heavily.
template <typename T, int N, int V, int NC, int MODE> constexpr int xstage_waves()
{
    // FP32 at nx = 256: 8 waves (64 VGPRs, a 12-byte spill) ran 178.5 us
    // against 181-182 at 7 (70 VGPRs, no spill), the mixed step +0.8%
    if constexpr (V == 2 && NC == 1 && MODE == MODE_RK) return sizeof(T) == 4 ? (N == 256 ? 8 : 7) : 5;
    return (sizeof(T) == 4 && V == 1 && NC == 1 && MODE == MODE_RK && N <= 256 && Plan<N, V>::E <= 8) ? 4 : 1;
}

// Accesses of the x stage's pointwise operands (state, stage state, RK
// accumulator, trap).  With NT they are non-temporal: these streams are read
// or written once per stage, and when the working set is larger than the
// 256 MB memory-side cache, letting them allocate there evicts the spectral
// channels that the x stage itself and the passes around it re-read
// (measured at 256^3: x stage -8%, step +5% FP64 / +6% mixed).  When it all
// fits, the hint only sends hits to HBM, so the launcher picks NT from the
// working-set size.  Every access of these arrays in the kernel goes through
// here, so they are always reached through the same (native vector) type.
template <bool NT, typename V> __device__ __forceinline__ V ldn(gp<const V> p, long long i)
{
    if constexpr (!NT) {
        return p[i];
    } else if constexpr (std::is_arithmetic_v<V>) {
        return __builtin_nontemporal_load(&p[i]);
    } else {
        using NV = typename V::Native_vec_;
        return __builtin_bit_cast(V, __builtin_nontemporal_load((gp<const NV>)(&p[i])));
    }
}
template <bool NT, typename V> __device__ __forceinline__ V ldn(gp<V> p, long long i)
{
    return ldn<NT, V>((gp<const V>)p, i);
}
template <bool NT, typename V> __device__ __forceinline__ void stn(gp<V> p, long long i, const V& v)
{
    if constexpr (!NT) {
        p[i] = v;
    } else if constexpr (std::is_arithmetic_v<V>) {
        __builtin_nontemporal_store(v, &p[i]);
    } else {
        using NV = typename V::Native_vec_;
        __builtin_nontemporal_store(__builtin_bit_cast(NV, v), (gp<NV>)(&p[i]));
    }
}

// The user term's inputs at linear index g (rows of N points along x).
template <int N, typename T>
__device__ __forceinline__ UserTermIn user_in(const XArgs<T>& a, long long g, int c, double pr, double pi, double n,
                                              double no)
{
    const long long r = g / N;
    return UserTermIn{a.t, (int)(g - r * N), (int)(r % a.ny), (int)(r / a.ny), c, pr, pi, n, no, a.user};
}

// UT: the user term compiled in (UserTerm<UT>; 0: none).
template <typename T, int N, int V, int NC, bool DENS, int MODE, bool NT = false, int UT = 0>
__global__ void __launch_bounds__((px_of<T, N, V>() * Plan<N, V>::T))
__attribute__((amdgpu_waves_per_eu(xstage_waves<T, N, V, NC, MODE>())))
k_xstage(XArgs<T> a0, const c2_t<T>* __restrict__ twg, const XArgs<T>* __restrict__ tab)
{
    using UTerm = UserTerm<UT>;
    using T2 = c2_t<T>;
    constexpr int TT = Plan<N, V>::T, E = Plan<N, V>::E, P = px_of<T, N, V>();
    constexpr int NV = 3 * NC;
    __shared__ T2 lds[P * lds_stride<N>()];
    const XArgs<T> a = args_of(a0, tab, blockIdx.y);
    if (a.stop && *glob(a.stop)) return;
    __shared__ T2 twl[N];
    const lds_table<T2> tw = stage_twiddles<T2, N>(twl, twg);
    const gp<T2> spec = glob(a.spec);
    const gp<double2> P0[2] = {glob(a.P0[0]), glob(a.P0[1])};
    const gp<T2> PS[2] = {glob(a.PS[0]), glob(a.PS[1])};
    const gp<T2> PSo[2] = {glob(a.PSo[0]), glob(a.PSo[1])};
    const gp<const T2> PS1[2] = {glob((const T2*)a.PS1[0]), glob((const T2*)a.PS1[1])};
    const gp<T2> K[2] = {glob(a.K[0]), glob(a.K[1])};
    const gp<const T> Vr[2] = {glob(a.Vr[0]), glob(a.Vr[1])};
    const gp<const T> Vi[2] = {glob(a.Vi[0]), glob(a.Vi[1])};
    const int t = threadIdx.x % TT, p = threadIdx.x / TT;
    // row group of this block (reversed: the kernel starts on what the passes
    // before it wrote last); partial sums stay indexed by the row group.  The
    // last group may hold pencils past the last row (grids with fewer rows
    // than a block's pencils, e.g. 1D): their threads skip the pencil work
    // (the x exchanges are wave-scoped, so idle lanes are harmless) and only
    // take part in the block's sum with zeros.
    const unsigned bx = a.rev ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
    const long long rowi = (long long)bx * P + p;
    const bool pv = rowi < a.ntot / N;
    const long long row = pv ? rowi : 0;
    const long long row0 = row * N;
    T2* my = lds + p * lds_stride<N>();
    const gp<T2> specd = glob(a.specd) + row * a.dpitch;   // this row's half spectrum

    double sums[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) sums[i] = 0.0;

    if (pv) {
        T2 v[E];
        T nd[E];             // density of the new state (DENS)
#pragma unroll
        for (int q = 0; q < E; ++q) nd[q] = (T)0;

        if (MODE == MODE_PROLOGUE) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    v[q] = from_d2<T>(P0[c][row0 + elemA<N, V>(t, q)]);
                    if (DENS) nd[q] += v[q].x * v[q].x + v[q].y * v[q].y;
                }
                fft_forward<T2, N, V, true>(v, t, my, tw);
                const gp<T2> sc = spec + c * a.ntot + row0;
#pragma unroll
                for (int q = 0; q < E; ++q) sc[elemB<N, V>(t, q)] = v[q];
            }
        } else {
            // phi, the k-space potential of the density, is held from the start
            // of the stage to the pointwise update, across the inverse transform
            // of psi.  Its first QL register slots live in LDS instead: in
            // registers it costs the FP64 kernel 66 VGPRs (202 -> 136 with all of
            // it in LDS), but all of it in LDS makes the FP64 block too large for
            // three blocks per CU, so the FP64 form keeps half of it in registers.
            constexpr int QL = sizeof(T) == 8 ? E / 2 : E;
            __shared__ T sphi[DENS ? P * TT * QL : 1];
            T phi[E];
            auto phi_at = [&](int q) -> T& { return q < QL ? sphi[q * (P * TT) + threadIdx.x] : phi[q]; };
            if (DENS) {
                // kx > N/2 from the conjugate symmetry of a real row's transform
#pragma unroll
                for (int q = 0; q < E; ++q) {
                    const int e = elemB<N, V>(t, q);
                    if (e <= N / 2) {
                        v[q] = specd[e];
                    } else {
                        const T2 w = specd[N - e];
                        v[q] = T2{w.x, -w.y};
                    }
                }
                fft_inverse<T2, N, V, true>(v, t, my, tw);
#pragma unroll
                for (int q = 0; q < E; ++q) phi_at(q) = v[q].x;
            }
            double nsave[E];     // NC == 2: density of component 0's stage input
            const bool from_p0 = (MODE == MODE_APPLY || MODE == MODE_HB || a.stage == 1);

#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const gp<T2> sc = spec + c * a.ntot + row0;
#pragma unroll
                for (int q = 0; q < E; ++q) v[q] = sc[elemB<N, V>(t, q)];
                fft_inverse<T2, N, V, true>(v, t, my, tw);   // v = KE psi_c (layout A)
                const double sig = a.scale ? glob(a.scale)[c] : 1.0;
                const double sig1 = (NC == 2 && a.scale) ? glob(a.scale)[1] : 1.0;
                const double g4s = a.g4[c][c], g2s = a.g2[c][c];
                const double g4o = a.g4[c][1 - c], g2o = a.g2[c][1 - c];
                if constexpr (!(MODE == MODE_RK && sizeof(T) == 4)) {
        #pragma unroll
                    for (int q = 0; q < E; ++q) {
                        const long long g = row0 + elemA<N, V>(t, q);
                        double2 po = double2{0.0, 0.0};
                        double2 ps;
                        if (from_p0) {
                            const double2 r = ldn<NT>(P0[c], g);
                            po = double2{sig * r.x, sig * r.y};
                            ps = po;
                        } else {
                            ps = to_d2<T>(ldn<NT>(PS[c], g));
                        }
                        const double n = ps.x * ps.x + ps.y * ps.y;
                        double no = 0.0;
                        if (NC == 2) {
                            if (c == 0) {
                                double2 o;
                                if (from_p0) {
                                    const double2 r = ldn<NT>(P0[1], g);
                                    o = double2{sig1 * r.x, sig1 * r.y};
                                } else {
                                    o = to_d2<T>(ldn<NT>(PS[1], g));
                                }
                                no = o.x * o.x + o.y * o.y;
                                nsave[q] = n;
                            } else {
                                no = nsave[q];
                            }
                        }
                        const double vr = (double)ldn<NT>(Vr[c], g);
                        const double vi = Vi[c] ? (double)ldn<NT>(Vi[c], g) : 0.0;
                        const double2 ke = to_d2<T>(v[q]);
                        double rX = (vr + g4s * n);
                        if (NC == 2) rX = rX + g4o * no;
                        if (DENS) rX += (double)phi_at(q);
                        double2 h = double2{ke.x + rX * ps.x - vi * ps.y, ke.y + rX * ps.y + vi * ps.x};
                        double2 he = double2{0.0, 0.0};
                        if (MODE == MODE_APPLY || MODE == MODE_HB) {
                            double eX = (vr + g2s * n);
                            if (NC == 2) eX = eX + g2o * no;
                            if (DENS) eX += 0.5 * (double)phi_at(q);
                            he = double2{ke.x + eX * ps.x - vi * ps.y, ke.y + eX * ps.y + vi * ps.x};
                        }
                        if constexpr (UTerm::enabled) {
                            const UserTermOut w = UTerm::eval(user_in<N>(a, g, c, ps.x, ps.y, n, no));
                            h = double2{h.x + (w.wr * ps.x - w.wi * ps.y), h.y + (w.wr * ps.y + w.wi * ps.x)};
                            if (MODE == MODE_APPLY || MODE == MODE_HB)
                                he = double2{he.x + (w.er * ps.x - w.ei * ps.y), he.y + (w.er * ps.y + w.ei * ps.x)};
                        }
                        if (MODE == MODE_APPLY) {
                            if (a.hmu) glob(a.hmu)[c * a.ntot + g] = h;
                            if (a.hen) glob(a.hen)[c * a.ntot + g] = he;
                            sums[3 * c + 0] += n;
                            sums[3 * c + 1] += ps.x * h.x + ps.y * h.y;
                            sums[3 * c + 2] += ps.x * he.x + ps.y * he.y;
                        } else if (MODE == MODE_HB) {
                            const double2 old = to_d2<T>(ldn<NT>(PS[c], g));
                            const double m = a.mu ? glob(a.mu)[c] : 0.0;
                            const double2 nw = double2{(1.0 + m * a.dt) * ps.x - a.dt * h.x + a.beta * (ps.x - old.x),
                                                       (1.0 + m * a.dt) * ps.y - a.dt * h.y + a.beta * (ps.y - old.y)};
                            stn<NT>(PS[c], g, from_d2<T>(ps));
                            stn<NT>(P0[c], g, nw);
                            const double nh = nw.x * h.x + nw.y * h.y;
                            sums[3 * c + 0] += nw.x * nw.x + nw.y * nw.y;
                            sums[3 * c + 1] += nh;
                            if constexpr (UTerm::enabled) {
                                sums[3 * c + 2] += nw.x * he.x + nw.y * he.y;
                            } else {
                                // he = h - d psi with d = rX - eX, so <nw, he> = <nw, h> - d <nw, psi>:
                                // he itself is not formed (holding it for every element of the
                                // pencil doubled the kernel's registers at E >= 8)
                                double d = (g4s - g2s) * n;
                                if (NC == 2) d += (g4o - g2o) * no;
                                if (DENS) d += 0.5 * (double)phi_at(q);
                                sums[3 * c + 2] += nh - d * (nw.x * ps.x + nw.y * ps.y);
                            }
                            v[q] = from_d2<T>(nw);
                            if (DENS) nd[q] += v[q].x * v[q].x + v[q].y * v[q].y;
                            // Two elements at a time: without the fence the scheduler starts every
                            // element's loads and FP64 arithmetic at once and the kernel runs at
                            // one wave per SIMD (128^3 FP32: 512 -> 164 registers with both changes).
                            if (q % 2 == 1) __builtin_amdgcn_sched_barrier(0);
                        } else {
                            if (a.stage != 1) {
                                const double2 r = ldn<NT>(P0[c], g);
                                po = double2{sig * r.x, sig * r.y};
                            }
                            if (a.stage < 4) {
                                // UpdateRKStep: psi = old + c*h ; k = h (stage 1) or k + 2h
                                const double2 nw = double2{po.x + (a.cr * h.x - a.ci * h.y),
                                                           po.y + (a.cr * h.y + a.ci * h.x)};
                                if (a.stage == 1) {
                                    stn<NT>(K[c], g, from_d2<T>(h));
                                } else {
                                    const double2 k = to_d2<T>(ldn<NT>(K[c], g));
                                    stn<NT>(K[c], g, from_d2<T>(double2{k.x + 2.0 * h.x, k.y + 2.0 * h.y}));
                                }
                                const T2 st = from_d2<T>(nw);
                                stn<NT>(PSo[c] ? PSo[c] : PS[c], g, st);
                                v[q] = st;
                            } else {
                                // FinalRKStep: psi = old + c*(k + h)/6
                                const double2 k = to_d2<T>(ldn<NT>(K[c], g));
                                const double aX = k.x + h.x, aY = k.y + h.y;
                                const double2 nw = double2{po.x + (a.cr * aX - a.ci * aY) / 6.0,
                                                           po.y + (a.cr * aY + a.ci * aX) / 6.0};
                                stn<NT>(P0[c], g, nw);
                                v[q] = from_d2<T>(nw);
                                if (a.want_norm) sums[3 * c] += nw.x * nw.x + nw.y * nw.y;
                            }
                            if (DENS) nd[q] += v[q].x * v[q].x + v[q].y * v[q].y;
                        }
                    }
                } else {
                    // Mixed precision RK stage: H, the stage state and the K
                    // accumulator in FP32; only stage 4 adds its FP32 increment
                    // c (K + h) / 6 to the FP64 state.
                    const float f4s = (float)g4s, f4o = (float)g4o;
                    const float fcr = (float)a.cr, fci = (float)a.ci;
#pragma unroll
                    for (int q = 0; q < E; ++q) {
                        const long long g = row0 + elemA<N, V>(t, q);
                        float2 ps;
                        if (a.stage == 1) {
                            const double2 r = ldn<NT>(P0[c], g);
                            ps = float2{(float)(sig * r.x), (float)(sig * r.y)};
                        } else {
                            ps = ldn<NT>(PS[c], g);
                        }
                        const float n = ps.x * ps.x + ps.y * ps.y;
                        float no = 0.f;
                        if (NC == 2) {
                            if (c == 0) {
                                float2 o;
                                if (a.stage == 1) {
                                    const double2 r = ldn<NT>(P0[1], g);
                                    o = float2{(float)(sig1 * r.x), (float)(sig1 * r.y)};
                                } else {
                                    o = ldn<NT>(PS[1], g);
                                }
                                no = o.x * o.x + o.y * o.y;
                                nsave[q] = n;
                            } else {
                                no = (float)nsave[q];
                            }
                        }
                        const float vr = (float)ldn<NT>(Vr[c], g);
                        const float vi = Vi[c] ? (float)ldn<NT>(Vi[c], g) : 0.f;
                        const float2 ke = v[q];
                        float rX = vr + f4s * n;
                        if (NC == 2) rX = rX + f4o * no;
                        if (DENS) rX += (float)phi_at(q);
                        float2 h = float2{ke.x + rX * ps.x - vi * ps.y, ke.y + rX * ps.y + vi * ps.x};
                        if constexpr (UTerm::enabled) {
                            const UserTermOut w = UTerm::eval(user_in<N>(a, g, c, ps.x, ps.y, n, no));
                            h = float2{h.x + (float)(w.wr * ps.x - w.wi * ps.y), h.y + (float)(w.wr * ps.y + w.wi * ps.x)};
                        }
                        if (a.stage < 4) {
                            float2 po = ps;   // stage 1: the (rounded) state itself
                            float2 k = float2{0.f, 0.f};
                            if (a.stage != 1) k = ldn<NT>(K[c], g);
                            if (a.stage == 2) {
                                // The state from the stage-1 output and its increment: PS = psi_n + c h1
                                // and K = h1, and stage 2 has the same c, so the FP64 state is not read
                                // (it would only be rounded to FP32 here).
                                po = float2{ps.x - (fcr * k.x - fci * k.y), ps.y - (fcr * k.y + fci * k.x)};
                            } else if (a.stage == 3) {
                                if (PS1[c]) {
                                    // With PS1 = psi_n + c2 h1, PS = psi_n + c2 h2, K = h1 + 2 h2 and
                                    // c2 = c3 / 2: psi_n = (PS1 + 2 PS - c2 K) / 3, again without
                                    // reading the FP64 state.  The division by 3 must not be biased:
                                    // the product with the FP32 value of 1/3 alone (3e-8 high) would
                                    // scale every stage-3 state by that bias, which the steps
                                    // accumulate (2e-6 in energy after 2000 steps at 256^3, against
                                    // 3e-8), so the product is corrected by its exact residual
                                    // r = s - 3 q (one FMA), which leaves the quotient within about
                                    // half an ulp of s / 3.
                                    const float2 p1 = ldn<NT>(PS1[c], g);
                                    const float hr = 0.5f * fcr, hi = 0.5f * fci;
                                    po = float2{div3(p1.x + 2.f * ps.x - (hr * k.x - hi * k.y)),
                                                div3(p1.y + 2.f * ps.y - (hr * k.y + hi * k.x))};
                                } else {
                                    const double2 r = ldn<NT>(P0[c], g);
                                    po = float2{(float)(sig * r.x), (float)(sig * r.y)};
                                }
                            }
                            const float2 nw = float2{po.x + (fcr * h.x - fci * h.y), po.y + (fcr * h.y + fci * h.x)};
                            if (a.stage == 1) stn<NT>(K[c], g, h);
                            else stn<NT>(K[c], g, float2{k.x + 2.f * h.x, k.y + 2.f * h.y});
                            stn<NT>(PSo[c] ? PSo[c] : PS[c], g, nw);
                            v[q] = nw;
                        } else {
                            const float2 k = ldn<NT>(K[c], g);
                            const float ax = k.x + h.x, ay = k.y + h.y;
                            const float dx = (fcr * ax - fci * ay) / 6.f, dy = (fcr * ay + fci * ax) / 6.f;
                            const double2 r = ldn<NT>(P0[c], g);
                            const double2 nw = double2{sig * r.x + (double)dx, sig * r.y + (double)dy};
                            stn<NT>(P0[c], g, nw);
                            v[q] = from_d2<T>(nw);
                            if (a.want_norm) sums[3 * c] += nw.x * nw.x + nw.y * nw.y;
                        }
                        if (DENS) nd[q] += v[q].x * v[q].x + v[q].y * v[q].y;
                    }
                }
                if (MODE != MODE_APPLY) {
                    fft_forward<T2, N, V, true>(v, t, my, tw);
#pragma unroll
                    for (int q = 0; q < E; ++q) sc[elemB<N, V>(t, q)] = v[q];
                }
            }
        }
        if (DENS && MODE != MODE_APPLY) {
#pragma unroll
            for (int q = 0; q < E; ++q) v[q] = T2{nd[q], (T)0};
            fft_forward<T2, N, V, true>(v, t, my, tw);
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int e = elemB<N, V>(t, q);
                if (e <= N / 2) specd[e] = v[q];
            }
        }
    }
    if (MODE == MODE_APPLY || MODE == MODE_HB || (MODE == MODE_RK && a.want_norm))
        block_sum_store<NV>(sums, glob(a.partials) + (long long)NV * bx);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the RK4 step, one or two components.  With g(psi) =
// H(psi) psi, H_c = T + V_c + sum_c' g4[c][c'] n_c' + Phi[n] (real V,
// n = n_1 (+ n_2)), the transposed Jacobian (real inner product Re<u, v>)
// acting on abar is, per component,
//     (G^T abar)_c = H_c abar_c + 2 psi_c (sum_c' g4[c][c'] r_c' + Phi[r]),
//     r_c = Re(conj(psi_c) abar_c),  r = r_1 (+ r_2)
// (g4 symmetric), and the parameter derivatives contribute dV_c += r_c,
// d(4 pi a_cc) += n_c r_c, d(4 pi a_12) += n_1 r_2 + n_2 r_1, and
// d(strength) += Phi1[n] r for the k-space kernel (Phi1 = Phi / strength: the
// kernel is linear in a_dd, or in the soft-core U).  One VJP is the prologue
// below, the spectral passes on NC + 2 channels (abar_c: kinetic; n and r:
// the k-space kernel, both real fields kept as half spectra like the
// solver's density), and the epilogue.  With both kernels on (`both`), n is
// also written to a third half channel that the passes multiply by the
// dipolar kernel alone: Phi_dd[n] r separates the two strengths' sums.
// Pencils past the last row (grids with fewer rows than a block's pencils,
// e.g. 1D) do no work: the x exchanges are wave-scoped.

struct AdjArgs {
    void* spec;             // channels of T2: abar_c (c < NC), then n, r
    long long ntot;
    long long hchan;        // n and r: half spectra (kx = 0..N/2), n at spec + NC ntot,
    int hpitch;             // r hchan elements on (both: n again 2 hchan on); row r at r * hpitch
    int both;               // dipolar and soft-core kernels on together
    const void* y[2];       // state components the Jacobian is taken at (YT2)
    const double2* src[2];  // prologue: abar_c = alpha * src_c (null: abar read from `abar`)
    double2 alpha;
    double2* abar[2];       // abar of this VJP (written by the prologue when src is set)
    const void* Vr[2];      // Re V_c (T)
    double g4[2][2];        // 4 pi a_cc'
    // epilogue
    const double2* lam_in[2];   // adjoint of the step output
    double2* lam[2];            // accumulator: lam = lam_in + G (init) or lam += G
    int init;
    double2* gb_out[2];         // next abar = ga * lam_in + gg * G (null: none; may be abar itself)
    double2 ga, gg;
    double* vbar[2];            // += r_c
    double* partials;           // per block, NC = 1: (n r, Phi[n] r, Phi_dd[n] r); NC = 2: (n1 r1, n2 r2,
                                // n1 r2 + n2 r1, Phi[n] r, Phi_dd[n] r); Phi_dd[n] r is 0 unless both
};

__device__ __forceinline__ double2 zmul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

template <typename T, typename YT, int N, int V, bool DENS, int NC>
__global__ void __launch_bounds__((px_of<T, N, V>() * Plan<N, V>::T))
k_adj_pro(AdjArgs a, const c2_t<T>* __restrict__ twg)
{
    using T2 = c2_t<T>;
    using Y2 = c2_t<YT>;
    constexpr int TT = Plan<N, V>::T, E = Plan<N, V>::E, P = px_of<T, N, V>();
    __shared__ T2 lds[P * lds_stride<N>()];
    __shared__ T2 twl[N];
    const lds_table<T2> tw = stage_twiddles<T2, N>(twl, twg);
    const int t = threadIdx.x % TT, p = threadIdx.x / TT;
    const long long row = (long long)blockIdx.x * P + p;
    if (row >= a.ntot / N) return;
    const long long row0 = row * N;
    T2* my = lds + p * lds_stride<N>();
    const gp<T2> sp = glob((T2*)a.spec);
    T2 v[E];
    double nn[E], rr[E];
#pragma unroll
    for (int q = 0; q < E; ++q) nn[q] = rr[q] = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const gp<const Y2> y = glob((const Y2*)a.y[c]);
        const gp<double2> abar = glob(a.abar[c]);
        const gp<const double2> src = glob(a.src[c]);
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const long long g = row0 + elemA<N, V>(t, q);
            double2 ab;
            if (a.src[0]) {
                ab = zmul(a.alpha, src[g]);
                abar[g] = ab;
            } else {
                ab = abar[g];
            }
            const Y2 yr = y[g];
            const double yx = (double)yr.x, yy = (double)yr.y;
            nn[q] += yx * yx + yy * yy;
            rr[q] += yx * ab.x + yy * ab.y;
            v[q] = T2{(T)ab.x, (T)ab.y};
        }
        fft_forward<T2, N, V, true>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) sp[c * a.ntot + row0 + elemB<N, V>(t, q)] = v[q];
    }
    if (DENS) {
        const gp<T2> hn = sp + NC * a.ntot + row * a.hpitch, hr = hn + a.hchan;
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = T2{(T)nn[q], (T)0};
        fft_forward<T2, N, V, true>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int e = elemB<N, V>(t, q);
            if (e <= N / 2) hn[e] = v[q];
        }
        if (a.both) {
            const gp<T2> hd = hn + 2 * a.hchan;
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const int e = elemB<N, V>(t, q);
                if (e <= N / 2) hd[e] = v[q];
            }
        }
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = T2{(T)rr[q], (T)0};
        fft_forward<T2, N, V, true>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int e = elemB<N, V>(t, q);
            if (e <= N / 2) hr[e] = v[q];
        }
    }
}

template <typename T, typename YT, int N, int V, bool DENS, int NC>
__global__ void __launch_bounds__((px_of<T, N, V>() * Plan<N, V>::T))
k_adj_epi(AdjArgs a, const c2_t<T>* __restrict__ twg)
{
    using T2 = c2_t<T>;
    using Y2 = c2_t<YT>;
    constexpr int TT = Plan<N, V>::T, E = Plan<N, V>::E, P = px_of<T, N, V>();
    constexpr int NS = NC == 1 ? 3 : 5;
    __shared__ T2 lds[P * lds_stride<N>()];
    __shared__ T2 twl[N];
    const lds_table<T2> tw = stage_twiddles<T2, N>(twl, twg);
    const int t = threadIdx.x % TT, p = threadIdx.x / TT;
    const long long rowi = (long long)blockIdx.x * P + p;
    const bool pv = rowi < a.ntot / N;
    const long long row = pv ? rowi : 0;
    const long long row0 = row * N;
    T2* my = lds + p * lds_stride<N>();
    const gp<T2> sp = glob((T2*)a.spec);
    T2 v[E];
    double phin[E], phir[E];
    double s[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = 0.0;
    // the half spectra of a real field: kx > N/2 from the conjugate of N - kx
    auto load_half = [&](gp<const T2> h) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int e = elemB<N, V>(t, q);
            if (e <= N / 2) {
                v[q] = h[e];
            } else {
                const T2 w = h[N - e];
                v[q] = T2{w.x, -w.y};
            }
        }
    };
    if (DENS && pv) {
        const gp<const T2> hn = sp + NC * a.ntot + row * a.hpitch;
        load_half(hn);
        fft_inverse<T2, N, V, true>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) phin[q] = (double)v[q].x;
        load_half(hn + a.hchan);
        fft_inverse<T2, N, V, true>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) phir[q] = (double)v[q].x;
    }
    // NC == 2: component 0's n and r, kept for component 1's coupling terms
    // (gb_out may alias abar, so component 0's abar is overwritten by then)
    double nsv[E], rsv[E];
    double rt[E];   // r summed over the components
#pragma unroll
    for (int q = 0; q < E; ++q) rt[q] = 0.0;
    if (pv) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const gp<const Y2> y = glob((const Y2*)a.y[c]);
            const gp<const Y2> yo = glob((const Y2*)a.y[1]);
            const gp<const double2> abar = glob((const double2*)a.abar[c]);
            const gp<const double2> abo = glob((const double2*)a.abar[1]);
            const gp<const T> Vr = glob((const T*)a.Vr[c]);
            const gp<double2> lam = glob(a.lam[c]), gb_out = glob(a.gb_out[c]);
            const gp<const double2> lam_in = glob(a.lam_in[c]);
            const gp<double> vbar = glob(a.vbar[c]);
            const double gs = a.g4[c][c], go = a.g4[c][1 - c];
#pragma unroll
            for (int q = 0; q < E; ++q) v[q] = sp[c * a.ntot + row0 + elemB<N, V>(t, q)];
            fft_inverse<T2, N, V, true>(v, t, my, tw);   // T abar_c
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const long long g = row0 + elemA<N, V>(t, q);
                const Y2 yr = y[g];
                const double2 yv = double2{(double)yr.x, (double)yr.y};
                const double2 ab = abar[g];
                const double n = yv.x * yv.x + yv.y * yv.y;
                const double r = yv.x * ab.x + yv.y * ab.y;
                double no = 0.0, ro = 0.0;   // the other component's density and r
                if (NC == 2) {
                    if (c == 0) {
                        const Y2 yq = yo[g];
                        const double2 yw = double2{(double)yq.x, (double)yq.y};
                        const double2 aq = abo[g];
                        no = yw.x * yw.x + yw.y * yw.y;
                        ro = yw.x * aq.x + yw.y * aq.y;
                        nsv[q] = n;
                        rsv[q] = r;
                    } else {
                        no = nsv[q];
                        ro = rsv[q];
                    }
                }
                const double pn = DENS ? phin[q] : 0.0;
                const double pr = DENS ? phir[q] : 0.0;
                const double w = (double)Vr[g] + gs * n + go * no + pn;
                const double c2 = 2.0 * (gs * r + go * ro + pr);
                const double2 G = double2{(double)v[q].x + w * ab.x + c2 * yv.x, (double)v[q].y + w * ab.y + c2 * yv.y};
                const double2 li = lam_in[g];
                const double2 base = a.init ? li : lam[g];
                lam[g] = double2{base.x + G.x, base.y + G.y};
                if (a.gb_out[0]) {
                    const double2 u = zmul(a.ga, li), wv = zmul(a.gg, G);
                    gb_out[g] = double2{u.x + wv.x, u.y + wv.y};
                }
                vbar[g] += r;
                rt[q] += r;
                if constexpr (NC == 1) {
                    s[0] += n * r;
                    s[1] += pn * r;
                } else {
                    s[c] += n * r;           // a_cc
                    if (c == 0) s[2] += n * ro + no * r;   // a_12: n1 r2 + n2 r1, once
                    s[3] += pn * r;          // the kernel: Phi[n] (r1 + r2)
                }
            }
        }
    }
    if (DENS && pv && a.both) {
        load_half(sp + NC * a.ntot + row * a.hpitch + 2 * a.hchan);
        fft_inverse<T2, N, V, true>(v, t, my, tw);   // Phi_dd[n]
#pragma unroll
        for (int q = 0; q < E; ++q) s[NS - 1] += (double)v[q].x * rt[q];
    }
    block_sum_store<NS>(s, glob(a.partials) + NS * (long long)blockIdx.x);
}

// ---------------------------------------------------------------------------
// Final reductions over the per-block partial sums (3 per component), summed
// in a fixed order by one block.
//   FIN_OBS:    out[3c..3c+2] = norm_c = S0 dv, S1 dv / norm_c, S2 dv / norm_c
//   FIN_RENORM: scale[c] = sqrt(npart_c / (S0 dv)); out[3c] = S0 dv
//   FIN_HB:     the bookkeeping of one heavy-ball iteration
//               (src/oneComponentGPSolver.cu:68-88, src/twoComponentGPSolver.cu
//               :59-94): norm_c, scale_c = sqrt(npart_c/norm_c),
//               mu_c, E_c = scale_c S1,2 dv / normarg_c (normarg = npart_c for
//               one component, the unscaled norm_c for two), out <- (norm, mu,
//               E) per component, convergence |dE|/E_prev < eps (one
//               component) or |dE_c| < eps for both (two) sets *stop.
enum { FIN_OBS = 0, FIN_RENORM = 1, FIN_HB = 2 };
struct FinArgs {
    const double* partials;
    int nblocks, nc, kind;
    double dv;
    double npart[2];
    double* out;
    double* scale;
    double* mu;
    double* eprev;
    int* stop;
    int* done;
    double eps;
};

// one block of kFinThreads threads per instance (blockIdx.x indexes tab when it is non-null)
constexpr int kFinThreads = 1024;
__global__ void k_finalize(FinArgs f0, const FinArgs* __restrict__ tab);   // defined in fgpe_capi.hip

template <typename TD, typename TS>
__global__ void k_convert(TD* __restrict__ dst, const TS* __restrict__ src, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (TD)src[i];
}

// soft-core Vt(k) table (src/parallel.cu:371-477), spatial-index order
template <typename T>
__global__ void k_softcore_table(T* __restrict__ out, const double* kx, const double* ky, const double* kz,
                                 int nx, int ny, long long n, int dim, double R, double U)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int ix = (int)(idx % nx);
    const int iy = (int)((idx / nx) % ny);
    const long long iz = idx / ((long long)nx * ny);
    const double a = R;
    const double k = 2 * kPi * sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy] + kz[iz] * kz[iz]);
    double v;
    if (dim == 1) {
        v = (k < 1E-6) ? U * 2 * a * (1 - (a * k) * (a * k) / 6.0) : U * 2 * sin(a * k) / (k);
    } else if (dim == 2) {
        v = (k < 1E-6) ? U * kPi * a * a : 2 * kPi * a * U * j1(k * a) / k;
    } else {
        v = (k < 1E-6) ? U * (4. / 3.) * kPi * pow(a, 3)
                       : 4 * kPi * U * pow(a, 3) * (sin(k * a) - k * a * cos(k * a)) / (pow(k * a, 3));
    }
    out[idx] = (T)v;
}

}  // namespace fgpe
