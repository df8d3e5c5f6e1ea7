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

#include "fft_core.hpp"

namespace fgpe {

constexpr double kPi = 3.141592653589793;  // hpp/header.hpp:28

// pencils per block of the x kernels and tile width of the strided kernels
template <typename T, int N> __host__ __device__ constexpr int px_of()
{
    return Plan<N>::T >= 128 ? 1 : 128 / Plan<N>::T;
}
template <typename T, int N> __host__ __device__ constexpr int ps_of()
{
    // 8 consecutive x columns (128 B of FP64 complex) per row segment; 16 for
    // FP32; fewer for the long pencils so that the LDS tile stays <= 64 KB
    return (sizeof(T) == 8) ? (N >= 512 ? 4 : 8) : (N >= 512 ? 8 : 16);
}

template <typename T> __device__ __forceinline__ double2 to_d2(c2_t<T> a) { return double2{(double)a.x, (double)a.y}; }
template <typename T> __device__ __forceinline__ c2_t<T> from_d2(double2 a) { return c2_t<T>{(T)a.x, (T)a.y}; }

// ---------------------------------------------------------------------------
// Block reduction of NV doubles per thread into out[0..NV): wave64 shuffles,
// then one LDS slot per wave, summed in wave order (deterministic).
template <int NV> __device__ __forceinline__ void block_sum_store(double (&s)[NV], double* out)
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
template <typename T, int N, int DIR>
__global__ void __launch_bounds__((px_of<T, N>() * Plan<N>::T))
k_fft_x(const c2_t<T>* in, c2_t<T>* out, const c2_t<T>* __restrict__ tw)
{
    using T2 = c2_t<T>;
    constexpr int TT = Plan<N>::T, E = Plan<N>::E, P = px_of<T, N>();
    __shared__ T2 lds[P * lds_stride<N>()];
    const int t = threadIdx.x % TT, p = threadIdx.x / TT;
    const long long row = (long long)blockIdx.x * P + p;
    const T2* src = in + row * N;
    T2* dst = out + row * N;
    T2* my = lds + p * lds_stride<N>();
    T2 v[E];
    if (DIR < 0) {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = src[elemA<N>(t, q)];
        fft_forward<T2, N>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) dst[elemB<N>(t, q)] = v[q];
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = src[elemB<N>(t, q)];
        fft_inverse<T2, N>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) dst[elemA<N>(t, q)] = v[q];
    }
}

// ---------------------------------------------------------------------------
// Strided axis (y or z) transform, in place.  Block = tile of P consecutive x
// columns at one `outer` index; channel = blockIdx.y.  Thread (p, t) with p
// fastest, so a row segment of P columns is one coalesced access.
// scale (forward only, may be null): channel c < ncomp is multiplied by
// scale[0], the density channel by scale[0]^2 (the pending renormalisation
// of the imaginary-time schemes).
struct StridedArgs {
    int nx;                 // columns (x extent)
    long long stride;       // element stride along the axis
    long long outer_stride; // stride of the outer index
    long long chan_stride;  // elements per channel
    const double* scale;
    int ncomp;
};

template <typename T, int N, int DIR>
__global__ void __launch_bounds__((ps_of<T, N>() * Plan<N>::T))
k_strided(c2_t<T>* __restrict__ data, const c2_t<T>* __restrict__ tw, StridedArgs a)
{
    using T2 = c2_t<T>;
    constexpr int TT = Plan<N>::T, E = Plan<N>::E, P = ps_of<T, N>();
    __shared__ T2 lds[P * lds_stride<N>()];
    const int p = threadIdx.x % P, t = threadIdx.x / P;
    const int ntile = a.nx / P;
    const int xt = blockIdx.x % ntile;
    const long long outer = blockIdx.x / ntile;
    const int ch = blockIdx.y;
    T2* base = data + ch * a.chan_stride + outer * a.outer_stride + (long long)xt * P + p;
    T2* my = lds + p * lds_stride<N>();
    T2 v[E];
    if (DIR < 0) {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = base[(long long)elemA<N>(t, q) * a.stride];
        if (a.scale) {
            double s = a.scale[0];
            if (ch >= a.ncomp) s = s * s;
            const T st = (T)s;
#pragma unroll
            for (int q = 0; q < E; ++q) v[q] = T2{v[q].x * st, v[q].y * st};
        }
        fft_forward<T2, N>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) base[(long long)elemB<N>(t, q) * a.stride] = v[q];
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = base[(long long)elemB<N>(t, q) * a.stride];
        fft_inverse<T2, N>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) base[(long long)elemA<N>(t, q) * a.stride] = v[q];
    }
}

// ---------------------------------------------------------------------------
// Last axis: forward FFT, k-space multiply, inverse FFT, in place.
// Channel c < ncomp (psi):  v *= (19.739208802/Ntot) * |k|^2   (src/parallel.cu:194-205)
// Channel c = ncomp (density): v *= Vt(k) / Ntot, Vt = dipolar (inline,
//   src/parallel.cu:326-352) + table (soft-core or user kernel, may be null).
struct LastArgs {
    StridedArgs s;
    const double* kA;   // k along this axis
    const double* kX;   // k along x (column index)
    const double* kO;   // k along the outer axis
    double dA, dX, dO;  // dipole direction components along A, X, O
    double lap_scale;   // 19.739208802 / Ntot
    double inv_ntot;    // 1 / Ntot
    double c_dd;        // 12 pi a_dd (0: no dipolar term)
    const void* wtab;   // T table of Vt (spatial-index order) or null
};

template <typename T, int N>
__global__ void __launch_bounds__((ps_of<T, N>() * Plan<N>::T))
k_last(c2_t<T>* __restrict__ data, const c2_t<T>* __restrict__ tw, LastArgs a)
{
    using T2 = c2_t<T>;
    constexpr int TT = Plan<N>::T, E = Plan<N>::E, P = ps_of<T, N>();
    __shared__ T2 lds[P * lds_stride<N>()];
    const int p = threadIdx.x % P, t = threadIdx.x / P;
    const int ntile = a.s.nx / P;
    const int xt = blockIdx.x % ntile;
    const long long outer = blockIdx.x / ntile;
    const int ch = blockIdx.y;
    const int ix = xt * P + p;
    const long long sp_off = outer * a.s.outer_stride + ix;   // spatial offset
    T2* base = data + ch * a.s.chan_stride + sp_off;
    T2* my = lds + p * lds_stride<N>();
    T2 v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = base[(long long)elemA<N>(t, q) * a.s.stride];
    if (a.s.scale) {
        double s = a.s.scale[0];
        if (ch >= a.s.ncomp) s = s * s;
        const T st = (T)s;
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = T2{v[q].x * st, v[q].y * st};
    }
    fft_forward<T2, N>(v, t, my, tw);

    const double kx = a.kX[ix], ko = a.kO[outer];
    const double kxo2 = kx * kx + ko * ko;
    if (ch < a.s.ncomp) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const double ka = a.kA[elemB<N>(t, q)];
            const double k2 = kxo2 + ka * ka;
            if (sizeof(T) == 8) {
                v[q] = T2{(T)((a.lap_scale * (double)v[q].x) * k2), (T)((a.lap_scale * (double)v[q].y) * k2)};
            } else {
                const T f = (T)(a.lap_scale * k2);
                v[q] = T2{v[q].x * f, v[q].y * f};
            }
        }
    } else {
        const T* wt = (const T*)a.wtab;
        const double num0 = kx * a.dX + ko * a.dO;
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const int ea = elemB<N>(t, q);
            const double ka = a.kA[ea];
            double vt = 0.0;
            if (a.c_dd != 0.0) {
                const double num = num0 + ka * a.dA;
                const double den = sqrt(kxo2 + ka * ka);
                if (!(den < 1E-6)) vt = a.c_dd * (((num / den) * (num / den)) - 1.0 / 3.0);
            }
            if (wt) vt = vt + (double)wt[sp_off + (long long)ea * a.s.stride];
            const double f = vt * a.inv_ntot;
            if (sizeof(T) == 8) {
                v[q] = T2{(T)((double)v[q].x * f), (T)((double)v[q].y * f)};
            } else {
                const T ff = (T)f;
                v[q] = T2{v[q].x * ff, v[q].y * ff};
            }
        }
    }
    fft_inverse<T2, N>(v, t, my, tw);
#pragma unroll
    for (int q = 0; q < E; ++q) base[(long long)elemA<N>(t, q) * a.s.stride] = v[q];
}

// ---------------------------------------------------------------------------
// x-axis stage kernel (one component).
//   MODE_PROLOGUE: spec <- Fx(psi) [, Fx(|psi|^2)] of the stored (raw) state;
//                  a pending scale is applied by the next forward pass
//   MODE_RK:       inverse x of spec -> KE [, Phi]; H at every point; RK4
//                  stage `stage` (src/oneComponentGPSolver.cu:165-187:
//                  UpdateRKStep / FinalRKStep, src/parallel.cu:230-271);
//                  spec <- Fx(next state) [, Fx(its density)]
//   MODE_APPLY:    inverse x -> H and Hen (BECHamiltonian,
//                  hamiltonians/BEC_oneComponent.cu:128-150) to hmu/hen (if
//                  non-null) and the sums |psi|^2, Re(psi* H), Re(psi* Hen).
// H = KE + (V.x + g4 n + Phi) psi + i V.y psi, Hen with g2 and Phi/2.
enum { MODE_PROLOGUE = 0, MODE_RK = 1, MODE_APPLY = 2 };

template <typename T> struct XArgs {
    c2_t<T>* spec;
    long long ntot;
    double2* P0;            // FP64 state psi_n (psi_old of the step)
    c2_t<T>* PS;            // stage state
    c2_t<T>* K;             // RK accumulator
    const T* Vr;            // Re V
    const T* Vi;            // Im V or null
    double g4, g2;          // 4 pi a(t), 2 pi a(t)
    double cr, ci;          // stage coefficient c
    int stage;              // 1..4
    const double* scale;    // pending renormalisation of P0 (null: 1)
    double* partials;       // 3 doubles per block (MODE_APPLY, or want_norm)
    int want_norm;          // MODE_RK stage 4: write sum |psi_new|^2 per block
    double2* hmu;
    double2* hen;
};

template <typename T, int N, bool DENS, int MODE>
__global__ void __launch_bounds__((px_of<T, N>() * Plan<N>::T))
k_xstage(XArgs<T> a, const c2_t<T>* __restrict__ tw)
{
    using T2 = c2_t<T>;
    constexpr int TT = Plan<N>::T, E = Plan<N>::E, P = px_of<T, N>();
    __shared__ T2 lds[P * lds_stride<N>()];
    const int t = threadIdx.x % TT, p = threadIdx.x / TT;
    const long long row0 = ((long long)blockIdx.x * P + p) * N;
    T2* my = lds + p * lds_stride<N>();
    T2* spec0 = a.spec + row0;
    T2* spec1 = a.spec + a.ntot + row0;
    const double sig = a.scale ? a.scale[0] : 1.0;

    T2 v[E];
    double nd[E];        // density of the new state (DENS)

    if (MODE == MODE_PROLOGUE) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            v[q] = from_d2<T>(a.P0[row0 + elemA<N>(t, q)]);
            if (DENS) { const T2 w = v[q]; nd[q] = (double)(w.x * w.x + w.y * w.y); }
        }
    } else {
        double phi[E];
        if (DENS) {
#pragma unroll
            for (int q = 0; q < E; ++q) v[q] = spec1[elemB<N>(t, q)];
            fft_inverse<T2, N>(v, t, my, tw);
#pragma unroll
            for (int q = 0; q < E; ++q) phi[q] = (double)v[q].x;
        }
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = spec0[elemB<N>(t, q)];
        fft_inverse<T2, N>(v, t, my, tw);   // v = KE psi_s (layout A)

        double sums[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const long long g = row0 + elemA<N>(t, q);
            // stage state psi_s
            double2 po = double2{0.0, 0.0};
            double2 ps;
            if (MODE == MODE_APPLY || a.stage == 1) {
                const double2 r = a.P0[g];
                po = double2{sig * r.x, sig * r.y};
                ps = po;
            } else {
                ps = to_d2<T>(a.PS[g]);
            }
            const double vr = (double)a.Vr[g];
            const double vi = a.Vi ? (double)a.Vi[g] : 0.0;
            const double n = ps.x * ps.x + ps.y * ps.y;
            const double2 ke = to_d2<T>(v[q]);
            const double ph = DENS ? phi[q] : 0.0;
            double rX = (vr + a.g4 * n);
            if (DENS) rX += ph;
            const double2 h = double2{ke.x + rX * ps.x - vi * ps.y, ke.y + rX * ps.y + vi * ps.x};
            if (MODE == MODE_APPLY) {
                double eX = (vr + a.g2 * n);
                if (DENS) eX += 0.5 * ph;
                const double2 he = double2{ke.x + eX * ps.x - vi * ps.y, ke.y + eX * ps.y + vi * ps.x};
                if (a.hmu) a.hmu[g] = h;
                if (a.hen) a.hen[g] = he;
                sums[0] += n;
                sums[1] += ps.x * h.x + ps.y * h.y;
                sums[2] += ps.x * he.x + ps.y * he.y;
            } else {
                if (a.stage != 1) {
                    const double2 r = a.P0[g];
                    po = double2{sig * r.x, sig * r.y};
                }
                double2 nw;
                if (a.stage < 4) {
                    // UpdateRKStep: psi = old + c*h ; k = h (stage 1) or k + 2h
                    nw = double2{po.x + (a.cr * h.x - a.ci * h.y), po.y + (a.cr * h.y + a.ci * h.x)};
                    if (a.stage == 1) {
                        a.K[g] = from_d2<T>(h);
                    } else {
                        const double2 k = to_d2<T>(a.K[g]);
                        a.K[g] = from_d2<T>(double2{k.x + 2.0 * h.x, k.y + 2.0 * h.y});
                    }
                    const T2 st = from_d2<T>(nw);
                    a.PS[g] = st;
                    v[q] = st;
                } else {
                    // FinalRKStep: psi = old + c*(k + h)/6
                    const double2 k = to_d2<T>(a.K[g]);
                    const double aX = k.x + h.x, aY = k.y + h.y;
                    nw = double2{po.x + (a.cr * aX - a.ci * aY) / 6.0, po.y + (a.cr * aY + a.ci * aX) / 6.0};
                    a.P0[g] = nw;
                    v[q] = from_d2<T>(nw);
                    if (a.want_norm) sums[0] += nw.x * nw.x + nw.y * nw.y;
                }
                if (DENS) { const T2 w = v[q]; nd[q] = (double)(w.x * w.x + w.y * w.y); }
            }
        }
        if (MODE == MODE_APPLY) {
            block_sum_store<3>(sums, a.partials + 3 * blockIdx.x);
            return;
        }
        if (a.want_norm) {
            double s1[1] = {sums[0]};
            block_sum_store<1>(s1, a.partials + blockIdx.x);
        }
    }
    // forward x-FFT of the next state and of its density
    fft_forward<T2, N>(v, t, my, tw);
#pragma unroll
    for (int q = 0; q < E; ++q) spec0[elemB<N>(t, q)] = v[q];
    if (DENS) {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = T2{(T)nd[q], (T)0};
        fft_forward<T2, N>(v, t, my, tw);
#pragma unroll
        for (int q = 0; q < E; ++q) spec1[elemB<N>(t, q)] = v[q];
    }
}

// ---------------------------------------------------------------------------
// Final reductions over the per-block partial sums, in a fixed order.
//   kind 0 (observables): out[0] = S0 dv, out[1] = S1 dv / out[0], out[2] = S2 dv / out[0]
//   kind 1 (renormalise): scale[0] = sqrt(npart / (S0 dv)), out[0] = S0 dv
enum { FIN_OBS = 0, FIN_RENORM = 1 };
__global__ void __launch_bounds__(256)
k_finalize(const double* __restrict__ partials, int nblocks, int nv, int kind, double dv,
           double npart, double* out, double* scale)
{
    __shared__ double sh[3][256];
    const int t = threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = t; b < nblocks; b += 256)
        for (int q = 0; q < nv; ++q) s[q] += partials[(long long)b * nv + q];
    for (int q = 0; q < 3; ++q) sh[q][t] = s[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 3; ++q) sh[q][t] += sh[q][t + w];
        __syncthreads();
    }
    if (t == 0) {
        const double norm = sh[0][0] * dv;
        if (kind == FIN_OBS) {
            out[0] = norm;
            out[1] = sh[1][0] * dv / norm;
            out[2] = sh[2][0] * dv / norm;
        } else {
            out[0] = norm;
            scale[0] = sqrt(npart / norm);
        }
    }
}

// psi <- scale * psi (applies a pending renormalisation), then scale <- 1
__global__ void k_apply_scale(double2* __restrict__ psi, long long n, double* scale)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const double s = scale[0];
    if (i < n) psi[i] = double2{s * psi[i].x, s * psi[i].y};
}
__global__ void k_set_scalar(double* p, double v) { p[0] = v; }

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
