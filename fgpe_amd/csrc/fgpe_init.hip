// fgpe_init.hip -- the initial state and the trap built on the device: the
// O(N) host loops of Psi (src/psi.cu:80-372) and Vext (src/vext.cu:60-199)
// as one pass over HBM each, written straight into the context's buffers.
//
// The reference evaluates these formulas on the host, where nothing fuses a
// multiply into an add.  The kernels keep its operation order and turn
// contraction off, so they reproduce its values up to the last-ulp
// differences of exp, cos, sin and tanh between the device and host math
// libraries.  Factors that depend on one axis only (the trap's per-axis
// terms, the vortex phase, the cosine seeds, the box profile, the Gaussian's
// squared coordinates) are evaluated once per axis index on the host, with
// the same operations and math library as the reference's loops, and the
// kernels combine them per point in the reference's order; what is left per
// point is the Gaussian's exp and the lattice profiles.  The one deliberate
// difference is the noise: the reference
// draws it from a sequential std::mt19937 stream, here every point draws its
// own Philox4x32-10 block (counter-based, Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), so the kernels need
// no sequential pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "ctx.hpp"

#pragma clang fp contract(off)

using namespace fgpe;

namespace {

constexpr int IB = 256;                     // threads per block
constexpr int IPT = 4;                      // points per thread
constexpr long long IPB = (long long)IB * IPT;

struct GeomD {
    double xmax[3], dx[3];
    int nx, ny;
    long long n;
};

struct Pt {
    long long i;
    int ix, iy, iz;
};

__device__ __forceinline__ Pt point(const GeomD& g, long long i)
{
    const long long r = i / g.nx;
    return Pt{i, (int)(i - r * g.nx), (int)(r % g.ny), (int)(r / g.ny)};
}

// -XMAX + i DX + DX/2, in the reference's order (src/psi.cu:134-150)
__device__ __forceinline__ double centre(const GeomD& g, int a, int i)
{
    return -g.xmax[a] + i * g.dx[a] + 0.5 * g.dx[a];
}

// Philox4x32-10: ten rounds of two 32x32 -> 64 multiplies, the key bumped
// by the Weyl constants between rounds.
struct U4 {
    uint32_t x, y, z, w;
};
__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
    }
    return c;
}

// 53 bits of (hi:lo) -> uniform on [-1, 1) (exact: r 2^-52 - 1)
__device__ __forceinline__ double unit_pm1(uint32_t lo, uint32_t hi)
{
    const uint64_t r = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)r * 0x1p-52 - 1.0;
}

// Noise pair (u1, u2) of point i in stream s.
__device__ __forceinline__ void noise_of(long long i, uint32_t s, uint32_t k0, uint32_t k1, double& u1, double& u2)
{
    const U4 o = philox4x32_10(U4{(uint32_t)i, (uint32_t)((unsigned long long)i >> 32), s, 0u}, k0, k1);
    u1 = unit_pm1(o.x, o.y);
    u2 = unit_pm1(o.z, o.w);
}

__device__ __forceinline__ void block_partial(double s, double* partials)
{
    double ss[1] = {s};
    block_sum_store<1>(ss, glob(partials) + blockIdx.x);
}

// psi = (1 + noise) * envelope (Psi::initHarmonic / initTube / initVoid,
// src/psi.cu:80-131); per-block sums of |psi|^2.  The complex product is
// the one of std::complex (__muldc3 order): (a + ib)(e + i0) = (a e - b 0,
// a 0 + b e).
template <int KIND>
__global__ void __launch_bounds__(IB) k_psi_init(double2* __restrict__ psi, GeomD g, const double* __restrict__ sx,
                                                 const double* __restrict__ sy, const double* __restrict__ sz,
                                                 double rr, double ci, uint32_t k0, uint32_t k1,
                                                 double* __restrict__ partials)
{
    double s = 0.0;
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= g.n) break;
        const Pt p = point(g, i);
        double u1, u2;
        noise_of(i, 0u, k0, k1, u1, u2);
        const double a = 1.0 + rr * u1, b = 0.0 + ci * u2;
        double re = a, im = b;
        if (KIND != FGPE_INIT_VOID) {
            // (y/ay)^2 + (z/az)^2 [+ (x/ax)^2], each square from the axis table
            double q = sy[p.iy] + sz[p.iz];
            if (KIND == FGPE_INIT_HARMONIC) q = q + sx[p.ix];
            const double e = exp(-0.5 * 0.1 * q);
            re = a * e - b * 0.0;
            im = a * 0.0 + b * e;
        }
        psi[i] = double2{re, im};
        s += re * re + im * im;
    }
    block_partial(s, partials);
}

// psi *= 1 + noise (Psi::addNoise, src/psi.cu:167-176), after the pending
// scale of this component
__global__ void __launch_bounds__(IB) k_psi_noise(double2* __restrict__ psi, long long n, double rr, double ci,
                                                  uint32_t stream, uint32_t k0, uint32_t k1,
                                                  const double* __restrict__ pend)
{
    const double sc = *pend;
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= n) break;
        double u1, u2;
        noise_of(i, stream, k0, k1, u1, u2);
        const double a = 1.0 + rr * u1, b = 0.0 + ci * u2;
        const double2 v = psi[i];
        const double c = sc * v.x, d = sc * v.y;
        psi[i] = double2{a * c - b * d, a * d + b * c};
    }
}

// psi *= exp(i ncirc 2 pi x / (2 xmax_x)) (Psi::imprintVortexTube,
// src/psi.cu:155-164); cs[ix] = (cos, sin) of the phase
__global__ void __launch_bounds__(IB) k_psi_vortex(double2* __restrict__ psi, GeomD g, const double2* __restrict__ cs,
                                                   const double* __restrict__ pend)
{
    const double sc = *pend;
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= g.n) break;
        const Pt p = point(g, i);
        const double2 w = cs[p.ix];
        const double2 v = psi[i];
        const double re = sc * v.x, im = sc * v.y;
        psi[i] = double2{w.x * re - w.y * im, w.x * im + w.y * re};
    }
}

// The lattice "force" profiles of src/psi.cu:179-347, in their operation
// order.  r3 = sqrt(3), r2 = sqrt(2) (correctly rounded, from the host).
template <int L>
__device__ __forceinline__ double force_of(double x, double y, double z, double amp, double k, double r3, double r2)
{
    if constexpr (L == FGPE_LATTICE_1D) {
        return amp * (1 - cos(k * x)) / 2.;
    } else if constexpr (L == FGPE_LATTICE_2D_TRIANGULAR) {
        return amp * (cos(k * x) + cos(k * (-0.5 * x + (r3 / 2) * y)) + cos(k * (-0.5 * x - (r3 / 2) * y))) / 3;
    } else if constexpr (L == FGPE_LATTICE_2D_SQUARE) {
        return amp * (cos(k * y) + cos(k * x)) / 2;
    } else if constexpr (L == FGPE_LATTICE_3D_HEX) {
        const double x1 = (-x + r3 * y) / 2., x4 = (-x - r3 * y) / 2., x2 = x, x3 = z;
        return amp * (cos(k * x1) + cos(k * x2) + cos(k * x3) + cos(k * x4)) / 4.0;
    } else if constexpr (L == FGPE_LATTICE_3D_FCC) {
        const double x1 = (x + y - z) / r3, x2 = (y + z - x) / r3, x3 = (z + x - y) / r3;
        const double x4 = (x - y - z) / r3, x5 = (y - z - x) / r3, x6 = (z - x - y) / r3;
        const double x7 = (x + y + z) / r3, x8 = (y + z + x) / r3, x9 = (z + x + y) / r3;
        return amp *
               (cos(k * x1) + cos(k * x2) + cos(k * x3) + cos(k * x4) + cos(k * x5) + cos(k * x6) + cos(k * x7) +
                cos(k * x8) + cos(k * x9)) /
               9.0;
    } else if constexpr (L == FGPE_LATTICE_3D_BCC) {
        const double x1 = (x + y) / r2, x2 = (z + y) / r2, x3 = (z + x) / r2;
        const double x4 = (x - y) / r2, x5 = (z - y) / r2, x6 = (z - x) / r2;
        return amp * (cos(k * x1) + cos(k * x2) + cos(k * x3) + cos(k * x4) + cos(k * x5) + cos(k * x6)) / 6;
    } else {
        return amp * (cos(k * z) + cos(k * y) + cos(k * x)) / 3;
    }
}

// psi = force * avg + psi (src/psi.cu:179-347; the complex + double adds to
// the real part)
template <int L>
__global__ void __launch_bounds__(IB) k_psi_force(double2* __restrict__ psi, GeomD g, double amp, double k, double ox,
                                                  double oy, double oz, double avg, double r3, double r2,
                                                  const double* __restrict__ pend)
{
    const double sc = *pend;
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= g.n) break;
        const Pt p = point(g, i);
        const double x = centre(g, 0, p.ix) - ox, y = centre(g, 1, p.iy) - oy, z = centre(g, 2, p.iz) - oz;
        const double f = force_of<L>(x, y, z, amp, k, r3, r2);
        const double2 v = psi[i];
        psi[i] = double2{f * avg + sc * v.x, sc * v.y};
    }
}

// per-block sums of |psi|^2 of the stored state times its pending scale
__global__ void __launch_bounds__(IB) k_psi_sumsq(const double2* __restrict__ psi, long long n,
                                                  const double* __restrict__ pend, double* __restrict__ partials)
{
    const double sc = *pend;
    double s = 0.0;
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= n) break;
        const double2 v = psi[i];
        const double re = sc * v.x, im = sc * v.y;
        s += re * re + im * im;
    }
    block_partial(s, partials);
}

// out[0] = sum of nb partials (one block, fixed order)
__global__ void __launch_bounds__(IB) k_sum_partials(const double* __restrict__ partials, long long nb,
                                                     double* __restrict__ out)
{
    double s = 0.0;
    for (long long b = threadIdx.x; b < nb; b += IB) s += partials[b];
    double ss[1] = {s};
    block_sum_store<1>(ss, glob(out));
}

// psi *= sqrt(npart / (norm dx dy dz)) with norm = *sum (Psi::normalize,
// src/psi.cu:351-372), then per-block sums of |psi|^2 dx dy dz
__global__ void __launch_bounds__(IB) k_psi_rescale(double2* __restrict__ psi, long long n,
                                                    const double* __restrict__ sum, double npart, double dx0,
                                                    double dx1, double dx2, const double* __restrict__ pend,
                                                    double* __restrict__ partials)
{
    const double sc = *pend;
    const double f = sqrt(npart / (*sum * dx0 * dx1 * dx2));
    double s = 0.0;
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= n) break;
        const double2 v = psi[i];
        const double re = (sc * v.x) * f, im = (sc * v.y) * f;
        psi[i] = double2{re, im};
        s += (re * re + im * im) * dx0 * dx1 * dx2;
    }
    block_partial(s, partials);
}

__global__ void k_set_one(double* p) { *p = 1.0; }

// Vext::initHarmonic / initTube / initVoid (src/vext.cu:60-118):
// 0.5 (tx + ty + tz) with t[i] = (x_i / ax)^2 / ax^2 per axis (no x term for
// the tube, 0 for void)
template <int KIND>
__global__ void __launch_bounds__(IB) k_vext_init(double* __restrict__ v, GeomD g, const double* __restrict__ tx,
                                                  const double* __restrict__ ty, const double* __restrict__ tz)
{
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= g.n) break;
        const Pt p = point(g, i);
        double r = 0;
        if (KIND == FGPE_INIT_HARMONIC) r = 0.5 * (tx[p.ix] + ty[p.iy] + tz[p.iz]);
        else if (KIND == FGPE_INIT_TUBE) r = 0.5 * (ty[p.iy] + tz[p.iz]);
        v[i] = r;
    }
}

// Vext::initBox (src/vext.cu:88-109): 0.5 V (bx by bz) with
// b[i] = tanh(B (x_i - w)) - tanh(B (x_i + w)) per axis
__global__ void __launch_bounds__(IB) k_vext_box(double* __restrict__ v, GeomD g, double V,
                                                 const double* __restrict__ bx, const double* __restrict__ by,
                                                 const double* __restrict__ bz)
{
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= g.n) break;
        const Pt p = point(g, i);
        v[i] = 0.5 * V * (bx[p.ix] * by[p.iy] * bz[p.iz]);
    }
}

// V += A cos(n 2 pi / L x), L = 2 xmax_x (src/vext.cu:122-179); cx[ix] is
// that whole term
__global__ void __launch_bounds__(IB) k_vext_cos(double* __restrict__ v, GeomD g, const double* __restrict__ cx)
{
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= g.n) break;
        const Pt p = point(g, i);
        v[i] += cx[p.ix];
    }
}

template <typename T>
__global__ void __launch_bounds__(IB) k_vext_store(const double* __restrict__ v, T* __restrict__ out, long long n)
{
    const long long b0 = (long long)blockIdx.x * IPB + threadIdx.x;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const long long i = b0 + (long long)j * IB;
        if (i >= n) break;
        out[i] = (T)v[i];
    }
}

unsigned nblocks(const fgpe_ctx* c) { return (unsigned)((c->ntot + IPB - 1) / IPB); }

GeomD geom_of(const fgpe_ctx* c, const fgpe_geom* g)
{
    GeomD d{};
    for (int a = 0; a < 3; ++a) {
        d.xmax[a] = g->xmax[a];
        d.dx[a] = g->dx[a];
    }
    d.nx = c->n[0];
    d.ny = c->n[1];
    d.n = c->ntot;
    return d;
}

int check_comp(fgpe_ctx* c, int comp, bool need_psi)
{
    if (!c) return fail(FGPE_EINVAL, "NULL context");
    if (comp < 0 || comp > 1) return fail(FGPE_EINVAL, "comp must be 0 or 1");
    if (need_psi && !c->P0[comp]) return fail(FGPE_ESTATE, "component has no state yet");
    return FGPE_OK;
}

// The state of comp changed outside the time steppers: its pending scale is
// folded in (the kernels multiply by it) and reset, cached spectra and a
// heavy-ball run are void.
int state_changed(fgpe_ctx* c, int comp)
{
    hipLaunchKernelGGL(k_set_one, dim3(1), dim3(1), 0, c->stream, c->scal + SC_SCALE + comp);
    FGPE_LAUNCHCHK("k_set_one");
    c->spec_valid = false;
    c->hb_active = false;
    c->imag_iter = 0;
    return FGPE_OK;
}

const double* pend_of(fgpe_ctx* c, int comp) { return c->scal + SC_SCALE + comp; }

int vstage_of(fgpe_ctx* c, int comp, double** v)
{
    if (!c->vstage[comp]) FGPE_RET(dalloc(c, (void**)&c->vstage[comp], sizeof(double) * c->ntot));
    *v = c->vstage[comp];
    return FGPE_OK;
}

// -XMAX + i DX + DX/2 on the host (src/psi.cu:134-150)
double centre_h(const fgpe_geom* g, int a, int i) { return -g->xmax[a] + i * g->dx[a] + 0.5 * g->dx[a]; }

// Per-axis factor tables: f(a, i) for the three axes, laid out x, y, z in
// the context's axis buffer.  The copy is ordered on the context's stream
// after the launches that read the previous tables, and the host waits for
// it, so the next call may rewrite the buffer.
template <typename F> int axis_tables(fgpe_ctx* c, F f, const double** tab)
{
    const int n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
    const size_t cap = 2 * (size_t)(n0 + n1 + n2);
    if (!c->axbuf) FGPE_RET(dalloc(c, (void**)&c->axbuf, sizeof(double) * cap));
    std::vector<double> h((size_t)(n0 + n1 + n2));
    for (int i = 0; i < n0; ++i) h[i] = f(0, i);
    for (int i = 0; i < n1; ++i) h[n0 + i] = f(1, i);
    for (int i = 0; i < n2; ++i) h[n0 + n1 + i] = f(2, i);
    FGPE_HIPCHK(hipMemcpyAsync(c->axbuf, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    tab[0] = c->axbuf;
    tab[1] = c->axbuf + n0;
    tab[2] = c->axbuf + n0 + n1;
    return FGPE_OK;
}

// (cos, sin) pairs per x index, in the same buffer
template <typename F> int x_pairs(fgpe_ctx* c, F f, const double2** tab)
{
    const int n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
    const size_t cap = 2 * (size_t)(n0 + n1 + n2);
    if (!c->axbuf) FGPE_RET(dalloc(c, (void**)&c->axbuf, sizeof(double) * cap));
    std::vector<double2> h((size_t)n0);
    for (int i = 0; i < n0; ++i) h[i] = f(i);
    FGPE_HIPCHK(hipMemcpyAsync(c->axbuf, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, c->stream));
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    *tab = (const double2*)c->axbuf;
    return FGPE_OK;
}

}  // namespace

extern "C" {

int fgpe_psi_init(fgpe_ctx* c, int comp, int kind, const fgpe_geom* g, const double ax[3], double rand_re,
                  double rand_im, unsigned long long seed)
{
    FGPE_RET(check_comp(c, comp, false));
    if (!g || (kind != FGPE_INIT_VOID && !ax)) return fail(FGPE_EINVAL, "NULL geometry or ax");
    if (kind < FGPE_INIT_HARMONIC || kind > FGPE_INIT_VOID) return fail(FGPE_EINVAL, "unknown psi kind");
    if (!c->P0[comp]) FGPE_RET(dalloc(c, (void**)&c->P0[comp], sizeof(double2) * c->ntot));
    const GeomD d = geom_of(c, g);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const double* t[3] = {nullptr, nullptr, nullptr};
    if (kind != FGPE_INIT_VOID) {
        // (x / ax)^2 per axis, as getX(p, i) / ax[a] squared in the reference
        FGPE_RET(axis_tables(c, [&](int a, int i) {
            const double x = centre_h(g, a, i) / ax[a];
            return x * x;
        }, t));
    }
    switch (kind) {
    case FGPE_INIT_HARMONIC:
        hipLaunchKernelGGL(k_psi_init<FGPE_INIT_HARMONIC>, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->P0[comp], d,
                           t[0], t[1], t[2], rand_re, rand_im, k0, k1, c->partials);
        break;
    case FGPE_INIT_TUBE:
        hipLaunchKernelGGL(k_psi_init<FGPE_INIT_TUBE>, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->P0[comp], d,
                           t[0], t[1], t[2], rand_re, rand_im, k0, k1, c->partials);
        break;
    default:
        hipLaunchKernelGGL(k_psi_init<FGPE_INIT_VOID>, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->P0[comp], d,
                           t[0], t[1], t[2], rand_re, rand_im, k0, k1, c->partials);
    }
    FGPE_LAUNCHCHK("k_psi_init");
    return state_changed(c, comp);
}

int fgpe_psi_add_noise(fgpe_ctx* c, int comp, double amp_re, double amp_im, unsigned long long seed,
                       unsigned int stream)
{
    FGPE_RET(check_comp(c, comp, true));
    if (stream == 0) return fail(FGPE_EINVAL, "stream 0 is the initial state's noise");
    hipLaunchKernelGGL(k_psi_noise, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->P0[comp], c->ntot, amp_re, amp_im,
                       (uint32_t)stream, (uint32_t)seed, (uint32_t)(seed >> 32), pend_of(c, comp));
    FGPE_LAUNCHCHK("k_psi_noise");
    return state_changed(c, comp);
}

int fgpe_psi_imprint_vortex_tube(fgpe_ctx* c, int comp, const fgpe_geom* g, double ncirc)
{
    FGPE_RET(check_comp(c, comp, true));
    if (!g) return fail(FGPE_EINVAL, "NULL geometry");
    const double xlen = 2 * g->xmax[0];
    const double2* cs = nullptr;
    FGPE_RET(x_pairs(c, [&](int i) {
        const double x = centre_h(g, 0, i);
        return double2{std::cos(ncirc * 2 * kPi / xlen * x), std::sin(ncirc * 2 * kPi / xlen * x)};
    }, &cs));
    hipLaunchKernelGGL(k_psi_vortex, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->P0[comp], geom_of(c, g), cs,
                       pend_of(c, comp));
    FGPE_LAUNCHCHK("k_psi_vortex");
    return state_changed(c, comp);
}

int fgpe_psi_force_lattice(fgpe_ctx* c, int comp, int lattice, const fgpe_geom* g, double amp, double k,
                           const double off[3], double npart)
{
    FGPE_RET(check_comp(c, comp, true));
    if (!g || !off) return fail(FGPE_EINVAL, "NULL geometry or offset");
    const GeomD d = geom_of(c, g);
    // avg = pow(npart / (DX0 DX1 DX2), 0.5) on the host, as the reference
    const double avg = std::pow(npart / (g->dx[0] * g->dx[1] * g->dx[2]), 0.5);
    const double r3 = std::sqrt(3.0), r2 = std::sqrt(2.0);
    const dim3 gr(nblocks(c)), bl(IB);
    double2* p = c->P0[comp];
    const double* pe = pend_of(c, comp);
    switch (lattice) {
#define FGPE_FORCE_CASE(L_)                                                                                         \
    case L_:                                                                                                        \
        hipLaunchKernelGGL(k_psi_force<L_>, gr, bl, 0, c->stream, p, d, amp, k, off[0], off[1], off[2], avg, r3, r2, \
                           pe);                                                                                     \
        break;
        FGPE_FORCE_CASE(FGPE_LATTICE_1D)
        FGPE_FORCE_CASE(FGPE_LATTICE_2D_TRIANGULAR)
        FGPE_FORCE_CASE(FGPE_LATTICE_2D_SQUARE)
        FGPE_FORCE_CASE(FGPE_LATTICE_3D_HEX)
        FGPE_FORCE_CASE(FGPE_LATTICE_3D_FCC)
        FGPE_FORCE_CASE(FGPE_LATTICE_3D_BCC)
        FGPE_FORCE_CASE(FGPE_LATTICE_3D_SC)
#undef FGPE_FORCE_CASE
    default:
        return fail(FGPE_EINVAL, "unknown lattice");
    }
    FGPE_LAUNCHCHK("k_psi_force");
    return state_changed(c, comp);
}

int fgpe_psi_normalize(fgpe_ctx* c, int comp, double npart, const fgpe_geom* g, double* host_norm)
{
    FGPE_RET(check_comp(c, comp, true));
    if (!g) return fail(FGPE_EINVAL, "NULL geometry");
    const unsigned nb = nblocks(c);
    double* tmp = c->scal + SC_TMP;
    hipLaunchKernelGGL(k_psi_sumsq, dim3(nb), dim3(IB), 0, c->stream, c->P0[comp], c->ntot, pend_of(c, comp),
                       c->partials);
    FGPE_LAUNCHCHK("k_psi_sumsq");
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(IB), 0, c->stream, c->partials, (long long)nb, tmp);
    FGPE_LAUNCHCHK("k_sum_partials");
    hipLaunchKernelGGL(k_psi_rescale, dim3(nb), dim3(IB), 0, c->stream, c->P0[comp], c->ntot, tmp, npart, g->dx[0],
                       g->dx[1], g->dx[2], pend_of(c, comp), c->partials);
    FGPE_LAUNCHCHK("k_psi_rescale");
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(IB), 0, c->stream, c->partials, (long long)nb, tmp + 1);
    FGPE_LAUNCHCHK("k_sum_partials");
    FGPE_RET(state_changed(c, comp));
    if (host_norm) {
        FGPE_HIPCHK(hipMemcpyAsync(host_norm, tmp + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    }
    return FGPE_OK;
}

int fgpe_vext_init(fgpe_ctx* c, int comp, int kind, const fgpe_geom* g, const double ax[3])
{
    FGPE_RET(check_comp(c, comp, false));
    if (!g || (kind != FGPE_INIT_VOID && !ax)) return fail(FGPE_EINVAL, "NULL geometry or ax");
    if (kind < FGPE_INIT_HARMONIC || kind > FGPE_INIT_VOID) return fail(FGPE_EINVAL, "unknown vext kind");
    double* v;
    FGPE_RET(vstage_of(c, comp, &v));
    const GeomD d = geom_of(c, g);
    const double* t[3] = {nullptr, nullptr, nullptr};
    if (kind != FGPE_INIT_VOID) {
        // pow(x, 2) / pow(ax, 2) with x = getX / ax (src/vext.cu:72-85)
        FGPE_RET(axis_tables(c, [&](int a, int i) {
            const double x = centre_h(g, a, i) / ax[a];
            return (x * x) / (ax[a] * ax[a]);
        }, t));
    }
    if (kind == FGPE_INIT_HARMONIC)
        hipLaunchKernelGGL(k_vext_init<FGPE_INIT_HARMONIC>, dim3(nblocks(c)), dim3(IB), 0, c->stream, v, d, t[0], t[1],
                           t[2]);
    else if (kind == FGPE_INIT_TUBE)
        hipLaunchKernelGGL(k_vext_init<FGPE_INIT_TUBE>, dim3(nblocks(c)), dim3(IB), 0, c->stream, v, d, t[0], t[1],
                           t[2]);
    else
        hipLaunchKernelGGL(k_vext_init<FGPE_INIT_VOID>, dim3(nblocks(c)), dim3(IB), 0, c->stream, v, d, t[0], t[1],
                           t[2]);
    FGPE_LAUNCHCHK("k_vext_init");
    return FGPE_OK;
}

int fgpe_vext_init_box(fgpe_ctx* c, int comp, const fgpe_geom* g, double V, double B, const double w[3])
{
    FGPE_RET(check_comp(c, comp, false));
    if (!g || !w) return fail(FGPE_EINVAL, "NULL geometry or box half-widths");
    double* v;
    FGPE_RET(vstage_of(c, comp, &v));
    const double* t[3] = {nullptr, nullptr, nullptr};
    FGPE_RET(axis_tables(c, [&](int a, int i) {
        const double x = centre_h(g, a, i);
        return std::tanh(B * (x - w[a])) - std::tanh(B * (x + w[a]));
    }, t));
    hipLaunchKernelGGL(k_vext_box, dim3(nblocks(c)), dim3(IB), 0, c->stream, v, geom_of(c, g), V, t[0], t[1], t[2]);
    FGPE_LAUNCHCHK("k_vext_box");
    return FGPE_OK;
}

int fgpe_vext_add_cos(fgpe_ctx* c, int comp, const fgpe_geom* g, double A, double n)
{
    FGPE_RET(check_comp(c, comp, false));
    if (!g) return fail(FGPE_EINVAL, "NULL geometry");
    if (!c->vstage[comp]) return fail(FGPE_ESTATE, "no trap being assembled: call fgpe_vext_init first");
    const double L = 2 * g->xmax[0];
    const double* t[3] = {nullptr, nullptr, nullptr};
    FGPE_RET(axis_tables(c, [&](int a, int i) {
        return a == 0 ? A * std::cos(n * 2 * kPi / L * centre_h(g, 0, i)) : 0.0;
    }, t));
    hipLaunchKernelGGL(k_vext_cos, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->vstage[comp], geom_of(c, g), t[0]);
    FGPE_LAUNCHCHK("k_vext_cos");
    return FGPE_OK;
}

int fgpe_vext_commit(fgpe_ctx* c, int comp, double* host_vext)
{
    FGPE_RET(check_comp(c, comp, false));
    if (!c->vstage[comp]) return fail(FGPE_ESTATE, "no trap being assembled: call fgpe_vext_init first");
    const size_t N = (size_t)c->ntot;
    if (!c->Vr[comp]) FGPE_RET(dalloc(c, &c->Vr[comp], c->tsz() * N));
    if (c->mixed())
        hipLaunchKernelGGL(k_vext_store<float>, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->vstage[comp],
                           (float*)c->Vr[comp], c->ntot);
    else
        hipLaunchKernelGGL(k_vext_store<double>, dim3(nblocks(c)), dim3(IB), 0, c->stream, c->vstage[comp],
                           (double*)c->Vr[comp], c->ntot);
    FGPE_LAUNCHCHK("k_vext_store");
    c->vcomplex[comp] = false;
    if (host_vext) {
        std::vector<double> re(N);
        FGPE_HIPCHK(hipMemcpyAsync(re.data(), c->vstage[comp], sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
        FGPE_HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < N; ++i) {
            host_vext[2 * i] = re[i];
            host_vext[2 * i + 1] = 0.0;
        }
    }
    FGPE_HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, (void**)&c->vstage[comp], sizeof(double) * N);
    return FGPE_OK;
}

}  // extern "C"
