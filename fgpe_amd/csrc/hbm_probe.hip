// Achievable HBM bandwidth of the access patterns the time step is built from,
// on the 256^3 FP64 workload's footprint: two complex channels of 256^3
// double2 (512 MiB, twice the 256 MB memory-side cache, so that nothing is
// reused between launches).  Each pattern is timed over 10 launches after 3
// warm-up launches with hipEvents; GB/s counts every byte read and written
// once.
//
//   read-only        contiguous loads, 8 in flight per thread
//   copy             contiguous loads into a second 512 MiB buffer, 4 per thread
//   in-place         contiguous load, scale, store to the same address, 4 per thread
//   hipMemcpy        the runtime's device-to-device copy
//   tile z / tile y  the FFT passes' tile pattern without the FFT: a block reads
//                    W consecutive x columns of all 256 points along z (pencils
//                    1 MiB apart) or along y (4 KiB apart), holds them in
//                    registers and writes them back scaled
//   chain            one RK stage's pass sequence without the FFTs, on the
//                    step's own footprint (a complex channel and a half channel
//                    of 128 stored columns, row pitch 144: 419 MB): a
//                    contiguous in-place pass over both (the x stage's spectral
//                    traffic), then y, z and y tile passes (16 columns, 512
//                    threads), each pass's blocks forward or reversed (the 16
//                    block orders) and the tiles enumerated outer index by outer
//                    index (as the step does) or column group by column group;
//                    every pass starts on what the one before it left in the
//                    memory-side cache
//
// Build: python -m fgpe_amd.build (fgpe_amd/hbm_probe).  Run: fgpe_amd/hbm_probe
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

#define PROBE_CHK(x)                                                                              \
    do {                                                                                          \
        const hipError_t err_ = (x);                                                              \
        if (err_ != hipSuccess) {                                                                 \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(err_), __LINE__);         \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

constexpr int NX = 256, NY = 256, NZ = 256, NCH = 2;
constexpr long long PLANE = (long long)NX * NY;
constexpr long long NTOT = PLANE * NZ;
constexpr long long NELEM = NCH * NTOT;                         // double2 elements
constexpr double BYTES = (double)NELEM * sizeof(double2);       // one sweep over the data

template <int U> __global__ void __launch_bounds__(256) k_read(const double2* __restrict__ a, double* __restrict__ sink)
{
    const long long i0 = (long long)blockIdx.x * 256 * U + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const double2 v = a[i0 + q * 256];
        s += v.x + v.y;
    }
    if (s == 1.2345e300) sink[0] = s;   // keeps the loads; never true for the zeroed data
}

template <int U> __global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ a, double2* __restrict__ b)
{
    const long long i0 = (long long)blockIdx.x * 256 * U + threadIdx.x;
    double2 v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) v[q] = a[i0 + q * 256];
#pragma unroll
    for (int q = 0; q < U; ++q) b[i0 + q * 256] = v[q];
}

template <int U> __global__ void __launch_bounds__(256) k_inplace(double2* __restrict__ d)
{
    const long long i0 = (long long)blockIdx.x * 256 * U + threadIdx.x;
    double2 v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) v[q] = d[i0 + q * 256];
#pragma unroll
    for (int q = 0; q < U; ++q) d[i0 + q * 256] = double2{v[q].x * 0.5, v[q].y * 0.5};
}

// One block per tile: channel, outer index (y for ZAXIS, else z) and a group of W
// x columns; thread (col, r0) holds points r0, r0 + R, ... of its column.
template <int W, int THREADS, bool ZAXIS> __global__ void __launch_bounds__(THREADS) k_tile(double2* __restrict__ d)
{
    constexpr int R = THREADS / W, PER = 256 / R;
    const int ntx = NX / W;
    long long t = blockIdx.x;
    const int ch = (int)(t / ((long long)ntx * 256));
    t -= (long long)ch * ntx * 256;
    const int outer = (int)(t / ntx), xt = (int)(t % ntx);
    const int col = threadIdx.x % W, r0 = threadIdx.x / W;
    const long long base = ch * NTOT + (ZAXIS ? (long long)outer * NX : (long long)outer * PLANE) + xt * W + col;
    const long long st = ZAXIS ? PLANE : NX;
    double2 v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) v[q] = d[base + (long long)(r0 + q * R) * st];
#pragma unroll
    for (int q = 0; q < PER; ++q) d[base + (long long)(r0 + q * R) * st] = double2{v[q].x * 0.5, v[q].y * 0.5};
}

// ---- chain: the passes of one RK stage without the FFTs ---------------------
constexpr int CH_W = 16, CH_TH = 512, CH_R = CH_TH / CH_W, CH_PER = 256 / CH_R;
constexpr int CH_HP = 144, CH_HC = 128;                       // half channel: pitch, stored columns used
constexpr long long CH_HALF = (long long)NX * NY * CH_HP;     // half-channel elements
constexpr int CH_NTF = NX / CH_W, CH_NHF = CH_HC / CH_W, CH_PERO = CH_NTF + CH_NHF;
constexpr int CH_TILES = 256 * CH_PERO;

// tile bx -> (outer index, column group, channel): outer index by outer index
// (all column groups of both channels of one outer index together) or column
// group by column group (all outer indices of one group, both channels, together)
__device__ __forceinline__ void chain_tile(int bx, bool colmaj, int& outer, int& g, int& ch)
{
    if (!colmaj) {
        outer = bx / CH_PERO;
        const int w = bx % CH_PERO;
        ch = w >= CH_NTF;
        g = ch ? w - CH_NTF : w;
    } else if (bx < CH_NHF * 2 * 256) {
        g = bx / 512;
        outer = (bx % 512) / 2;
        ch = bx % 2;
    } else {
        const int b2 = bx - CH_NHF * 512;
        g = CH_NHF + b2 / 256;
        outer = b2 % 256;
        ch = 0;
    }
}

template <bool ZAXIS> __global__ void __launch_bounds__(CH_TH)
k_chain_tile(double2* __restrict__ psi, double2* __restrict__ half, int rev, int colmaj, double sc)
{
    const int bx = rev ? CH_TILES - 1 - (int)blockIdx.x : (int)blockIdx.x;
    int outer, g, ch;
    chain_tile(bx, colmaj != 0, outer, g, ch);
    const int col = threadIdx.x % CH_W, r0 = threadIdx.x / CH_W;
    const long long row = ch ? CH_HP : NX, plane = row * NY;
    double2* d = ch ? half : psi;
    const long long base = (ZAXIS ? (long long)outer * row : (long long)outer * plane) + g * CH_W + col;
    const long long st = ZAXIS ? plane : row;
    double2 v[CH_PER];
#pragma unroll
    for (int q = 0; q < CH_PER; ++q) v[q] = d[base + (long long)(r0 + q * CH_R) * st];
#pragma unroll
    for (int q = 0; q < CH_PER; ++q) d[base + (long long)(r0 + q * CH_R) * st] = double2{v[q].x * sc, v[q].y * sc};
}

// contiguous in-place pass over the complex channel, then the half channel
__global__ void __launch_bounds__(256) k_chain_x(double2* __restrict__ psi, double2* __restrict__ half, int rev,
                                                 double sc)
{
    const long long nb_psi = NTOT / 1024, nb = nb_psi + CH_HALF / 1024;
    const long long b = rev ? nb - 1 - blockIdx.x : blockIdx.x;
    double2* d = b < nb_psi ? psi + b * 1024 : half + (b - nb_psi) * 1024;
    double2 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = d[threadIdx.x + q * 256];
#pragma unroll
    for (int q = 0; q < 4; ++q) d[threadIdx.x + q * 256] = double2{v[q].x * sc, v[q].y * sc};
}

struct Timer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

template <typename F> int timed(const char* name, double bytes, Timer& tm, F launch)
{
    for (int w = 0; w < 3; ++w) launch();
    PROBE_CHK(hipGetLastError());
    PROBE_CHK(hipDeviceSynchronize());
    PROBE_CHK(hipEventRecord(tm.e0, 0));
    for (int r = 0; r < 10; ++r) launch();
    PROBE_CHK(hipEventRecord(tm.e1, 0));
    PROBE_CHK(hipEventSynchronize(tm.e1));
    float ms = 0.f;
    PROBE_CHK(hipEventElapsedTime(&ms, tm.e0, tm.e1));
    const double s = 1e-3 * ms / 10;
    std::printf("%-34s %8.1f us  %6.0f GB/s  %5.1f%% of 8 TB/s\n", name, 1e6 * s, bytes / s / 1e9, bytes / s / 8e12 * 100);
    std::fflush(stdout);
    return 0;
}

template <int W, int THREADS, bool Z> int tile(double2* d, Timer& tm)
{
    static_assert(THREADS % W == 0 && 256 % (THREADS / W) == 0 && NX % W == 0, "tile shape");
    const unsigned blocks = (unsigned)(NCH * (NX / W) * 256);
    char name[64];
    std::snprintf(name, sizeof name, "tile %s, %d columns, %d threads", Z ? "z" : "y", W, THREADS);
    return timed(name, 2 * BYTES, tm, [&] { hipLaunchKernelGGL((k_tile<W, THREADS, Z>), dim3(blocks), dim3(THREADS), 0, 0, d); });
}

// All 16 block orders (bit 0: y forward, 1: z, 2: y inverse, 3: x pass reversed)
// of one enumeration; prints the best order's per-pass times (10 sequences each,
// events between the passes).
int chain(double2* psi, double2* half, int colmaj)
{
    static_assert(NTOT % 1024 == 0 && CH_HALF % 1024 == 0, "x-pass blocks");
    hipEvent_t e[5];
    for (auto& x : e) PROBE_CHK(hipEventCreate(&x));
    const unsigned xblocks = (unsigned)((NTOT + CH_HALF) / 1024);
    const double tile_bytes = 2.0 * sizeof(double2) * ((double)NTOT + (double)NX * NY * CH_HC);
    const double x_bytes = 2.0 * sizeof(double2) * ((double)NTOT + (double)CH_HALF);
    float best = 1e30f, bt[4] = {};
    int besto = 0;
    for (int o = 0; o < 16; ++o) {
        auto seq = [&](bool ev) {
            if (ev) (void)hipEventRecord(e[0], 0);
            hipLaunchKernelGGL(k_chain_x, dim3(xblocks), dim3(256), 0, 0, psi, half, (o >> 3) & 1, 1.0);
            if (ev) (void)hipEventRecord(e[1], 0);
            hipLaunchKernelGGL(k_chain_tile<false>, dim3(CH_TILES), dim3(CH_TH), 0, 0, psi, half, o & 1, colmaj, 1.0);
            if (ev) (void)hipEventRecord(e[2], 0);
            hipLaunchKernelGGL(k_chain_tile<true>, dim3(CH_TILES), dim3(CH_TH), 0, 0, psi, half, (o >> 1) & 1, colmaj,
                               1.0);
            if (ev) (void)hipEventRecord(e[3], 0);
            hipLaunchKernelGGL(k_chain_tile<false>, dim3(CH_TILES), dim3(CH_TH), 0, 0, psi, half, (o >> 2) & 1, colmaj,
                               1.0);
            if (ev) (void)hipEventRecord(e[4], 0);
        };
        for (int w = 0; w < 3; ++w) seq(false);
        PROBE_CHK(hipGetLastError());
        PROBE_CHK(hipDeviceSynchronize());
        float tot = 0.f, part[4] = {};
        for (int r = 0; r < 10; ++r) {
            seq(true);
            PROBE_CHK(hipEventSynchronize(e[4]));
            for (int k = 0; k < 4; ++k) {
                float ms = 0.f;
                PROBE_CHK(hipEventElapsedTime(&ms, e[k], e[k + 1]));
                part[k] += ms / 10;
                tot += ms / 10;
            }
        }
        if (tot < best) {
            best = tot;
            besto = o;
            for (int k = 0; k < 4; ++k) bt[k] = part[k];
        }
    }
    std::printf("chain, %s tiles, best block order %2d: x %.1f us (%.0f GB/s), y fwd %.1f us (%.0f GB/s), "
                "z %.1f us (%.0f GB/s), y inv %.1f us (%.0f GB/s)\n",
                colmaj ? "column-group-major" : "outer-major", besto, 1e3 * bt[0], x_bytes / bt[0] / 1e6, 1e3 * bt[1],
                tile_bytes / bt[1] / 1e6, 1e3 * bt[2], tile_bytes / bt[2] / 1e6, 1e3 * bt[3], tile_bytes / bt[3] / 1e6);
    std::fflush(stdout);
    for (auto& x : e) PROBE_CHK(hipEventDestroy(x));
    return 0;
}

}  // namespace

int main()
{
    double2 *a = nullptr, *b = nullptr;
    PROBE_CHK(hipMalloc(&a, (size_t)BYTES));
    PROBE_CHK(hipMalloc(&b, (size_t)BYTES));
    PROBE_CHK(hipMemset(a, 0, (size_t)BYTES));
    PROBE_CHK(hipMemset(b, 0, (size_t)BYTES));
    Timer tm;
    PROBE_CHK(hipEventCreate(&tm.e0));
    PROBE_CHK(hipEventCreate(&tm.e1));
    std::printf("footprint: %d x %d x %d x %d channels of double2 = %.0f MiB\n", NX, NY, NZ, NCH, BYTES / 1048576.0);
    int rc = 0;
    rc |= timed("read-only (8 per thread)", BYTES, tm, [&] {
        hipLaunchKernelGGL(k_read<8>, dim3((unsigned)(NELEM / (256 * 8))), dim3(256), 0, 0, a, (double*)b);
    });
    rc |= timed("copy (4 per thread)", 2 * BYTES, tm, [&] {
        hipLaunchKernelGGL(k_copy<4>, dim3((unsigned)(NELEM / (256 * 4))), dim3(256), 0, 0, a, b);
    });
    rc |= timed("copy (16 per thread)", 2 * BYTES, tm, [&] {
        hipLaunchKernelGGL(k_copy<16>, dim3((unsigned)(NELEM / (256 * 16))), dim3(256), 0, 0, a, b);
    });
    rc |= timed("in-place (4 per thread)", 2 * BYTES, tm, [&] {
        hipLaunchKernelGGL(k_inplace<4>, dim3((unsigned)(NELEM / (256 * 4))), dim3(256), 0, 0, a);
    });
    rc |= timed("in-place (16 per thread)", 2 * BYTES, tm, [&] {
        hipLaunchKernelGGL(k_inplace<16>, dim3((unsigned)(NELEM / (256 * 16))), dim3(256), 0, 0, a);
    });
    rc |= timed("hipMemcpy device to device", 2 * BYTES, tm,
                [&] { (void)hipMemcpyAsync(b, a, (size_t)BYTES, hipMemcpyDeviceToDevice, 0); });
    rc |= tile<8, 256, true>(a, tm);
    rc |= tile<16, 256, true>(a, tm);
    rc |= tile<16, 1024, true>(a, tm);
    rc |= tile<32, 512, true>(a, tm);
    rc |= tile<64, 1024, true>(a, tm);
    rc |= tile<16, 256, false>(a, tm);
    rc |= tile<16, 1024, false>(a, tm);
    rc |= tile<32, 1024, false>(a, tm);
    // the chain runs on the first 268 MB of a (complex channel) and of b (half channel)
    static_assert(CH_HALF <= NELEM, "half channel fits in b");
    rc |= chain(a, b, 0);
    rc |= chain(a, b, 1);
    PROBE_CHK(hipFree(a));
    PROBE_CHK(hipFree(b));
    return rc;
}
