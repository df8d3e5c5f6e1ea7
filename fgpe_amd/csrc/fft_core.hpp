// fft_core.hpp -- Stockham FFT of one pencil held in registers, with the
// stage-to-stage transposes staged through LDS.  Written for CDNA4 (gfx950):
// wave64, 16-byte LDS accesses, padding so that 16 consecutive elements and
// consecutive pencils fall in different banks.
//
// One pencil of length N is owned by T threads; thread t holds E = N/T
// elements.  Stage s (radix r, Ns = product of the earlier radices) performs
// N/r butterflies; thread t performs butterflies j = t + T*m, m < E/r:
//     v[i] = x[j + i*N/r] * W_{Ns*r}^{i*(j mod Ns)}      (i < r)
//     v    = DFT_r(v)
//     y[(j/Ns)*Ns*r + (j mod Ns) + i*Ns] = v[i]
// (Stockham autosort, Van Loan, "Computational Frameworks for the FFT",
// sec. 1.7).  Because the first stage reads x[j + i*N/r0] and the last one
// writes y[j + i*N/rL], the register layouts at the two ends are fixed:
//   layout A: q = m*r0 + i  <->  element (t + T*m) + i*N/r0
//   layout B: q = m*rL + i  <->  element (t + T*m) + i*N/rL
// The forward transform runs the radices in plan order (A in, B out), the
// inverse in reverse order (B in, A out), so a forward transform followed by
// an inverse one needs no reshuffle, and neither does an inverse followed by a
// forward one.  Transforms are unnormalised, forward = exp(-2 pi i jk/N).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "radix_consts.hpp"

namespace fgpe {

// Device arrays through global-address-space pointers: a pointer read from
// an argument table in memory is otherwise generic, and generic (flat)
// accesses also count against the LDS wait counter, which would make every
// LDS exchange wait for outstanding HBM traffic.
// (The host pass of hipcc only parses kernel bodies; there the alias is a
// plain pointer, since address space 1 means something else on the host.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T> using gp = __attribute__((address_space(1))) T*;
#else
template <typename T> using gp = T*;
#endif
template <typename T> __device__ __forceinline__ gp<T> glob(T* p) { return (gp<T>)p; }

template <typename T> struct c2;
template <> struct c2<double> { using type = double2; };
template <> struct c2<float> { using type = float2; };
template <typename T> using c2_t = typename c2<T>::type;

template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return T2{a.x + b.x, a.y + b.y}; }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { return T2{a.x - b.x, a.y - b.y}; }
template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b)
{
    return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// i*S*a with S = +1 or -1
template <int S, typename T2> __device__ __forceinline__ T2 mul_i(T2 a)
{
    return S > 0 ? T2{-a.y, a.x} : T2{a.y, -a.x};
}

// cos/sin(2 pi m / R) of the codelet constants
template <int R> __device__ __forceinline__ double kcos(int m);
template <int R> __device__ __forceinline__ double ksin(int m);
#define FGPE_KTAB(R)                                                                   \
    template <> __device__ __forceinline__ double kcos<R>(int m) { return kCos##R[m]; } \
    template <> __device__ __forceinline__ double ksin<R>(int m) { return kSin##R[m]; }
FGPE_KTAB(3)
FGPE_KTAB(6)
FGPE_KTAB(8)
FGPE_KTAB(12)
FGPE_KTAB(16)
#undef FGPE_KTAB

// ---- in-register DFT codelets, W_R = exp(DIR * 2 pi i / R) -----------------
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<1, DIR> {
    template <typename T2> __device__ __forceinline__ static void run(T2*) {}
};

template <int DIR> struct Dft<2, DIR> {
    template <typename T2> __device__ __forceinline__ static void run(T2* v)
    {
        T2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR> struct Dft<3, DIR> {
    template <typename T2> __device__ __forceinline__ static void run(T2* v)
    {
        using T = decltype(v[0].x);
        const T s3 = (T)kSin3[1];
        T2 s = cadd(v[1], v[2]);
        T2 d = csub(v[1], v[2]);
        T2 m = T2{v[0].x - (T)0.5 * s.x, v[0].y - (T)0.5 * s.y};
        T2 rot = mul_i<DIR>(T2{s3 * d.x, s3 * d.y});
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, rot);
        v[2] = csub(m, rot);
    }
};

template <int DIR> struct Dft<4, DIR> {
    template <typename T2> __device__ __forceinline__ static void run(T2* v)
    {
        T2 t0 = cadd(v[0], v[2]);
        T2 t1 = csub(v[0], v[2]);
        T2 t2 = cadd(v[1], v[3]);
        T2 t3 = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

// Cooley-Tukey R = A*B inside registers: n = n1 + A n2, k = B k1 + k2,
// X[B k1 + k2] = sum_n1 W_A^{n1 k1} W_R^{n1 k2} sum_n2 x[n1 + A n2] W_B^{n2 k2}
template <int A, int B, int DIR> struct DftCT {
    template <typename T2> __device__ __forceinline__ static void run(T2* v)
    {
        using T = decltype(v[0].x);
        constexpr int R = A * B;
        T2 y[R];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) {
            T2 tmp[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) tmp[n2] = v[n1 + A * n2];
            Dft<B, DIR>::run(tmp);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) {
                const int m = (n1 * k2) % R;
                if (m == 0) {
                    y[n1 * B + k2] = tmp[k2];
                } else {
                    T2 w{(T)kcos<R>(m), (T)(DIR * ksin<R>(m))};
                    y[n1 * B + k2] = cmul(tmp[k2], w);
                }
            }
        }
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            T2 tmp[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) tmp[n1] = y[n1 * B + k2];
            Dft<A, DIR>::run(tmp);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) v[B * k1 + k2] = tmp[k1];
        }
    }
};

template <int DIR> struct Dft<6, DIR> : DftCT<2, 3, DIR> {};
template <int DIR> struct Dft<8, DIR> : DftCT<2, 4, DIR> {};
template <int DIR> struct Dft<12, DIR> : DftCT<4, 3, DIR> {};
template <int DIR> struct Dft<16, DIR> : DftCT<4, 4, DIR> {};

// ---- plans: radices in forward order, threads per pencil T -------------------
// Variant 0 favours few threads per pencil (E = 16 elements per thread for
// the long pencils); variant 1 uses more threads and fewer registers (E = 8)
// where a second decomposition exists.  `distinct` says whether variant 1
// differs from variant 0.  Variant 2 (the x pencils only) takes E = 4
// elements per thread in radix-4 stages where that is yet another plan: at
// N = 256 a pencil is then exactly one wave; on other lengths it repeats
// variant 1 (`distinct` false: differs from variant 1 or not).
template <int N, int V = 0> struct Plan;
#define FGPE_PLAN(N_, V_, L_, T_, R0, R1, R2, D_)                 \
    template <> struct Plan<N_, V_> {                            \
        static constexpr int L = L_, T = T_, E = N_ / T_;        \
        static constexpr int r[4] = {R0, R1, R2, 1};             \
        static constexpr bool distinct = D_;                     \
    };
#define FGPE_PLAN4(N_, V_, T_, R0, R1, R2, R3, D_)                \
    template <> struct Plan<N_, V_> {                            \
        static constexpr int L = 4, T = T_, E = N_ / T_;         \
        static constexpr int r[4] = {R0, R1, R2, R3};            \
        static constexpr bool distinct = D_;                     \
    };
FGPE_PLAN(16, 0, 2, 4, 4, 4, 1, false)
FGPE_PLAN(32, 0, 2, 4, 8, 4, 1, false)
FGPE_PLAN(48, 0, 2, 4, 12, 4, 1, false)
FGPE_PLAN(64, 0, 2, 8, 8, 8, 1, false)
FGPE_PLAN(96, 0, 3, 8, 6, 4, 4, false)
FGPE_PLAN(128, 0, 2, 8, 16, 8, 1, false)
FGPE_PLAN(192, 0, 3, 16, 12, 4, 4, false)
FGPE_PLAN(256, 0, 2, 16, 16, 16, 1, false)
FGPE_PLAN(384, 0, 3, 16, 12, 8, 4, false)
FGPE_PLAN(512, 0, 3, 64, 8, 8, 8, false)
FGPE_PLAN(1024, 0, 3, 64, 16, 8, 8, false)
FGPE_PLAN(16, 1, 2, 4, 4, 4, 1, false)
FGPE_PLAN(32, 1, 2, 4, 8, 4, 1, false)
FGPE_PLAN(48, 1, 2, 4, 12, 4, 1, false)
FGPE_PLAN(64, 1, 3, 16, 4, 4, 4, true)
FGPE_PLAN(96, 1, 3, 8, 6, 4, 4, false)
FGPE_PLAN(128, 1, 3, 16, 8, 4, 4, true)
FGPE_PLAN(192, 1, 3, 16, 12, 4, 4, false)
FGPE_PLAN(256, 1, 3, 32, 8, 4, 8, true)
FGPE_PLAN4(384, 1, 32, 12, 4, 4, 2, true)
FGPE_PLAN(512, 1, 3, 64, 8, 8, 8, false)
FGPE_PLAN(1024, 1, 3, 64, 16, 8, 8, false)
template <int N> struct Plan<N, 2> : Plan<N, 1> {
    static constexpr bool distinct = false;
};
FGPE_PLAN4(128, 2, 32, 4, 4, 4, 2, true)
FGPE_PLAN4(256, 2, 64, 4, 4, 4, 4, true)
#undef FGPE_PLAN
#undef FGPE_PLAN4
template <int N, int V, int DIR, int S> __host__ __device__ constexpr int radix_of()
{
    return DIR < 0 ? Plan<N, V>::r[S] : Plan<N, V>::r[Plan<N, V>::L - 1 - S];
}
template <int N, int V, int DIR, int S> __host__ __device__ constexpr int ns_of()
{
    if constexpr (S == 0) return 1;
    else return ns_of<N, V, DIR, S - 1>() * radix_of<N, V, DIR, S - 1>();
}

// LDS placement: one pad slot after every 16 elements; pencil stride odd.
template <int N> __host__ __device__ constexpr int lds_stride() { return (N + N / 16) | 1; }
__device__ __forceinline__ int lds_phys(int a) { return a + (a >> 4); }

// register slot q of layout A / B -> element index within the pencil
template <int N, int V> __device__ __forceinline__ int elemA(int t, int q)
{
    constexpr int r = Plan<N, V>::r[0];
    return (t + Plan<N, V>::T * (q / r)) + (q % r) * (N / r);
}
template <int N, int V> __device__ __forceinline__ int elemB(int t, int q)
{
    constexpr int r = Plan<N, V>::r[Plan<N, V>::L - 1];
    return (t + Plan<N, V>::T * (q / r)) + (q % r) * (N / r);
}

// Twiddle table entry W_N^m = exp(-2 pi i m / N) (forward), conjugated for the
// inverse.
template <int DIR, typename T2, typename TP> __device__ __forceinline__ T2 twid(TP tw, int m)
{
    T2 w = tw[m];
    if (DIR > 0) w.y = -w.y;
    return w;
}

// A twiddle table staged in LDS (the kernel copies it there once per block).
// The host pass of the compiler only parses the kernels, and HIP's vector
// types have no copy from an address-space-qualified reference, so there the
// table is a plain pointer.
#ifdef __HIP_DEVICE_COMPILE__
template <typename T2> using lds_table = __attribute__((address_space(3))) const T2*;
#else
template <typename T2> using lds_table = const T2*;
#endif

// Copies the N-entry twiddle table into the block's LDS array twl (every
// thread of the block must call it) and returns it as an LDS table, so that
// the stages read twiddles with ds loads rather than through the vector cache.
template <typename T2, int N>
__device__ __forceinline__ lds_table<T2> stage_twiddles(T2* twl, const T2* __restrict__ twg)
{
    const gp<const T2> g = glob(twg);
    for (int i = threadIdx.x; i < N; i += blockDim.x) twl[i] = g[i];
    __syncthreads();
    return (lds_table<T2>)twl;
}

// Ordering of the stage-to-stage exchange.  A pencil owned by consecutive
// lanes of one wave (x pencils: T <= 64 threads, t = lane % T) needs only
// wavefront-scope ordering: a wave's LDS operations are performed in program
// order, so the fences below just keep the compiler from moving the reads
// above the writes (and the next writes above the reads).  Pencils spread
// over several waves (the strided tiles) need the block barrier.
template <bool WAVE> __device__ __forceinline__ void exchange_sync()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Stage-to-stage exchange through LDS: write / sync / read / sync.
template <typename T2, int N, int V, int DIR, int S, bool WAVE>
__device__ __forceinline__ void exchange(T2* v, int t, T2* lds)
{
    constexpr int TT = Plan<N, V>::T, E = N / TT;
    constexpr int r = radix_of<N, V, DIR, S>();
    constexpr int Ns = ns_of<N, V, DIR, S>();
    constexpr int M = E / r;
    constexpr int r2 = radix_of<N, V, DIR, S + 1>();
    constexpr int M2 = E / r2;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int j = t + TT * m;
        const int k = j % Ns;
        const int base = (j / Ns) * Ns * r + k;
#pragma unroll
        for (int i = 0; i < r; ++i) lds[lds_phys(base + i * Ns)] = v[m * r + i];
    }
    exchange_sync<WAVE>();
#pragma unroll
    for (int m = 0; m < M2; ++m) {
        const int j = t + TT * m;
#pragma unroll
        for (int i = 0; i < r2; ++i) v[m * r2 + i] = lds[lds_phys(j + i * (N / r2))];
    }
    exchange_sync<WAVE>();
}

// The same exchange through a scalar LDS buffer of the same length, half the
// bytes: the real parts go through it first, then the imaginary parts, each
// with its own write / sync / read / sync.  A kernel whose LDS limits its
// blocks per CU fits more of them this way, at twice the barriers.
template <typename T2, int N, int V, int DIR, int S, bool WAVE>
__device__ __forceinline__ void exchange(T2* v, int t, decltype(T2{}.x)* lds)
{
    constexpr int TT = Plan<N, V>::T, E = N / TT;
    constexpr int r = radix_of<N, V, DIR, S>();
    constexpr int Ns = ns_of<N, V, DIR, S>();
    constexpr int M = E / r;
    constexpr int r2 = radix_of<N, V, DIR, S + 1>();
    constexpr int M2 = E / r2;
#pragma unroll
    for (int part = 0; part < 2; ++part) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int j = t + TT * m;
            const int k = j % Ns;
            const int base = (j / Ns) * Ns * r + k;
#pragma unroll
            for (int i = 0; i < r; ++i) lds[lds_phys(base + i * Ns)] = part ? v[m * r + i].y : v[m * r + i].x;
        }
        exchange_sync<WAVE>();
#pragma unroll
        for (int m = 0; m < M2; ++m) {
            const int j = t + TT * m;
#pragma unroll
            for (int i = 0; i < r2; ++i) {
                const auto w = lds[lds_phys(j + i * (N / r2))];
                if (part) v[m * r2 + i].y = w;
                else v[m * r2 + i].x = w;
            }
        }
        exchange_sync<WAVE>();
    }
}

template <typename T2, int N, int V, int DIR, int S, bool WAVE, typename TP, typename LT>
__device__ __forceinline__ void fft_stages(T2* v, int t, LT* lds, TP tw)
{
    constexpr int L = Plan<N, V>::L, TT = Plan<N, V>::T, E = N / TT;
    constexpr int r = radix_of<N, V, DIR, S>();
    constexpr int Ns = ns_of<N, V, DIR, S>();
    constexpr int M = E / r;
    static_assert(E % r == 0, "plan: E must be a multiple of every radix");
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int j = t + TT * m;
        if constexpr (Ns > 1) {
            const int k = j % Ns;
#pragma unroll
            for (int i = 1; i < r; ++i)
                v[m * r + i] = cmul(v[m * r + i], twid<DIR, T2>(tw, i * k * (N / (Ns * r))));
        }
        Dft<r, DIR>::run(v + m * r);
    }
    if constexpr (S + 1 < L) {
        exchange<T2, N, V, DIR, S, WAVE>(v, t, lds);
        fft_stages<T2, N, V, DIR, S + 1, WAVE, TP, LT>(v, t, lds, tw);
    }
}

// The FP32 twiddle pointer as a value the compiler cannot see through.  A
// kernel that runs a forward and an inverse transform of the same length
// loads the same table entries in both; left visible, those loads are merged
// and the values stay live in registers across everything between the two
// transforms, which costs the FP32 kernels occupancy (reloading them is an
// L1/L2 hit).  The FP64 kernels measured faster with the merged loads.  An
// LDS copy of the table is left alone.
template <typename T2, typename TP> __device__ __forceinline__ TP opaque_table(TP p)
{
    if constexpr (sizeof(T2) == 8 && !std::is_same_v<TP, lds_table<T2>>) asm volatile("" : "+s"(p));
    return p;
}

// Forward: v in layout A -> layout B.  Inverse: layout B -> layout A.
// lds: this pencil's lds_stride<N>() elements, complex (T2) or, for the
// split exchange, scalar.  WAVE: the pencil's T threads
// are consecutive lanes of one wave (then only that wave must call);
// otherwise every thread of the block must call (block barriers).  tw: the
// twiddle table of length N, in global memory (gp) or staged in LDS.
template <typename T2, int N, int V, bool WAVE = false, typename TP, typename LT>
__device__ __forceinline__ void fft_forward(T2* v, int t, LT* lds, TP tw)
{
    static_assert(!WAVE || (Plan<N, V>::T <= 64 && 64 % Plan<N, V>::T == 0), "pencil must fit in a wave");
    fft_stages<T2, N, V, -1, 0, WAVE>(v, t, lds, opaque_table<T2>(tw));
}
template <typename T2, int N, int V, bool WAVE = false, typename TP, typename LT>
__device__ __forceinline__ void fft_inverse(T2* v, int t, LT* lds, TP tw)
{
    static_assert(!WAVE || (Plan<N, V>::T <= 64 && 64 % Plan<N, V>::T == 0), "pencil must fit in a wave");
    fft_stages<T2, N, V, +1, 0, WAVE>(v, t, lds, opaque_table<T2>(tw));
}

}  // namespace fgpe
