// launch.hpp -- template launchers of the pass kernels, dispatched over the
// compiled pencil lengths and the two plan variants (the variant of each
// kernel class and axis is a per-context choice, set by fgpe_autotune or
// fgpe_set_variant).  Included once per precision (launch_f64.hip,
// launch_f32.hip) so the two sets of kernels compile in parallel.
#pragma once
#include "ctx.hpp"

namespace fgpe {
namespace detail {

template <typename T> const c2_t<T>* tw_of(fgpe_ctx* c, int axis)
{
    return (const c2_t<T>*)(sizeof(T) == 8 ? c->tw64[axis] : c->tw32[axis]);
}

template <typename T, int N, int V> int fft_x_n(fgpe_ctx* c, const void* in, void* out, int dir)
{
    constexpr int P = px_of<T, N, V>();
    const long long rows = c->ntot / N;
    if (rows % P) return fail(FGPE_EUNSUP, "x rows not a multiple of the pencils per block");
    dim3 grid((unsigned)(rows / P)), block(P * Plan<N, V>::T);
    using T2 = c2_t<T>;
    TScope ts(c, FGPE_KCLASS_FFTX);
    if (dir < 0)
        hipLaunchKernelGGL((k_fft_x<T, N, V, -1>), grid, block, 0, c->stream, (const T2*)in, (T2*)out,
                           tw_of<T>(c, 0));
    else
        hipLaunchKernelGGL((k_fft_x<T, N, V, +1>), grid, block, 0, c->stream, (const T2*)in, (T2*)out,
                           tw_of<T>(c, 0));
    FGPE_LAUNCHCHK("k_fft_x");
    return FGPE_OK;
}

template <typename T, int N, int V> int strided_n(fgpe_ctx* c, void* data, int axis, int dir, int nch, int ncomp,
                                                  const double* scale, const int* stop, bool half, const void* tab,
                                                  int ninst, int slab, int nslab)
{
    constexpr int P = ps_of<T, N, V>();
    StridedArgs a = slab_of(strided_args(c, data, axis, scale, nch, ncomp, half, stop), slab, nslab);
    a.rev = (c->order >> (dir < 0 ? 0 : 2)) & 1;
    if (a.ncols % P) return fail(FGPE_EUNSUP, "slab width not a multiple of the strided tile width");
    dim3 grid((unsigned)strided_blocks<P>(a), (unsigned)ninst), block(P * Plan<N, V>::T);
    const StridedArgs* t = (const StridedArgs*)tab;
    TScope ts(c, FGPE_KCLASS_STRIDED);
    if (dir < 0)
        hipLaunchKernelGGL((k_strided<T, N, V, -1>), grid, block, 0, c->stream, tw_of<T>(c, axis), a, t);
    else
        hipLaunchKernelGGL((k_strided<T, N, V, +1>), grid, block, 0, c->stream, tw_of<T>(c, axis), a, t);
    FGPE_LAUNCHCHK("k_strided");
    return FGPE_OK;
}

template <typename T, int N, int V> int last_n(fgpe_ctx* c, void* data, int axis, int nch, int ncomp,
                                               const double* scale, bool filter, const int* stop, bool half,
                                               const void* tab, int ninst, int slab, int nslab, bool vpx)
{
    constexpr int P = ps_of<T, N, V>();
    LastArgs a = last_args(c, data, axis, nch, ncomp, scale, filter, stop, half, vpx);
    a.s = slab_of(a.s, slab, nslab);
    a.s.rev = (c->order >> 1) & 1;
    if (a.s.ncols % P) return fail(FGPE_EUNSUP, "slab width not a multiple of the strided tile width");
    dim3 grid((unsigned)strided_blocks<P>(a.s), (unsigned)ninst), block(P * Plan<N, V>::T);
    TScope ts(c, FGPE_KCLASS_LAST);
    hipLaunchKernelGGL((k_last<T, N, V>), grid, block, 0, c->stream, tw_of<T>(c, axis), a, (const LastArgs*)tab);
    FGPE_LAUNCHCHK("k_last");
    return FGPE_OK;
}

template <typename T, int N, int V, int NC, bool DENS, int MODE>
int xstage_t(fgpe_ctx* c, const XArgs<T>& a, const void* tab, int ninst)
{
    constexpr int P = px_of<T, N, V>();
    const long long rows = c->ntot / N;
    if (rows % P) return fail(FGPE_EUNSUP, "x rows not a multiple of the pencils per block");
    dim3 grid((unsigned)(rows / P), (unsigned)ninst), block(P * Plan<N, V>::T);
    XArgs<T> b = a;
    b.rev = (c->order >> 3) & 1;
    TScope ts(c, FGPE_KCLASS_XSTAGE);
    // Non-temporal operand streams when the working set exceeds the
    // memory-side cache: a context's own (time-stepping modes, long rows), or
    // a batched heavy-ball launch, whose groups share the cache with every
    // other group of the sweep.
    if constexpr ((N >= 256 && MODE == MODE_RK) || MODE == MODE_HB) {
        if (c->working_set() * (size_t)ninst > kMallBytes || (MODE == MODE_HB && tab)) {
            hipLaunchKernelGGL((k_xstage<T, N, V, NC, DENS, MODE, true>), grid, block, 0, c->stream, b,
                               tw_of<T>(c, 0), (const XArgs<T>*)tab);
            FGPE_LAUNCHCHK("k_xstage");
            return FGPE_OK;
        }
    }
    hipLaunchKernelGGL((k_xstage<T, N, V, NC, DENS, MODE>), grid, block, 0, c->stream, b, tw_of<T>(c, 0),
                       (const XArgs<T>*)tab);
    FGPE_LAUNCHCHK("k_xstage");
    return FGPE_OK;
}

template <typename T, int N, int V, int NC, bool DENS>
int xstage_m(fgpe_ctx* c, const XArgs<T>& a, int mode, const void* tab, int ninst)
{
    switch (mode) {
    case MODE_PROLOGUE: return xstage_t<T, N, V, NC, DENS, MODE_PROLOGUE>(c, a, tab, ninst);
    case MODE_RK: return xstage_t<T, N, V, NC, DENS, MODE_RK>(c, a, tab, ninst);
    case MODE_APPLY: return xstage_t<T, N, V, NC, DENS, MODE_APPLY>(c, a, tab, ninst);
    case MODE_HB: return xstage_t<T, N, V, NC, DENS, MODE_HB>(c, a, tab, ninst);
    }
    return fail(FGPE_EINVAL, "bad x-stage mode");
}

template <typename T, int N, int V>
int xstage_n(fgpe_ctx* c, const XArgs<T>& a, int nc, bool dens, int mode, const void* tab, int ninst)
{
    if (nc == 1)
        return dens ? xstage_m<T, N, V, 1, true>(c, a, mode, tab, ninst)
                    : xstage_m<T, N, V, 1, false>(c, a, mode, tab, ninst);
    return dens ? xstage_m<T, N, V, 2, true>(c, a, mode, tab, ninst)
                : xstage_m<T, N, V, 2, false>(c, a, mode, tab, ninst);
}

template <typename T, typename YT, int N, int V, bool DENS> int adj_t(fgpe_ctx* c, const AdjArgs& a, bool epi)
{
    constexpr int P = px_of<T, N, V>();
    const long long rows = c->ntot / N;
    if (rows % P) return fail(FGPE_EUNSUP, "x rows not a multiple of the pencils per block");
    dim3 grid((unsigned)(rows / P)), block(P * Plan<N, V>::T);
    TScope ts(c, FGPE_KCLASS_XSTAGE);
    if (epi) hipLaunchKernelGGL((k_adj_epi<T, YT, N, V, DENS>), grid, block, 0, c->stream, a, tw_of<T>(c, 0));
    else hipLaunchKernelGGL((k_adj_pro<T, YT, N, V, DENS>), grid, block, 0, c->stream, a, tw_of<T>(c, 0));
    FGPE_LAUNCHCHK("k_adj");
    return FGPE_OK;
}

template <typename T, int N, int V> int adj_n(fgpe_ctx* c, const AdjArgs& a, bool y64, bool dens, bool epi)
{
    if (y64) return dens ? adj_t<T, double, N, V, true>(c, a, epi) : adj_t<T, double, N, V, false>(c, a, epi);
    return dens ? adj_t<T, T, N, V, true>(c, a, epi) : adj_t<T, T, N, V, false>(c, a, epi);
}

}  // namespace detail

#define FGPE_DISPATCH(len, var, CALL)                                        \
    switch (len) {                                                           \
    case 32: return (var) ? CALL(32, 1) : CALL(32, 0);                       \
    case 64: return (var) ? CALL(64, 1) : CALL(64, 0);                       \
    case 96: return (var) ? CALL(96, 1) : CALL(96, 0);                       \
    case 128: return (var) ? CALL(128, 1) : CALL(128, 0);                    \
    case 192: return (var) ? CALL(192, 1) : CALL(192, 0);                    \
    case 256: return (var) ? CALL(256, 1) : CALL(256, 0);                    \
    case 512: return (var) ? CALL(512, 1) : CALL(512, 0);                    \
    }

template <typename T> int Launch<T>::adj(fgpe_ctx* c, const AdjArgs& a, bool y64, bool dens, bool epi)
{
#define CALL(N_, V_) detail::adj_n<T, N_, V_>(c, a, y64, dens, epi)
    FGPE_DISPATCH(c->n[0], c->var[FGPE_KCLASS_XSTAGE][0], CALL)
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported nx");
}

template <typename T> long long Launch<T>::xblocks(fgpe_ctx* c)
{
#define CALL(N_, V_) (c->ntot / N_ / px_of<T, N_, V_>())
    FGPE_DISPATCH(c->n[0], c->var[FGPE_KCLASS_XSTAGE][0], CALL)
#undef CALL
    return 0;
}

template <typename T> int Launch<T>::fft_x(fgpe_ctx* c, const void* in, void* out, int dir)
{
#define CALL(N_, V_) detail::fft_x_n<T, N_, V_>(c, in, out, dir)
    FGPE_DISPATCH(c->n[0], c->var[FGPE_KCLASS_FFTX][0], CALL)
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported nx");
}

template <typename T> int Launch<T>::strided(fgpe_ctx* c, void* data, int axis, int dir, int nch, int ncomp,
                                             const double* scale, const int* stop, bool half, const void* tab,
                                             int ninst, int slab, int nslab)
{
#define CALL(N_, V_) \
    detail::strided_n<T, N_, V_>(c, data, axis, dir, nch, ncomp, scale, stop, half, tab, ninst, slab, nslab)
    FGPE_DISPATCH(c->n[axis], c->var[FGPE_KCLASS_STRIDED][axis], CALL)
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported strided axis length");
}

template <typename T> int Launch<T>::last(fgpe_ctx* c, void* data, int axis, int nch, int ncomp,
                                          const double* scale, bool filter, const int* stop, bool half,
                                          const void* tab, int ninst, int slab, int nslab, bool vpx)
{
#define CALL(N_, V_) \
    detail::last_n<T, N_, V_>(c, data, axis, nch, ncomp, scale, filter, stop, half, tab, ninst, slab, nslab, vpx)
    FGPE_DISPATCH(c->n[axis], c->var[FGPE_KCLASS_LAST][axis], CALL)
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported last axis length");
}

template <typename T> int Launch<T>::xstage(fgpe_ctx* c, const XArgs<T>& a, int nc, bool dens, int mode,
                                            const void* tab, int ninst)
{
#define CALL(N_, V_) detail::xstage_n<T, N_, V_>(c, a, nc, dens, mode, tab, ninst)
    FGPE_DISPATCH(c->n[0], c->var[FGPE_KCLASS_XSTAGE][0], CALL)
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported nx");
}

#undef FGPE_DISPATCH

}  // namespace fgpe
