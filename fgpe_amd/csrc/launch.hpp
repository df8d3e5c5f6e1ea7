// launch.hpp -- template launchers of the pass kernels, dispatched over the
// compiled pencil lengths and the two plan variants (the variant of each
// kernel class and axis is a per-context choice, set by fgpe_autotune or
// fgpe_set_variant).
//
// The per-length launchers (detail::*_n<T, N, V>) are instantiated one length
// and precision per translation unit (launch_inst.hip, compiled once per
// (T, N) by the build recipe) so that the kernels compile in parallel; the
// dispatch units (launch_f64.hip, launch_f32.hip) only declare them.
#pragma once
#include "ctx.hpp"

namespace fgpe {
namespace detail {

template <typename T> const c2_t<T>* tw_of(fgpe_ctx* c, int axis)
{
    return (const c2_t<T>*)(sizeof(T) == 8 ? c->tw64[axis] : c->tw32[axis]);
}

// blocks of an x-pencil launch: rows in groups of P pencils, the last group
// possibly partial
template <int P> inline long long x_groups(const fgpe_ctx* c, int n) { return (c->ntot / n + P - 1) / P; }

template <typename T, int N, int V> int fft_x_n(fgpe_ctx* c, const void* in, void* out, int dir)
{
    constexpr int P = px_of<T, N, V>();
    dim3 grid((unsigned)x_groups<P>(c, N)), block(P * Plan<N, V>::T);
    using T2 = c2_t<T>;
    TScope ts(c, FGPE_KCLASS_FFTX);
    const long long rows = c->ntot / N;
    if (dir < 0)
        hipLaunchKernelGGL((k_fft_x<T, N, V, -1>), grid, block, 0, c->stream, (const T2*)in, (T2*)out,
                           tw_of<T>(c, 0), rows);
    else
        hipLaunchKernelGGL((k_fft_x<T, N, V, +1>), grid, block, 0, c->stream, (const T2*)in, (T2*)out,
                           tw_of<T>(c, 0), rows);
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
    dim3 grid((unsigned)strided_blocks<P>(a.s), (unsigned)ninst), block(P * Plan<N, V>::T);
    TScope ts(c, FGPE_KCLASS_LAST);
    hipLaunchKernelGGL((k_last<T, N, V>), grid, block, 0, c->stream, tw_of<T>(c, axis), a, (const LastArgs*)tab);
    FGPE_LAUNCHCHK("k_last");
    return FGPE_OK;
}

template <typename T, int N, int V, int NC, bool DENS, int MODE, int UT>
int xstage_u(fgpe_ctx* c, const XArgs<T>& a, const void* tab, int ninst)
{
    constexpr int P = px_of<T, N, V>();
    dim3 grid((unsigned)x_groups<P>(c, N), (unsigned)ninst), block(P * Plan<N, V>::T);
    XArgs<T> b = a;
    b.rev = (c->order >> 3) & 1;
    TScope ts(c, FGPE_KCLASS_XSTAGE);
    // Non-temporal operand streams when the working set exceeds the
    // memory-side cache: a context's own (time-stepping modes, long rows), or
    // a batched heavy-ball launch, whose groups share the cache with every
    // other group of the sweep.
    if constexpr ((N >= 256 && MODE == MODE_RK) || MODE == MODE_HB) {
        if (c->working_set() * (size_t)ninst > kMallBytes || (MODE == MODE_HB && tab)) {
            hipLaunchKernelGGL((k_xstage<T, N, V, NC, DENS, MODE, true, UT>), grid, block, 0, c->stream, b,
                               tw_of<T>(c, 0), (const XArgs<T>*)tab);
            FGPE_LAUNCHCHK("k_xstage");
            return FGPE_OK;
        }
    }
    hipLaunchKernelGGL((k_xstage<T, N, V, NC, DENS, MODE, false, UT>), grid, block, 0, c->stream, b, tw_of<T>(c, 0),
                       (const XArgs<T>*)tab);
    FGPE_LAUNCHCHK("k_xstage");
    return FGPE_OK;
}

// the context's user term (fgpe_hparams::user_id, checked against the
// build's list when the Hamiltonian is set)
template <typename T, int N, int V, int NC, bool DENS, int MODE>
int xstage_t(fgpe_ctx* c, const XArgs<T>& a, const void* tab, int ninst)
{
#define FGPE_USER_CASE(ID, S) \
    if (c->hp.user_id == ID) return xstage_u<T, N, V, NC, DENS, MODE, ID>(c, a, tab, ninst);
    FGPE_USER_TERMS(FGPE_USER_CASE)
#undef FGPE_USER_CASE
    return xstage_u<T, N, V, NC, DENS, MODE, 0>(c, a, tab, ninst);
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

template <typename T, typename YT, int N, int V, bool DENS, int NC>
int adj_t(fgpe_ctx* c, const AdjArgs& a, bool epi)
{
    constexpr int P = px_of<T, N, V>();
    dim3 grid((unsigned)x_groups<P>(c, N)), block(P * Plan<N, V>::T);
    TScope ts(c, FGPE_KCLASS_XSTAGE);
    if (epi) hipLaunchKernelGGL((k_adj_epi<T, YT, N, V, DENS, NC>), grid, block, 0, c->stream, a, tw_of<T>(c, 0));
    else hipLaunchKernelGGL((k_adj_pro<T, YT, N, V, DENS, NC>), grid, block, 0, c->stream, a, tw_of<T>(c, 0));
    FGPE_LAUNCHCHK("k_adj");
    return FGPE_OK;
}

template <typename T, typename YT, int N, int V, int NC>
int adj_d(fgpe_ctx* c, const AdjArgs& a, bool dens, bool epi)
{
    return dens ? adj_t<T, YT, N, V, true, NC>(c, a, epi) : adj_t<T, YT, N, V, false, NC>(c, a, epi);
}

// y64: the state the Jacobian is taken at is FP64 (else T); the component
// count is the context's
template <typename T, int N, int V> int adj_n(fgpe_ctx* c, const AdjArgs& a, bool y64, bool dens, bool epi)
{
    if (c->hp.ncomp == 2)
        return y64 ? adj_d<T, double, N, V, 2>(c, a, dens, epi) : adj_d<T, T, N, V, 2>(c, a, dens, epi);
    return y64 ? adj_d<T, double, N, V, 1>(c, a, dens, epi) : adj_d<T, T, N, V, 1>(c, a, dens, epi);
}

}  // namespace detail

// Declarations (EXT = extern) or instantiations (EXT empty) of the
// per-length launchers of one precision, length and variant.
#define FGPE_LAUNCH_INST(EXT, T, N, V)                                                                            \
    EXT template int detail::fft_x_n<T, N, V>(fgpe_ctx*, const void*, void*, int);                                \
    EXT template int detail::strided_n<T, N, V>(fgpe_ctx*, void*, int, int, int, int, const double*, const int*,  \
                                                bool, const void*, int, int, int);                                \
    EXT template int detail::last_n<T, N, V>(fgpe_ctx*, void*, int, int, int, const double*, bool, const int*,    \
                                             bool, const void*, int, int, int, bool);                             \
    EXT template int detail::xstage_n<T, N, V>(fgpe_ctx*, const XArgs<T>&, int, bool, int, const void*, int);     \
    EXT template int detail::adj_n<T, N, V>(fgpe_ctx*, const AdjArgs&, bool, bool, bool);

// The x-class launchers of the lengths with a distinct variant 2.
#define FGPE_LAUNCH_INST_X(EXT, T, N, V)                                                                          \
    EXT template int detail::xstage_n<T, N, V>(fgpe_ctx*, const XArgs<T>&, int, bool, int, const void*, int);     \
    EXT template int detail::adj_n<T, N, V>(fgpe_ctx*, const AdjArgs&, bool, bool, bool);

#define FGPE_X2_CHECK(N_) static_assert(Plan<N_, 2>::distinct, "FGPE_FOR_EACH_N_X2 lists a length without a third plan");
FGPE_FOR_EACH_N_X2(FGPE_X2_CHECK)
#undef FGPE_X2_CHECK
#define FGPE_X2_CHECK(N_) static_assert(Plan<N_, 2>::distinct == x_variant2(N_), "x_variant2 out of date");
FGPE_FOR_EACH_N(FGPE_X2_CHECK)
#undef FGPE_X2_CHECK

// Dispatch on the runtime length to the compiled one.
#define FGPE_DISPATCH_CASE(N_) \
    case N_: return var ? CALL(N_, 1) : CALL(N_, 0);
// The x class: variant 2 where the length has a third plan (else variant 1).
#define FGPE_DISPATCH_X_CASE(N_)                                   \
    case N_:                                                       \
        if constexpr (Plan<N_, 2>::distinct) {                     \
            if (var == 2) return CALL(N_, 2);                      \
        }                                                          \
        return var ? CALL(N_, 1) : CALL(N_, 0);
#define FGPE_DISPATCH_X(len, variant)            \
    {                                            \
        const int var = (variant);               \
        switch (len) {                           \
            FGPE_FOR_EACH_N(FGPE_DISPATCH_X_CASE) \
        }                                        \
    }
#define FGPE_DISPATCH(len, variant)          \
    {                                        \
        const int var = (variant);           \
        switch (len) {                       \
            FGPE_FOR_EACH_N(FGPE_DISPATCH_CASE) \
        }                                    \
    }

template <typename T> int Launch<T>::adj(fgpe_ctx* c, const AdjArgs& a, bool y64, bool dens, bool epi)
{
#define CALL(N_, V_) detail::adj_n<T, N_, V_>(c, a, y64, dens, epi)
    FGPE_DISPATCH_X(c->n[0], c->var[FGPE_KCLASS_XSTAGE][0])
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported nx");
}

template <typename T> long long Launch<T>::xblocks(fgpe_ctx* c)
{
#define CALL(N_, V_) detail::x_groups<px_of<T, N_, V_>()>(c, N_)
    FGPE_DISPATCH_X(c->n[0], c->var[FGPE_KCLASS_XSTAGE][0])
#undef CALL
    return 0;
}

template <typename T> int Launch<T>::fft_x(fgpe_ctx* c, const void* in, void* out, int dir)
{
#define CALL(N_, V_) detail::fft_x_n<T, N_, V_>(c, in, out, dir)
    FGPE_DISPATCH(c->n[0], c->var[FGPE_KCLASS_FFTX][0])
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported nx");
}

template <typename T> int Launch<T>::strided(fgpe_ctx* c, void* data, int axis, int dir, int nch, int ncomp,
                                             const double* scale, const int* stop, bool half, const void* tab,
                                             int ninst, int slab, int nslab)
{
#define CALL(N_, V_) \
    detail::strided_n<T, N_, V_>(c, data, axis, dir, nch, ncomp, scale, stop, half, tab, ninst, slab, nslab)
    FGPE_DISPATCH(c->n[axis], c->var[FGPE_KCLASS_STRIDED][axis])
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported strided axis length");
}

template <typename T> int Launch<T>::last(fgpe_ctx* c, void* data, int axis, int nch, int ncomp,
                                          const double* scale, bool filter, const int* stop, bool half,
                                          const void* tab, int ninst, int slab, int nslab, bool vpx)
{
#define CALL(N_, V_) \
    detail::last_n<T, N_, V_>(c, data, axis, nch, ncomp, scale, filter, stop, half, tab, ninst, slab, nslab, vpx)
    FGPE_DISPATCH(c->n[axis], c->var[FGPE_KCLASS_LAST][axis])
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported last axis length");
}

template <typename T> int Launch<T>::xstage(fgpe_ctx* c, const XArgs<T>& a, int nc, bool dens, int mode,
                                            const void* tab, int ninst)
{
#define CALL(N_, V_) detail::xstage_n<T, N_, V_>(c, a, nc, dens, mode, tab, ninst)
    FGPE_DISPATCH_X(c->n[0], c->var[FGPE_KCLASS_XSTAGE][0])
#undef CALL
    return fail(FGPE_EUNSUP, "unsupported nx");
}

}  // namespace fgpe
