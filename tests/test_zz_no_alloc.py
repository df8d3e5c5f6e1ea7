"""fgpe_imag_steps_noalloc, the imaginary-time entry point that allocates
nothing (its history lives in a fixed device array of the library), against
fgpe_imag_steps on twin contexts, through the C ABI: two calls in a row, the
second with a smaller grid and history, for the heavy-ball and the RK4
schemes.  States, iteration counts and histories must be identical, and the
no-allocation context's device memory must not grow."""
import ctypes as C

import numpy as np
import pytest

from fgpe_amd import FGPE_IMAG_HEAVY_BALL, FGPE_IMAG_RK4_RENORM, configs
from fgpe_amd import _lib
from tests.cases import Case

pytestmark = pytest.mark.gpu


def _run(L, ctx, fn, scheme, dt, beta, eps, n):
    hist = np.zeros((n, 3))
    done = C.c_int(-1)
    npart = np.ones(2)
    rc = fn(ctx.h, scheme, dt, beta, npart.ctypes.data_as(_lib._dp), eps, n, hist.ctypes.data_as(_lib._dp),
            C.byref(done))
    _lib.check(rc)
    return done.value, hist[:done.value]


def _twins(case, seed):
    out = []
    for _ in range(2):
        c = case.ctx(0)
        c.set_hamiltonian(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0))
        c.upload_psi(configs.gaussian_noisy(case.spec, seed, width=1.2))
        c.upload_vext(configs.harmonic_trap(case.spec, (1.0, 1.0, 1.3)))
        out.append(c)
    return out


@pytest.mark.parametrize("scheme", [FGPE_IMAG_HEAVY_BALL, FGPE_IMAG_RK4_RENORM])
def test_noalloc_matches_imag_steps(torch_cuda, scheme):
    L = _lib.load()
    dt, beta, eps = 2e-3, 0.9, 1e-30
    for shape, n, seed in (((64, 64, 64), 40, 51), ((32, 32, 32), 15, 52)):
        case = Case(*shape, L=(12.0, 12.0, 12.0))
        ref, mine = _twins(case, seed)
        try:
            print(f"scheme {scheme}, {shape}, {n} iterations then {n // 2} more", flush=True)
            bytes0 = mine.device_bytes()
            for m in (n, n // 2):      # a continuation on the same context, with a shorter history
                d0, h0 = _run(L, ref, L.fgpe_imag_steps, scheme, dt, beta, eps, m)
                d1, h1 = _run(L, mine, L.fgpe_imag_steps_noalloc, scheme, dt, beta, eps, m)
                assert d0 == d1 == m
                assert np.array_equal(h0, h1, equal_nan=True)
                assert np.array_equal(ref.download_psi(), mine.download_psi())
            assert mine.device_bytes() == bytes0
            assert ref.device_bytes() > bytes0   # the allocating form kept a history buffer
        finally:
            ref.close()
            mine.close()


def test_noalloc_rejects_too_long_history(torch_cuda):
    L = _lib.load()
    case = Case(32, 32, 32, L=(12.0, 12.0, 12.0))
    ctx = _twins(case, 53)[0]
    try:
        n = 16384 + 1
        hist = np.zeros((n, 3))
        done = C.c_int(-1)
        npart = np.ones(2)
        rc = L.fgpe_imag_steps_noalloc(ctx.h, FGPE_IMAG_HEAVY_BALL, 2e-3, 0.9, npart.ctypes.data_as(_lib._dp),
                                       0.0, n, hist.ctypes.data_as(_lib._dp), C.byref(done))
        assert rc == -1
        assert b"FGPE_NOALLOC_MAX_STEPS" in L.fgpe_last_error()
        assert done.value == -1
    finally:
        ctx.close()
