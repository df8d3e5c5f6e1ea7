"""Parity checks of the library built with the example user pointwise term
(libfgpe_user_example.so, fgpe_amd/csrc/user_term.hpp) against the oracle's
restatement of that term.  Run in a process of its own with FGPE_LIBRARY
set (tests/test_gpu_user_term.py), since a process loads one build of the
library.  Prints one line per check and exits non-zero on the first failure."""
import os
import sys

import numpy as np
import torch  # before the library: torch's HIP runtime must be the process's first

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from fgpe_amd import FGPE_IMAG_HEAVY_BALL, FGPE_IMAG_RK4_RENORM, FGPE_PREC_MIXED, FgpeError, configs  # noqa: E402
from fgpe_amd._lib import user_term_enabled, user_term_ids  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests.cases import Case, rel_max  # noqa: E402

USER = (0.8, 3.0, 0.4, 0.05)   # dg amplitude, dg frequency, K3, ramp per z cell


def say(msg):
    print(msg, flush=True)


def check(cond, msg):
    if not cond:
        say("FAIL " + msg)
        sys.exit(1)
    say("ok   " + msg)


def one_comp(c, h):
    vt = c.vtilde(h.get("a_dd", 0.0), h.get("dip", (0, 0, 1))) if h.get("a_dd") else None
    return orc.Ham(1, h["a"], vtilde=vt, user=USER)


def main():
    check(torch.cuda.is_available(), "HIP device visible")
    check(user_term_enabled(), "library reports a user term")
    c = Case(32, 32, 64, L=(12.0, 12.0, 16.0))
    h = dict(a=(0.7,), a_dd=0.5, dip=(0.0, 0.0, 1.0), user=USER, user_id=1)
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.5))
    psi = configs.gaussian_noisy(c.spec, 4, noise=0.02)

    say("apply_h 32x32x64")
    hmu_ref, hen_ref = orc.hpsi(c.grid, one_comp(c, h), 0.37, psi, V)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        hmu = torch.empty(c.shape, dtype=torch.complex128, device="cuda")
        hen = torch.empty_like(hmu)
        ctx.apply_h(0.37, hmu.data_ptr(), hen.data_ptr())
        ctx.sync()
        check(rel_max(hmu.cpu().numpy(), hmu_ref) < 1e-13, "H psi with the user term")
        check(rel_max(hen.cpu().numpy(), hen_ref) < 1e-13, "H_en psi with the user term")

    say("rk4 32x32x64 f64 and mixed")
    dt, n = 5e-4, 12
    ref = orc.rk4(c.grid, one_comp(c, h), psi, V, dt, n, t0=0.1)
    for prec, tol in ((0, 1e-12), (FGPE_PREC_MIXED, 1e-6)):
        with c.ctx(prec) as ctx:
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.rk4_steps(dt, n, t0=0.1)
            err = rel_max(ctx.download_psi(), ref)
        check(err < tol, f"RK4 precision {prec}: rel err {err:.2e} < {tol:.0e}")

    say("heavy ball 32x32x64")
    hh = dict(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0), user=(0.0, 0.0, 0.0, 0.02), user_id=1)
    ref, nref, href = orc.imag_heavy_ball(c.grid, orc.Ham(1, hh["a"], vtilde=c.vtilde(0.03), user=hh["user"]),
                                          psi, V, 2e-3, 0.9, 1.0, 1e-30, 60)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**hh)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(2e-3, 60, scheme=FGPE_IMAG_HEAVY_BALL, beta=0.9, npart=1.0, eps=1e-30,
                                    history=True)
        gpu = ctx.download_psi()
    check(done == nref == 60, "heavy ball ran 60 iterations")
    check(rel_max(gpu, ref) < 1e-11, "heavy-ball state with the user term")
    check(np.abs(hist - href).max() <= 1e-11 * np.abs(href).max(), "heavy-ball norm, mu, E history")

    say("heavy ball with a time-dependent user term, one call and three calls")
    ht = dict(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0), user=(0.6, 40.0, 0.0, 0.02), user_id=1)
    ref, nref, href = orc.imag_heavy_ball(c.grid, orc.Ham(1, ht["a"], vtilde=c.vtilde(0.03), user=ht["user"]),
                                          psi, V, 2e-3, 0.9, 1.0, 1e-30, 60)
    runs = {}
    for split in (1, 3):
        with c.ctx() as ctx:
            ctx.set_hamiltonian(**ht)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            hs = []
            for _ in range(split):
                d, hh_ = ctx.imag_steps(2e-3, 60 // split, scheme=FGPE_IMAG_HEAVY_BALL, beta=0.9, npart=1.0,
                                        eps=1e-30, history=True)
                check(d == 60 // split, f"call of {60 // split} iterations ran them all")
                hs.append(hh_)
            runs[split] = (ctx.download_psi(), np.concatenate(hs))
    for split, (st, hi) in runs.items():
        check(rel_max(st, ref) < 1e-11, f"{split} call(s): state at t = iteration dt against the oracle")
        check(np.abs(hi - href).max() <= 1e-11 * np.abs(href).max(), f"{split} call(s): history against the oracle")
    check(rel_max(runs[3][0], runs[1][0]) < 1e-12, "three calls of 20 equal one call of 60")

    say("RK4 imaginary time with a time-dependent user term, one call and two calls")
    hr = dict(a=(0.05,), user=(0.6, 40.0, 0.0, 0.0), user_id=1)
    ref = orc.rk4(c.grid, orc.Ham(1, hr["a"], user=hr["user"]), psi, V, 1e-3, 30, imaginary=True, npart=1.0)
    for split in (1, 2):
        with c.ctx() as ctx:
            ctx.set_hamiltonian(**hr)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            for _ in range(split):
                d, hh_ = ctx.imag_steps(1e-3, 30 // split, scheme=FGPE_IMAG_RK4_RENORM, npart=1.0, history=True)
            err = rel_max(ctx.download_psi(), ref)
        check(err < 1e-12, f"{split} call(s): RK4 imaginary time rel err {err:.2e}")
        check(np.all(np.isfinite(hh_[:, 0]) & (hh_[:, 0] > 0)) and np.all(np.isnan(hh_[:, 1:])),
              "RK4 history: norm before renormalisation, NaN in the mu and E slots")

    say("two components rk4")
    f = 1.0 / (4 * np.pi)
    V2 = np.stack([V, V + 0.1])
    p2 = np.stack([psi, configs.gaussian_noisy(c.spec, 8, offset=(-1.0, 0, 0), noise=0.02)])
    h2 = dict(ncomp=2, a=(100 * f, 90 * f, 80 * f), user=USER, user_id=1)
    ref = orc.rk4(c.grid, orc.Ham(2, h2["a"], user=USER), p2, V2, dt, 8)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h2)
        ctx.upload_psi(p2)
        ctx.upload_vext(V2)
        ctx.rk4_steps(dt, 8)
        err = rel_max(ctx.download_psi(), ref)
    check(err < 1e-12, f"two-component RK4 rel err {err:.2e}")

    say("a second term in the same library: two contexts, term 1 and term 2, stepped in turn")
    check(user_term_ids() == [1, 2], "the library lists user terms 1 and 2")
    DRIVE = (0.0, 0.0, 0.0, 0.0, 1.5, 2.0, 0.4, 0.3)   # A, gamma, kx, ky per cell
    h2d = dict(ncomp=2, a=(100 * f, 90 * f, 80 * f), user=DRIVE, user_id=2)
    ref2 = orc.rk4(c.grid, orc.Ham(2, h2d["a"], user=DRIVE, user_id=2), p2, V2, dt, 8, t0=0.05)
    ref1 = orc.rk4(c.grid, one_comp(c, h), psi, V, dt, 8, t0=0.05)
    with c.ctx() as c1, c.ctx() as c2:
        c1.set_hamiltonian(**h)
        c1.upload_psi(psi)
        c1.upload_vext(V)
        c2.set_hamiltonian(**h2d)
        c2.upload_psi(p2)
        c2.upload_vext(V2)
        for j in range(8):
            c1.rk4_steps(dt, 1, 0.05 + j * dt)
            c2.rk4_steps(dt, 1, 0.05 + j * dt)
        e1 = rel_max(c1.download_psi(), ref1)
        e2 = rel_max(c2.download_psi(), ref2)
    check(e1 < 1e-12, f"term 1 context rel err {e1:.2e}")
    check(e2 < 1e-12, f"term 2 context (two components, opposite signs) rel err {e2:.2e}")

    say("term 2 heavy ball, one component")
    hd = dict(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0), user=DRIVE, user_id=2)
    ref, nref, href = orc.imag_heavy_ball(c.grid, orc.Ham(1, hd["a"], vtilde=c.vtilde(0.03), user=DRIVE, user_id=2),
                                          psi, V, 2e-3, 0.9, 1.0, 1e-30, 40)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**hd)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(2e-3, 40, scheme=FGPE_IMAG_HEAVY_BALL, beta=0.9, npart=1.0, eps=1e-30,
                                    history=True)
        gpu = ctx.download_psi()
    check(done == nref == 40 and rel_max(gpu, ref) < 1e-11, "term 2 heavy-ball state")
    check(np.abs(hist - href).max() <= 1e-11 * np.abs(href).max(), "term 2 heavy-ball history")

    say("an id the library does not compile is refused")
    with c.ctx() as ctx:
        try:
            ctx.set_hamiltonian(a=(0.7,), user=USER, user_id=3)
            refused = False
        except FgpeError:
            refused = True
    check(refused, "fgpe_set_hamiltonian refuses user term 3")

    say("adjoint refuses the user term")
    with c.ctx() as ctx:
        ctx.set_hamiltonian(a=(0.7,), user=USER, user_id=1)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        try:
            ctx.adjoint_forward(dt, 4, 2)
            refused = False
        except FgpeError:
            refused = True
    check(refused, "fgpe_adjoint_forward returns an error")
    return 0


if __name__ == "__main__":
    sys.exit(main())
