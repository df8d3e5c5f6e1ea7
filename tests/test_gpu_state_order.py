"""Ordering of state changes against queued work on the context's stream,
and set-up calls that must leave a consistent context behind:

- a trap uploaded straight after an asynchronous fgpe_rk4_steps must not
  reach the device before the queued steps have read the old trap;
- a k space set after a soft-core Hamiltonian rebuilds the soft-core table;
- a group autotune in the middle of a batched heavy-ball run ends that run
  with its pending renormalisation applied, so the next run starts afresh
  from the normalised states (equal, bit for bit, to a fresh run from the
  same states)."""
import numpy as np
import pytest

from fgpe_amd import configs
from oracle import oracle as orc
from tests.cases import Case, rel_max

pytestmark = pytest.mark.gpu


def test_upload_vext_after_async_steps():
    c = Case(64, 64, 128, L=(12.0, 12.0, 16.0))
    h = dict(a=(0.4,), a_dd=0.2, dip=(0.0, 0.0, 1.0))
    V1 = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.5))
    V2 = configs.harmonic_trap(c.spec, (1.3, 0.8, 0.6))
    psi = configs.gaussian_noisy(c.spec, 31, noise=0.02)
    dt, n1, n2 = 5e-4, 30, 6
    orc.set_threads(16)
    ham = orc.Ham(1, h["a"], vtilde=c.vtilde(0.2))
    ref = orc.rk4(c.grid, ham, psi, V1, dt, n1)
    ref = orc.rk4(c.grid, ham, ref, V2, dt, n2, t0=n1 * dt)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V1)
        ctx.rk4_steps(dt, n1)          # queued, not waited for
        ctx.upload_vext(V2)            # must wait for the queued steps
        ctx.rk4_steps(dt, n2, n1 * dt)
        err = rel_max(ctx.download_psi(), ref)
    print(f"rk4 / upload_vext / rk4: rel err {err:.2e}")
    assert err < 1e-12


def test_set_kspace_rebuilds_softcore_table():
    c1 = Case(32, 32, 64, L=(12.0, 12.0, 16.0))
    c2 = Case(32, 32, 64, L=(10.0, 11.0, 14.0))   # the k space the run uses
    h = dict(a=(0.3,), sc_U=3.0, sc_R=1.1, sc_dim=3)
    V = configs.harmonic_trap(c2.spec, (1.0, 1.0, 0.7))
    psi = configs.gaussian_noisy(c2.spec, 32, noise=0.02)
    dt, n = 5e-4, 10
    ref = orc.rk4(c2.grid, orc.Ham(1, h["a"], vtilde=c2.vtilde(sc=(3, 1.1, 3.0))), psi, V, dt, n)
    with c1.ctx() as ctx:                            # k space of the first box
        ctx.set_hamiltonian(**h)                     # table built from it
        ctx.set_kspace(c2.kx, c2.ky, c2.kz, c2.dv)   # must rebuild the table
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.rk4_steps(dt, n)
        err = rel_max(ctx.download_psi(), ref)
    print(f"soft-core RK4 after a new k space: rel err {err:.2e}")
    assert err < 1e-12


def test_group_autotune_mid_run_starts_fresh(torch_cuda):
    from fgpe_amd import Group
    side = torch_cuda.cuda.Stream()
    case = Case(32, 32, 32, L=(12.0, 12.0, 12.0))
    insts = [(dict(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0)), (1.0, 1.0, 1.3), 41),
             (dict(a=(0.08,), a_dd=0.02, dip=(1.0, 0.0, 0.0)), (1.2, 1.0, 1.0), 42)]
    dt, beta, n1, n2 = 2e-3, 0.9, 15, 10

    def make():
        out = []
        for h, om, seed in insts:
            c = case.ctx(0)
            c.set_stream(side.cuda_stream)
            c.set_hamiltonian(**h)
            c.upload_psi(configs.gaussian_noisy(case.spec, seed, width=1.2))
            c.upload_vext(configs.harmonic_trap(case.spec, om))
            out.append(c)
        return out

    def pin_variants(ctxs):
        for c in ctxs:
            for k, axis in ((0, 0), (1, 1), (2, 2)):
                c.set_variant(k, axis, 0)

    a, b = make(), make()
    ga, gb = Group(a), Group(b)
    try:
        ga.imag_steps(dt, n1, beta=beta, npart=1.0, eps=0.0)
        ga.autotune(2)                     # ends the run, renormalised
        pin_variants(a)
        ga.imag_steps(dt, n2, beta=beta, npart=1.0, eps=0.0)
        ga.imag_finish()

        gb.imag_steps(dt, n1, beta=beta, npart=1.0, eps=0.0)
        gb.imag_finish()
        for c in b:                        # a fresh run from the same states
            c.upload_psi(c.download_psi())
        pin_variants(b)
        gb.imag_steps(dt, n2, beta=beta, npart=1.0, eps=0.0)
        gb.imag_finish()
        for i, (x, y) in enumerate(zip(a, b)):
            px, py = x.download_psi(), y.download_psi()
            nrm = float(np.sum(np.abs(px) ** 2) * case.dv)
            print(f"instance {i}: max diff {np.abs(px - py).max():.3e}, norm {nrm:.15f}")
            assert np.array_equal(px, py)
            assert abs(nrm - 1.0) < 1e-12
    finally:
        ga.close()
        gb.close()
        for c in a + b:
            c.close()
