"""The benchmark configurations at their full grid sizes.

Where the oracle finishes in seconds the device is compared with it at the
full size over a shortened horizon (1e-12); over the full horizon, where the
oracle would take minutes to hours, the checks are the properties the
equations fix independently of size: imaginary time lowers the energy at
fixed norm, real-time RK4 never raises the norm and, on smooth states,
conserves norm and energy to its truncation error, and the mixed-precision
form stays within 1e-6 relative energy of the FP64 form."""
import os

import numpy as np
import pytest

from fgpe_amd import FGPE_PREC_F64, FGPE_PREC_MIXED, FGPE_IMAG_RK4_RENORM, Context, configs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ctx_for(w, prec):
    g = w.grid
    kx, ky, kz = g.kaxes()
    c = Context(g.nx, g.ny, g.nz, prec)
    c.set_kspace(kx, ky, kz, g.dv)
    kw = dict(ncomp=w.ncomp, a=w.a)
    if w.a_dd:
        kw.update(a_dd=w.a_dd, dip=w.dip)
    if w.sc_U:
        kw.update(sc_U=w.sc_U, sc_R=w.sc_R, sc_dim=3)
    c.set_hamiltonian(**kw)
    c.upload_psi(w.initial_state())
    c.upload_vext(w.trap())
    return c


def oracle_grid(w):
    g = w.grid
    kx, ky, kz = g.kaxes()
    return orc.Grid(g.nx, g.ny, g.nz, kx, ky, kz, g.dv)


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def fixture(name):
    return np.load(os.path.join(GOLDEN, name))


def compare_to_fixture(fx, k, psi, obs, dv, tol=1e-12):
    """Checkpoint k of an oracle fixture (tests/golden/make_horizon.py):
    psi on its sampled points against max |psi|, sum |psi| dv, norm, mu, E."""
    sub = int(fx["sub"])
    e_psi = float(np.abs(psi[::sub, ::sub, ::sub] - fx["psi_s"][k]).max() / fx["psimax"][k])
    e_abs = abs(float(np.abs(psi).sum() * dv) - fx["absum"][k]) / fx["absum"][k]
    e_obs = np.abs(np.asarray(obs[:3]) - fx["obs"][k]) / np.abs(fx["obs"][k])
    print(f"  step {int(fx['steps'][k])}: psi {e_psi:.2e}, sum|psi| {e_abs:.2e}, norm/mu/E {e_obs}", flush=True)
    assert e_psi < tol and e_abs < tol and np.all(e_obs < tol), (k, e_psi, e_abs, e_obs)


def test_config1_imaginary_rk4_64cubed():
    """64^3 contact GPE, imaginary-time RK4 renormalised to N = 1, dt = 1e-3,
    over the full 2000 steps against the oracle (a fixture the oracle wrote,
    tests/golden/make_horizon.py): psi, sum |psi|, norm, mu and E within
    1e-12 at steps 500, 1000, 1500 and 2000, the energy falling between
    checkpoints at a norm of 1."""
    w = configs.contact_imag_64()
    print(f"config 1: {w.name}, {w.grid.shape}", flush=True)
    fx = fixture("config1_64cubed_2000.npz")
    energies = []
    with ctx_for(w, FGPE_PREC_F64) as c:
        for k, step in enumerate(fx["steps"]):
            c.imag_steps(w.dt, 500, scheme=FGPE_IMAG_RK4_RENORM, npart=1.0)
            o = c.observables(step * w.dt)
            compare_to_fixture(fx, k, c.download_psi(), o, w.grid.dv)
            assert abs(o[0] - 1.0) < 1e-12
            energies.append(o[2])
    assert all(e1 < e0 for e0, e1 in zip(energies, energies[1:])), energies


def test_config2_horizon_64x64x128():
    """The dipolar real-time physics (eps_dd = 1.3, cigar trap, dt = 5e-4)
    over its full 4000 steps on a 64 x 64 x 128 grid, against the oracle's
    run (a fixture, tests/golden/make_horizon.py) at steps 1000 ... 4000:
    FP64 within 1e-12 on psi, sum |psi|, norm, mu and E; the mixed form
    within 1e-6 of the FP64 oracle in energy (its psi error, which the
    instability amplifies from FP32 rounding, within 1e-3).  The oracle
    built with and without FMA contraction stays within 5e-14 of itself over
    these 4000 steps, so 1e-12 leaves room for rounding, not for a fault."""
    w = configs.dipolar_real(64, 64, 128, nsteps=4000)
    print(f"config 2 horizon: {w.name}", flush=True)
    fx = fixture("config2_64x64x128_4000.npz")
    every = int(fx["steps"][1] - fx["steps"][0])
    for prec in (FGPE_PREC_F64, FGPE_PREC_MIXED):
        with ctx_for(w, prec) as c:
            for k, step in enumerate(fx["steps"]):
                c.rk4_steps(w.dt, every, (step - every) * w.dt)
                o = c.observables(step * w.dt)
                if prec == FGPE_PREC_F64:
                    compare_to_fixture(fx, k, c.download_psi(), o, w.grid.dv)
                else:
                    sub = int(fx["sub"])
                    e_psi = float(np.abs(c.download_psi()[::sub, ::sub, ::sub] - fx["psi_s"][k]).max()
                                  / fx["psimax"][k])
                    e_en = abs(o[2] - fx["obs"][k][2]) / abs(fx["obs"][k][2])
                    print(f"  mixed step {int(step)}: psi {e_psi:.2e}, E {e_en:.2e}", flush=True)
                    # the energy is the bound; psi's error grows with the unstable
                    # dynamics (measured 1.6e-7, 7.6e-7, 2.9e-6, 8.0e-5 at steps 1000-4000)
                    assert e_en < 1e-6 and e_psi < 1e-3, (k, e_en, e_psi)


def test_config2_dipolar_real_time_full_grid():
    """128 x 128 x 256 dipolar real-time RK4 (eps_dd = 1.3), dt = 5e-4: two
    steps against the oracle at the full size; over 400 steps the norm never
    grows (|R(iz)| <= 1 for RK4 at |z| <= 2 sqrt 2, and the seeded white noise
    puts a little weight on the damped high-k modes), and the mixed form stays
    within 1e-6 of FP64 in energy."""
    w = configs.dipolar_real(128, 128, 256, nsteps=400)
    print(f"config 2: {w.name}", flush=True)
    og = oracle_grid(w)
    orc.set_threads(16)
    vt = orc.vdd(og, w.a_dd, w.dip)
    ref = orc.rk4(og, orc.Ham(1, w.a, vtilde=vt), w.initial_state(), w.trap(), w.dt, 2)
    out = {}
    for prec in (FGPE_PREC_F64, FGPE_PREC_MIXED):
        with ctx_for(w, prec) as c:
            o0 = c.observables(0.0)
            c.rk4_steps(w.dt, 2)
            if prec == FGPE_PREC_F64:
                assert rel(c.download_psi(), ref) < 1e-12
            c.rk4_steps(w.dt, w.nsteps - 2, 2 * w.dt)
            o1 = c.observables(w.nsteps * w.dt)
            out[prec] = (o0, o1, c.download_psi())
    o0, o1, p64 = out[FGPE_PREC_F64]
    assert o1[0] <= o0[0] and o0[0] - o1[0] < 1e-4 * o0[0]   # norm: non-increasing, small
    m0, m1, pm = out[FGPE_PREC_MIXED]
    assert abs(m1[2] - o1[2]) < 1e-6 * abs(o1[2])
    assert np.sqrt(np.sum(np.abs(pm - p64) ** 2) / np.sum(np.abs(p64) ** 2)) < 1e-5


def test_config3_oracle_steps_256cubed():
    """Two coupled components at the full 256^3 with the soft-core term on
    n1 + n2: two RK4 steps against the oracle, FP64 within 1e-12 and the
    mixed form (the configuration's FP32 bulk) within 1e-6, on both
    components' psi and on each component's norm and energy."""
    w = configs.twocomp_softcore_256(256, nsteps=2)
    print(f"config 3: {w.name}, oracle {w.nsteps} steps", flush=True)
    og = oracle_grid(w)
    orc.set_threads(16)
    vt = orc.vsoftcore(og, 3, w.sc_R, w.sc_U)
    ham = orc.Ham(2, w.a, vtilde=vt)
    psi0, V = w.initial_state(), w.trap()
    ref = orc.rk4(og, ham, psi0, V, w.dt, w.nsteps)
    ref_obs = []
    hmu, hen = orc.hpsi(og, ham, w.nsteps * w.dt, ref, V)
    for comp in range(2):
        nrm = orc.norm(og, ref[comp])
        ref_obs.append((nrm,) + orc.obs2(og, nrm, ref[comp], hmu[comp], hen[comp]))
    del hmu, hen
    for prec, tol in ((FGPE_PREC_F64, 1e-12), (FGPE_PREC_MIXED, 1e-6)):
        with ctx_for(w, prec) as c:
            c.rk4_steps(w.dt, w.nsteps)
            got = c.download_psi()
            o = c.observables(w.nsteps * w.dt)
        for comp in range(2):
            e_psi = rel(got[comp], ref[comp])
            e_obs = [abs(o[3 * comp + j] - ref_obs[comp][j]) / abs(ref_obs[comp][j]) for j in (0, 2)]
            print(f"  precision {prec} component {comp}: psi {e_psi:.2e}, norm / E {e_obs}", flush=True)
            assert e_psi < tol and max(e_obs) < tol, (prec, comp, e_psi, e_obs)


def test_config3_two_component_softcore_256cubed():
    """Two coupled components at 256^3 with the soft-core term: 50 RK4 steps
    conserve each component's norm and the total energy, in FP64 and in the
    mixed form (the configuration's FP32 bulk), which agree to 1e-6."""
    w = configs.twocomp_softcore_256(256, nsteps=50)
    print(f"config 3: {w.name}", flush=True)
    out = {}
    for prec in (FGPE_PREC_F64, FGPE_PREC_MIXED):
        with ctx_for(w, prec) as c:
            o0 = c.observables(0.0)
            c.rk4_steps(w.dt, w.nsteps)
            o1 = c.observables(w.nsteps * w.dt)
            out[prec] = (o0, o1)
    for prec, (o0, o1) in out.items():
        for comp in range(2):
            assert abs(o1[3 * comp] - o0[3 * comp]) < 1e-8 * o0[3 * comp], (prec, comp)
        e0 = o0[0] * o0[2] + o0[3] * o0[5]
        e1 = o1[0] * o1[2] + o1[3] * o1[5]
        assert abs(e1 - e0) < 1e-6 * abs(e0), prec
    f, m = out[FGPE_PREC_F64][1], out[FGPE_PREC_MIXED][1]
    for comp in range(2):
        assert abs(m[3 * comp + 2] - f[3 * comp + 2]) < 1e-6 * abs(f[3 * comp + 2])


def test_mixed_long_horizon_dipolar_256cubed():
    """The benchmark workload itself, 256^3 dipolar real-time RK4, over 2000
    steps (one time unit) in both forms: the mixed form (FP32 bulk, stages 2
    and 3 rebuilding psi_n from the stage states) stays within 1e-6 of the
    FP64 form in energy and within 1e-4 in L2 norm of psi, and within 1e-6 of
    its norm (RK4 damps the seeded high-k noise in both forms alike)."""
    w = configs.dipolar_real(256, 256, 256, nsteps=2000)
    print(f"{w.name}, {w.nsteps} steps", flush=True)
    out = {}
    for prec in (FGPE_PREC_F64, FGPE_PREC_MIXED):
        with ctx_for(w, prec) as c:
            o0 = c.observables(0.0)
            c.rk4_steps(w.dt, w.nsteps)
            out[prec] = (o0, c.observables(w.nsteps * w.dt), c.download_psi())
    f0, f1, pf = out[FGPE_PREC_F64]
    m0, m1, pm = out[FGPE_PREC_MIXED]
    de = abs(m1[2] - f1[2]) / abs(f1[2])
    l2 = float(np.sqrt(np.sum(np.abs(pm - pf) ** 2) / np.sum(np.abs(pf) ** 2)))
    print(f"mixed vs FP64 after {w.nsteps} steps: rel energy {de:.3e}, L2 psi {l2:.3e}, "
          f"norm drift FP64 {abs(f1[0] - f0[0]) / f0[0]:.3e} mixed {abs(m1[0] - m0[0]) / m0[0]:.3e}", flush=True)
    assert de < 1e-6
    assert l2 < 1e-4
    assert abs(m1[0] - f1[0]) < 1e-6 * f1[0]
