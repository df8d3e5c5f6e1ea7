"""Parity of the HIP path (called through the C ABI) with the CPU oracle on
seeded inputs.  FP64 form: 1e-12 relative or better per quantity; mixed
form: the error bounds written next to each assertion."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.cases import Case, rel_max

pytestmark = pytest.mark.gpu


def dev(torch, arr):
    return torch.from_numpy(np.ascontiguousarray(arr)).cuda()


FFT_SHAPES = [(32, 32, 32), (64, 32, 96), (128, 64, 32), (96, 192, 32), (512, 32, 32),
              (256, 32, 64), (64, 64, 1), (256, 128, 1),
              (32, 256, 64), (16, 64, 256)]   # 256-long y and z: the split-exchange FP64 pass


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_fft3d_f64(torch_cuda, shape):
    torch = torch_cuda
    nx, ny, nz = shape
    c = Case(nx, ny, nz)
    x = c.rand_psi(11)
    ref_f = np.fft.fftn(x)
    ref_b = np.fft.ifftn(x) * x.size
    with c.ctx() as ctx:
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        xi = dev(torch, x)
        out = torch.empty_like(xi)
        ctx.fft3d_f64(xi.data_ptr(), out.data_ptr(), -1)
        torch.cuda.synchronize()
        err = rel_max(out.cpu().numpy(), ref_f)
        assert err < 1e-14, err
        ctx.fft3d_f64(xi.data_ptr(), out.data_ptr(), +1)
        torch.cuda.synchronize()
        err = rel_max(out.cpu().numpy(), ref_b)
        assert err < 1e-14, err
        # in place, and against the oracle's transform
        y = xi.clone()
        ctx.fft3d_f64(y.data_ptr(), y.data_ptr(), -1)
        torch.cuda.synchronize()
        err = rel_max(y.cpu().numpy(), orc.fft3d(x, -1))
        assert err < 1e-14, err


@pytest.mark.parametrize("shape", [(32, 32, 32), (128, 64, 32), (256, 32, 64)])
def test_fft3d_f32(torch_cuda, shape):
    torch = torch_cuda
    c = Case(*shape)
    x = c.rand_psi(12)
    with c.ctx() as ctx:
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        xi = dev(torch, x.astype(np.complex64))
        out = torch.empty_like(xi)
        ctx.fft3d_f32(xi.data_ptr(), out.data_ptr(), -1)
        torch.cuda.synchronize()
        # FP32 FFT: error ~ eps32 * log2(N)
        err = rel_max(out.cpu().numpy().astype(np.complex128), np.fft.fftn(x))
        assert err < 2e-6, err


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 32, 128), (128, 128, 1)])
def test_laplacian_f64(torch_cuda, shape):
    torch = torch_cuda
    c = Case(*shape, L=(16.0, 12.0, 20.0))
    x = c.rand_psi(13)
    ref = orc.laplace(c.grid, x)
    with c.ctx() as ctx:
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        xi = dev(torch, x)
        out = torch.empty_like(xi)
        ctx.laplacian_f64(xi.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        err = rel_max(out.cpu().numpy(), ref)
        assert err < 1e-13, err


HAMS = {
    "contact": dict(a=(0.7,)),
    "quench": dict(a=(0.7,), quench=(0.5, 0.9, 0.3, 0.1, 0.2, 0.4)),
    "dipolar": dict(a=(1.2,), a_dd=1.5, dip=(0.0, 0.6, 0.8)),
    "softcore": dict(a=(0.4,), sc_U=5.0, sc_R=1.5, sc_dim=3),
    "dip+sc": dict(a=(0.4,), a_dd=0.6, dip=(0.0, 0.0, 1.0), sc_U=2.0, sc_R=1.2, sc_dim=3),
    "moving": dict(a=(0.7,), a_dd=0.5, dip=(0.0, 0.0, 1.0), vpx=(1.3,)),
}


def oracle_ham(c, h):
    vt = None
    if h.get("a_dd") or h.get("sc_U"):
        vt = c.vtilde(h.get("a_dd", 0.0), h.get("dip", (0, 0, 1)),
                      (h.get("sc_dim", 3), h.get("sc_R"), h.get("sc_U")) if h.get("sc_U") else None)
    return orc.Ham(1, h["a"], quench=h.get("quench"), vtilde=vt, vpx=h.get("vpx", (0.0, 0.0)))


@pytest.mark.parametrize("name", list(HAMS))
def test_apply_h_and_observables(torch_cuda, name):
    torch = torch_cuda
    h = HAMS[name]
    c = Case(32, 64, 32, L=(12.0, 16.0, 12.0))
    psi = c.rand_psi(21, 0.2)
    V = c.rand_psi(22, 3.0).real + 1j * (0.1 * c.rand_psi(23).real if name == "quench" else 0.0)
    t = 0.15
    hmu_ref, hen_ref = orc.hpsi(c.grid, oracle_ham(c, h), t, psi, V)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        hmu = torch.empty(c.shape, dtype=torch.complex128, device="cuda")
        hen = torch.empty_like(hmu)
        ctx.apply_h(t, hmu.data_ptr(), hen.data_ptr())
        ctx.sync()
        err = rel_max(hmu.cpu().numpy(), hmu_ref)
        assert err < 1e-13, err
        err = rel_max(hen.cpu().numpy(), hen_ref)
        assert err < 1e-13, err
        obs = ctx.observables(t)
    nrm = orc.norm(c.grid, psi)
    mu, en = orc.obs2(c.grid, nrm, psi, hmu_ref, hen_ref)
    assert abs(obs[0] - nrm) < 1e-13 * nrm
    assert abs(obs[1] - mu) < 1e-12 * abs(mu)
    assert abs(obs[2] - en) < 1e-12 * abs(en)


def run_both(c, h, psi, V, dt, nsteps, precision=0, imaginary=False):
    from fgpe_amd import Context  # noqa: F401
    with c.ctx(precision) as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        if imaginary:
            ctx.imag_steps(dt, nsteps)
        else:
            ctx.rk4_steps(dt, nsteps)
        gpu = ctx.download_psi()
        obs = ctx.observables(nsteps * dt)
    return gpu, obs


@pytest.mark.parametrize("name", ["contact", "dipolar", "softcore", "quench", "moving"])
def test_rk4_real_time_f64(name):
    h = dict(HAMS[name])
    c = Case(32, 32, 64, L=(10.0, 10.0, 20.0))
    from fgpe_amd import configs
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.25))
    if name == "quench":
        V = V + 0.05j * c.rand_psi(5).real
    psi = configs.normalise(np.exp(-0.5 * np.abs(configs.harmonic_trap(c.spec, (1, 1, 0.5)))) +
                            0.01 * c.rand_psi(4), c.dv)
    dt, n = 5e-4, 12
    gpu, obs = run_both(c, h, psi, V, dt, n)
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, dt, n)
    err = rel_max(np.abs(gpu), np.abs(ref))
    assert err < 1e-12, err
    err = rel_max(gpu, ref)
    assert err < 1e-12, err
    hmu, hen = orc.hpsi(c.grid, oracle_ham(c, h), n * dt, ref, V)
    nrm = orc.norm(c.grid, ref)
    mu, en = orc.obs2(c.grid, nrm, ref, hmu, hen)
    assert abs(obs[0] - nrm) < 1e-12 * nrm
    assert abs(obs[2] - en) < 1e-12 * abs(en)


def test_rk4_imag_renorm_f64():
    from fgpe_amd import configs
    c = Case(32, 32, 32, L=(16.0, 16.0, 16.0))
    h = dict(a=(100.0 / (4 * np.pi),))
    V = configs.harmonic_trap(c.spec)
    psi = configs.gaussian_noisy(c.spec, 1)
    dt, n = 1e-3, 30
    gpu, obs = run_both(c, h, psi, V, dt, n, imaginary=True)
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, dt, n, imaginary=True, npart=1.0)
    err = rel_max(gpu, ref)
    assert err < 1e-12, err
    assert abs(obs[0] - 1.0) < 1e-13


def test_mixed_precision_dipolar_error():
    """Mixed form against the FP64 GPU form: relative energy error <= 1e-6."""
    from fgpe_amd import configs
    w = configs.dipolar_real(64, 64, 128, nsteps=200)
    psi, V = w.initial_state(), w.trap()
    c = Case(64, 64, 128, L=(16.0, 16.0, 32.0))
    h = dict(a=w.a, a_dd=w.a_dd, dip=w.dip)
    g64, o64 = run_both(c, h, psi, V, w.dt, w.nsteps, precision=0)
    g32, o32 = run_both(c, h, psi, V, w.dt, w.nsteps, precision=1)
    l2 = float(np.sqrt(np.sum(np.abs(g32 - g64) ** 2) / np.sum(np.abs(g64) ** 2)))
    de = abs(o32[2] - o64[2]) / abs(o64[2])
    print(f"mixed vs f64 after {w.nsteps} steps: L2 psi {l2:.3e}, rel energy {de:.3e}")
    assert de < 1e-6
    assert l2 < 1e-4


def test_mixed_imaginary_rk4_renormalised():
    """The mixed form's imaginary-time RK4 with renormalisation (real stage
    coefficients, the renormalisation scale folded into the next step's stage
    states, from which stages 2 and 3 rebuild psi_n): 40 steps of a dipolar
    state within 1e-6 of the FP64 oracle, the norm held at N = 1."""
    from fgpe_amd import FGPE_IMAG_RK4_RENORM, FGPE_PREC_MIXED, configs
    c = Case(32, 32, 64, L=(12.0, 12.0, 16.0))
    h = dict(a=(0.6,), a_dd=0.4, dip=(0.0, 0.0, 1.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.5))
    psi = configs.gaussian_noisy(c.spec, 11, noise=0.02)
    dt, n = 1e-3, 40
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, dt, n, imaginary=True, npart=1.0)
    with c.ctx(FGPE_PREC_MIXED) as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.imag_steps(dt, n, scheme=FGPE_IMAG_RK4_RENORM, npart=1.0)
        gpu = ctx.download_psi()
        obs = ctx.observables()
    err = rel_max(gpu, ref)
    assert err < 1e-6, err
    assert abs(obs[0] - 1.0) < 1e-12, obs[0]


def test_heavy_ball_one_component_f64():
    """The reference's imaginary-time scheme, convergence test on the device."""
    from fgpe_amd import FGPE_IMAG_HEAVY_BALL, configs
    c = Case(32, 32, 32, L=(12.0, 12.0, 12.0))
    h = dict(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 1.3))
    psi = configs.gaussian_noisy(c.spec, 6, width=1.2)
    dt, beta, npart, eps, n = 2e-3, 0.9, 1.0, 1e-7, 3000
    ref, nref, href = orc.imag_heavy_ball(c.grid, oracle_ham(c, h), psi, V, dt, beta, npart, eps, n)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(dt, n, scheme=FGPE_IMAG_HEAVY_BALL, beta=beta, npart=npart,
                                    eps=eps, history=True)
        gpu = ctx.download_psi()
    assert nref < n, "reference run should converge within the step budget"
    assert done == nref, (done, nref)
    err = rel_max(gpu, ref)
    assert err < 1e-11, err
    assert np.abs(hist - href).max() <= 1e-11 * np.abs(href).max()


def test_heavy_ball_moving_frame_f64():
    """Heavy ball with the moving-frame term (alg_calcHpsi applies it every
    iteration): same converging iteration, state and history as the oracle."""
    from fgpe_amd import FGPE_IMAG_HEAVY_BALL, configs
    c = Case(32, 32, 32, L=(12.0, 12.0, 12.0))
    h = dict(a=(0.05,), a_dd=0.03, dip=(0.0, 0.0, 1.0), vpx=(0.8,))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 1.3))
    psi = configs.gaussian_noisy(c.spec, 6, width=1.2)
    dt, beta, npart, eps, n = 2e-3, 0.9, 1.0, 1e-7, 3000
    ref, nref, href = orc.imag_heavy_ball(c.grid, oracle_ham(c, h), psi, V, dt, beta, npart, eps, n)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(dt, n, scheme=FGPE_IMAG_HEAVY_BALL, beta=beta, npart=npart,
                                    eps=eps, history=True)
        gpu = ctx.download_psi()
    assert nref < n
    assert done == nref, (done, nref)
    err = rel_max(gpu, ref)
    assert err < 1e-11, err
    assert np.abs(hist - href).max() <= 1e-11 * np.abs(href).max()


def two_comp_case():
    from fgpe_amd import configs
    c = Case(32, 32, 64, L=(12.0, 12.0, 16.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.5))
    p1 = configs.gaussian_noisy(c.spec, 7, offset=(1.0, 0, 0), noise=0.02)
    p2 = configs.gaussian_noisy(c.spec, 8, offset=(-1.0, 0, 0), noise=0.02)
    return c, np.stack([V, V + 0.1]), np.stack([p1, p2 * np.exp(0.3j)])


def two_comp_ham(c, h):
    vt = None
    if h.get("sc_U"):
        vt = c.vtilde(0.0, sc=(h["sc_dim"], h["sc_R"], h["sc_U"]))
    return orc.Ham(2, h["a"], vtilde=vt, vpx=h.get("vpx", (0.0, 0.0)))


@pytest.mark.parametrize("name", ["contact", "softcore", "moving"])
def test_two_component_rk4_f64(name):
    c, V, psi = two_comp_case()
    f = 1.0 / (4 * np.pi)
    h = dict(ncomp=2, a=(100 * f, 100 * f, 80 * f))
    if name == "softcore":
        h.update(sc_U=5.0, sc_R=1.5, sc_dim=3)
    if name == "moving":   # two-component RK4 uses the MU form only: no moving-frame term
        h.update(vpx=(1.1, -0.6))
    dt, n = 5e-4, 10
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.rk4_steps(dt, n)
        gpu = ctx.download_psi()
        obs = ctx.observables(n * dt)
    ref = orc.rk4(c.grid, two_comp_ham(c, h), psi, V, dt, n)
    err = rel_max(gpu, ref)
    assert err < 1e-12, err
    if name == "softcore":
        # the mixed form on the same case: both components' stage states
        # rebuilt at stages 2 and 3, the coupling read from the other
        # component's stage state
        from fgpe_amd import FGPE_PREC_MIXED
        with c.ctx(FGPE_PREC_MIXED) as ctx:
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.rk4_steps(dt, n)
            gm = ctx.download_psi()
        errm = rel_max(gm, ref)
        assert errm < 1e-6, errm
    hmu, hen = orc.hpsi(c.grid, two_comp_ham(c, h), n * dt, ref, V)
    for k in range(2):
        nrm = orc.norm(c.grid, ref[k])
        mu, en = orc.obs2(c.grid, nrm, ref[k], hmu[k], hen[k])
        assert abs(obs[3 * k] - nrm) < 1e-12 * nrm
        assert abs(obs[3 * k + 1] - mu) < 1e-12 * abs(mu)
        assert abs(obs[3 * k + 2] - en) < 1e-12 * abs(en)


def test_two_component_heavy_ball_f64():
    from fgpe_amd import FGPE_IMAG_HEAVY_BALL
    c, V, psi = two_comp_case()
    h = dict(ncomp=2, a=(0.02, 0.03, 0.01))
    dt, beta, npart, eps, n = 2e-3, 0.9, 1.0, 1e-9, 60
    ref, nref, href = orc.imag_heavy_ball(c.grid, two_comp_ham(c, h), psi, V, dt, beta, npart, eps, n)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(dt, n, scheme=FGPE_IMAG_HEAVY_BALL, beta=beta, npart=npart,
                                    eps=eps, history=True)
        gpu = ctx.download_psi()
    assert done == nref
    err = rel_max(gpu, ref)
    assert err < 1e-11, err
    assert np.abs(hist - href).max() <= 1e-11 * np.abs(href).max()


def test_two_component_mixed_error():
    """Mixed form of the two-component soft-core step against FP64."""
    c, V, psi = two_comp_case()
    f = 1.0 / (4 * np.pi)
    h = dict(ncomp=2, a=(100 * f, 100 * f, 80 * f), sc_U=5.0, sc_R=1.5, sc_dim=3)
    res = []
    for prec in (0, 1):
        with c.ctx(prec) as ctx:
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.rk4_steps(5e-4, 100)
            res.append((ctx.download_psi(), ctx.observables(0.05)))
    (g64, o64), (g32, o32) = res
    l2 = float(np.sqrt(np.sum(np.abs(g32 - g64) ** 2) / np.sum(np.abs(g64) ** 2)))
    de = max(abs(o32[2] - o64[2]) / abs(o64[2]), abs(o32[5] - o64[5]) / abs(o64[5]))
    print(f"two-component mixed vs f64: L2 psi {l2:.3e}, rel energy {de:.3e}")
    assert de < 1e-6 and l2 < 1e-4


def test_px_and_cos_diagnostics():
    """<Px> and <cos> of the reference's status rows, against numpy."""
    c = Case(32, 64, 32, L=(12.0, 16.0, 10.0))
    psi = c.rand_psi(31, 0.3) * np.exp(1j * 2 * np.pi * 2 * np.arange(32) / 32)[None, None, :]
    x = -6.0 + (np.arange(32) + 0.5) * 12.0 / 32
    with c.ctx() as ctx:
        ctx.set_hamiltonian(a=(0.0,))
        ctx.upload_psi(psi)
        px = ctx.calc_px()
        cs = ctx.calc_cos(x, 12.0)
    F = np.fft.fftn(psi)
    ref_px = c.dv * float(np.sum(c.kx[None, None, :] * np.abs(F) ** 2))
    ref_cos = c.dv * float(np.sum(np.cos(2 * np.pi * x / 12.0)[None, None, :] * np.abs(psi) ** 2))
    assert abs(px - ref_px) < 1e-12 * np.sum(np.abs(F) ** 2) * c.dv * np.abs(c.kx).max()
    assert abs(cs - ref_cos) < 1e-12 * c.dv * np.sum(np.abs(psi) ** 2)


@pytest.mark.parametrize("shape", [(64, 128, 256), (128, 64, 64), (256, 32, 32)])
def test_variant1_plans_parity(torch_cuda, shape):
    """The second plan variant of every kernel class against numpy / the oracle."""
    torch = torch_cuda
    from fgpe_amd import configs
    c = Case(*shape, L=(10.0, 12.0, 20.0))
    x = c.rand_psi(51)
    with c.ctx() as ctx:
        for k in range(5):
            for axis in range(3):
                ctx.set_variant(k, axis, 1)
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        xi = dev(torch, x)
        out = torch.empty_like(xi)
        ctx.fft3d_f64(xi.data_ptr(), out.data_ptr(), -1)
        torch.cuda.synchronize()
        err = rel_max(out.cpu().numpy(), np.fft.fftn(x))
        assert err < 1e-14, err
    h = dict(a=(0.9,), a_dd=0.8, dip=(0.0, 0.0, 1.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.25))
    psi = configs.normalise(np.exp(-0.5 * np.abs(configs.harmonic_trap(c.spec, (1, 1, 0.5)))) +
                            0.01 * c.rand_psi(52), c.dv)
    with c.ctx() as ctx:
        for k in range(5):
            for axis in range(3):
                ctx.set_variant(k, axis, 1)
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.rk4_steps(5e-4, 3)
        gpu = ctx.download_psi()
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 5e-4, 3)
    err = rel_max(gpu, ref)
    assert err < 1e-12, err


@pytest.mark.parametrize("shape", [(128, 64, 64), (256, 32, 32), (256, 64, 1)])
def test_variant1_plans_mixed(shape):
    """The mixed form (FP32 bulk, FP64 state) with the second plan variant of
    every kernel class, dipolar and soft-core terms on (the density potential
    that the FP32 x stage keeps in LDS): RK4 steps within 1e-6 of the FP64
    oracle."""
    from fgpe_amd import FGPE_PREC_MIXED, configs
    nx, ny, nz = shape
    c = Case(*shape, L=(10.0, 12.0, 20.0 if nz > 1 else 1.0))
    h = dict(a=(0.9,), a_dd=0.8, dip=(0.0, 0.6, 0.8) if nz > 1 else (0.6, 0.8, 0.0),
             sc_U=1.5, sc_R=1.1, sc_dim=3 if nz > 1 else 2)
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.25))
    psi = configs.normalise(np.exp(-0.5 * np.abs(V)) + 0.01 * c.rand_psi(57), c.dv)
    with c.ctx(FGPE_PREC_MIXED) as ctx:
        for k in range(5):
            for axis in range(3):
                ctx.set_variant(k, axis, 1)
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.rk4_steps(5e-4, 3)
        gpu = ctx.download_psi()
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 5e-4, 3)
    err = rel_max(gpu, ref)
    assert err < 1e-6, (shape, err)


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_state(precision):
    """The autotuner's timing launches (stage-2 x stages on whatever the stage
    buffers hold, in the mixed form rebuilding the state from them) leave the
    state alone, and the steps after it match the oracle."""
    from fgpe_amd import configs
    c = Case(64, 64, 128, L=(16.0, 16.0, 32.0))
    h = dict(a=(1.0,), a_dd=1.3, dip=(0.0, 0.0, 1.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.25))
    psi = configs.normalise(np.exp(-0.5 * np.abs(V)) + 0.01 * c.rand_psi(53), c.dv)
    with c.ctx(precision) as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ms = ctx.autotune(3)
        assert all(t > 0 for v in ms.values() for t in v if t != 0)
        assert np.array_equal(ctx.download_psi(), psi)
        ctx.rk4_steps(5e-4, 4)
        gpu = ctx.download_psi()
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 5e-4, 4)
    err = rel_max(gpu, ref)
    assert err < (1e-12 if precision == 0 else 1e-6), err


@pytest.mark.parametrize("precision", [0, 1])
def test_column_slabs_are_bitwise_neutral(precision):
    """Splitting the y / last / y passes into 2 or 4 column slabs changes only
    the launch partition: the states after dipolar + soft-core RK4 steps are
    bit-identical to the unslabbed run (and, in FP64, at oracle parity)."""
    from fgpe_amd import configs
    c = Case(128, 32, 64, L=(16.0, 8.0, 16.0))
    h = dict(a=(0.9,), a_dd=0.7, dip=(0.3, 0.0, 0.95), sc_U=1.5, sc_R=1.1, sc_dim=3)
    V = configs.harmonic_trap(c.spec, (1.0, 1.2, 0.8))
    psi = configs.normalise(np.exp(-0.5 * np.abs(V)) + 0.01 * c.rand_psi(54), c.dv)
    outs = {}
    for ns in (1, 2, 4):
        with c.ctx(precision) as ctx:
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.set_slabs(ns)
            assert ctx.slabs() == ns
            ctx.rk4_steps(5e-4, 5)
            outs[ns] = ctx.download_psi()
    for ns in (2, 4):
        err = float(np.abs(outs[ns] - outs[1]).max())
        print(f"nslab {ns}: max |slabbed - unslabbed| = {err:.3e}")
        assert err == 0.0, (ns, err)
    if precision == 0:
        ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 5e-4, 5)
        err = rel_max(outs[1], ref)
        assert err < 1e-12, err


@pytest.mark.parametrize("precision", [0, 1])
def test_block_order_is_bitwise_neutral(precision):
    """Reversing the block order of any of the four passes only changes which
    block takes which tile: states and observables are bit-identical."""
    from fgpe_amd import configs
    c = Case(64, 32, 64, L=(16.0, 8.0, 16.0))
    h = dict(a=(0.9,), a_dd=0.7, dip=(0.3, 0.0, 0.95))
    V = configs.harmonic_trap(c.spec, (1.0, 1.2, 0.8))
    psi = configs.normalise(np.exp(-0.5 * np.abs(V)) + 0.01 * c.rand_psi(55), c.dv)
    outs = {}
    for mask in (0, 5, 10, 15):
        with c.ctx(precision) as ctx:
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.set_order(mask)
            assert ctx.order() == mask
            ctx.rk4_steps(5e-4, 4)
            outs[mask] = (ctx.download_psi(), ctx.observables(2e-3))
    for mask in (5, 10, 15):
        err = float(np.abs(outs[mask][0] - outs[0][0]).max())
        print(f"order {mask}: max |psi - psi(order 0)| = {err:.3e}")
        assert err == 0.0, (mask, err)
        assert np.array_equal(outs[mask][1], outs[0][1]), (mask, outs[mask][1], outs[0][1])


@pytest.mark.parametrize("shape", [(96, 32, 64), (192, 64, 32), (512, 32, 32), (32, 96, 192),
                                   (64, 64, 1), (96, 192, 1)])
def test_rk4_dipolar_softcore_all_plans_f64(shape):
    """Full RK4 steps through every pencil length's plan (radix-3 lengths, the
    512 plan, 2D grids), with a tilted dipole (the half-spectrum density and
    its Nyquist-plane symmetrisation) plus a soft-core term: 1e-12 against the
    oracle."""
    from fgpe_amd import configs
    nx, ny, nz = shape
    L = (12.0, 10.0, 14.0 if nz > 1 else 1.0)
    c = Case(nx, ny, nz, L=L)
    h = dict(a=(0.6,), a_dd=0.5, dip=(0.48, 0.6, 0.64) if nz > 1 else (0.6, 0.8, 0.0),
             sc_U=1.5, sc_R=1.1, sc_dim=3 if nz > 1 else 2)
    V = configs.harmonic_trap(c.spec, (1.0, 1.1, 0.9))
    psi = configs.normalise(np.exp(-0.5 * np.abs(V)) + 0.01 * c.rand_psi(56), c.dv)
    dt, n = 5e-4, 4
    gpu, obs = run_both(c, h, psi, V, dt, n)
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, dt, n)
    err = rel_max(gpu, ref)
    assert err < 1e-12, (shape, err)
    hmu, hen = orc.hpsi(c.grid, oracle_ham(c, h), n * dt, ref, V)
    nrm = orc.norm(c.grid, ref)
    mu, en = orc.obs2(c.grid, nrm, ref, hmu, hen)
    assert abs(obs[2] - en) < 1e-12 * abs(en), (shape, obs[2], en)


def test_rk4_nontemporal_x_stage_f64():
    """256x256x64 FP64 dipolar: the per-step working set (about 370 MB) is
    larger than the memory-side cache, so the x stage runs its non-temporal
    instantiation; two RK4 steps against the oracle at the drop-in bar."""
    from fgpe_amd import configs
    h = dict(HAMS["dipolar"])
    c = Case(256, 256, 64, L=(16.0, 16.0, 8.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 2.0))
    psi = configs.normalise(np.exp(-0.5 * np.abs(configs.harmonic_trap(c.spec, (1, 1, 2)))) +
                            0.01 * c.rand_psi(9), c.dv)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        print("working set MB", 256 * 256 * 64 * (56 + 2 * 16) / 2 ** 20, flush=True)
        ctx.rk4_steps(5e-4, 2)
        gpu = ctx.download_psi()
    orc.set_threads(16)
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 5e-4, 2)
    err = rel_max(gpu, ref)
    assert err < 1e-12, err


@pytest.mark.parametrize("prec", [0, 1])
def test_two_component_imaginary_time_with_softcore(prec):
    """Two components with the soft-core kernel on n1 + n2 in imaginary time
    (the renormalisation applied at once, since the two components' scales
    differ): heavy ball with the same iterations, state and norm / mu / E
    history as the oracle (1e-11), and RK4 with renormalisation (1e-12); the
    mixed form within 1e-6."""
    from fgpe_amd import FGPE_IMAG_HEAVY_BALL, FGPE_IMAG_RK4_RENORM
    c, V, psi = two_comp_case()
    h = dict(ncomp=2, a=(0.02, 0.03, 0.01), sc_U=2.0, sc_R=1.2, sc_dim=3)
    tol = 1e-11 if prec == 0 else 1e-6
    dt, beta, npart, eps, n = 2e-3, 0.9, 1.0, 1e-9, 50
    ref, nref, href = orc.imag_heavy_ball(c.grid, two_comp_ham(c, h), psi, V, dt, beta, npart, eps, n)
    with c.ctx(prec) as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(dt, n, scheme=FGPE_IMAG_HEAVY_BALL, beta=beta, npart=npart, eps=eps,
                                    history=True)
        gpu = ctx.download_psi()
    assert done == nref
    assert rel_max(gpu, ref) < tol
    assert np.abs(hist - href).max() <= tol * np.abs(href).max()
    ref = orc.rk4(c.grid, two_comp_ham(c, h), psi, V, 1e-3, 30, imaginary=True, npart=npart)
    with c.ctx(prec) as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.imag_steps(1e-3, 30, scheme=FGPE_IMAG_RK4_RENORM, npart=npart)
        err = rel_max(ctx.download_psi(), ref)
    assert err < (1e-12 if prec == 0 else 1e-6), err


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["dip+sc", "moving"])
def test_kmul_on_y_pass(precision, name):
    """3D grids with the k-space multiply on the y pass (z pencils run the
    strided forward and inverse passes): RK4 real time and renormalised
    imaginary time against the oracle (FP64 1e-12, mixed 1e-6), for both plan
    variants of the passes."""
    from fgpe_amd import configs
    c = Case(64, 32, 128, L=(10.0, 8.0, 20.0))
    h = HAMS[name]
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.25))
    psi = configs.normalise(np.exp(-0.5 * np.abs(configs.harmonic_trap(c.spec, (1, 1, 0.5)))) +
                            0.01 * c.rand_psi(71), c.dv)
    tol = 1e-12 if precision == 0 else 1e-6
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 5e-4, 3)
    for variant in (0, 1):
        with c.ctx(precision) as ctx:
            for k in range(5):
                for axis in range(3):
                    ctx.set_variant(k, axis, variant)
            ctx.set_kmul_axis(1)
            assert ctx.kmul_axis() == 1
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.rk4_steps(5e-4, 3)
            err = rel_max(ctx.download_psi(), ref)
            assert err < tol, (variant, err)
    if name == "dip+sc":
        with c.ctx(precision) as ctx:
            ctx.set_kmul_axis(1)
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.imag_steps(1e-3, 4)
            gpu = ctx.download_psi()
        refi = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 1e-3, 4, imaginary=True, npart=1.0)
        err = rel_max(gpu, refi)
        assert err < tol, err


def test_kmul_axis_choices():
    """A 3D grid takes the multiply on axis 1 or 2; 2D and 1D grids have one choice."""
    from fgpe_amd import FgpeError
    c3, c2, c1 = Case(32, 32, 32), Case(64, 32, 1), Case(64, 1, 1)
    with c3.ctx() as ctx:
        assert ctx.kmul_axis() == 2
        ctx.set_kmul_axis(1)
        assert ctx.kmul_axis() == 1
        ctx.set_kmul_axis(2)
        assert ctx.kmul_axis() == 2
        with pytest.raises(FgpeError):
            ctx.set_kmul_axis(0)
    with c2.ctx() as ctx:
        assert ctx.kmul_axis() == 1
        ctx.set_kmul_axis(1)
        with pytest.raises(FgpeError):
            ctx.set_kmul_axis(2)
    with c1.ctx() as ctx:
        assert ctx.kmul_axis() == 0
        with pytest.raises(FgpeError):
            ctx.set_kmul_axis(1)


XSTAGE = 0   # FGPE_KCLASS_XSTAGE


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("shape", [(256, 32, 32), (128, 32, 64), (256, 64, 1), (128, 128, 1)])
def test_x_stage_third_plan(precision, shape):
    """The x stage's third plan (4 elements per thread, radix-4 stages; at
    nx = 256 one wave per pencil): dipolar + soft-core RK4 against the oracle
    (FP64 1e-12, mixed 1e-6), and in FP64 the heavy ball with its on-device
    history (1e-11)."""
    from fgpe_amd import FGPE_IMAG_HEAVY_BALL, configs
    c = Case(*shape, L=(16.0, 8.0, 8.0) if shape[2] > 1 else (16.0, 12.0, 1.0))
    h = dict(HAMS["dip+sc"])
    if shape[2] == 1:
        h.update(sc_dim=2, dip=(1.0, 0.0, 0.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.5))
    psi = configs.normalise(np.exp(-0.5 * np.abs(configs.harmonic_trap(c.spec, (1, 1, 0.5)))) +
                            0.01 * c.rand_psi(81), c.dv)
    with c.ctx(precision) as ctx:
        ctx.set_variant(XSTAGE, 0, 2)
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.rk4_steps(5e-4, 3)
        gpu = ctx.download_psi()
    ref = orc.rk4(c.grid, oracle_ham(c, h), psi, V, 5e-4, 3)
    err = rel_max(gpu, ref)
    assert err < (1e-12 if precision == 0 else 1e-6), err
    if precision == 1:
        return
    dt, beta, npart, eps, n = 2e-3, 0.9, 1.0, 1e-7, 40
    ref, nref, href = orc.imag_heavy_ball(c.grid, oracle_ham(c, h), psi, V, dt, beta, npart, eps, n)
    with c.ctx() as ctx:
        ctx.set_variant(XSTAGE, 0, 2)
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(dt, n, scheme=FGPE_IMAG_HEAVY_BALL, beta=beta, npart=npart,
                                    eps=eps, history=True)
        gpu = ctx.download_psi()
    assert done == nref, (done, nref)
    assert rel_max(gpu, ref) < 1e-11
    assert np.abs(hist - href).max() <= 1e-11 * np.abs(href).max()


def test_x_stage_third_plan_two_components():
    from fgpe_amd import configs
    c = Case(128, 16, 32, L=(16.0, 8.0, 8.0))
    V1 = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.5))
    V = np.stack([V1, V1 + 0.1])
    p1 = configs.gaussian_noisy(c.spec, 7, offset=(1.0, 0, 0), noise=0.02)
    p2 = configs.gaussian_noisy(c.spec, 8, offset=(-1.0, 0, 0), noise=0.02)
    psi = np.stack([p1, p2 * np.exp(0.3j)])
    f = 1.0 / (4 * np.pi)
    h = dict(ncomp=2, a=(100 * f, 100 * f, 80 * f), sc_U=5.0, sc_R=1.5, sc_dim=3)
    with c.ctx() as ctx:
        ctx.set_variant(XSTAGE, 0, 2)
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        ctx.rk4_steps(5e-4, 4)
        gpu = ctx.download_psi()
    ref = orc.rk4(c.grid, two_comp_ham(c, h), psi, V, 5e-4, 4)
    err = rel_max(gpu, ref)
    assert err < 1e-12, err


def test_x_stage_variant_range():
    """Variant 2 exists for the x-stage class only; on a length without a
    third plan it runs variant 1."""
    from fgpe_amd import FgpeError
    c = Case(64, 32, 32)
    with c.ctx() as ctx:
        ctx.set_variant(XSTAGE, 0, 2)
        with pytest.raises(FgpeError):
            ctx.set_variant(1, 1, 2)
        with pytest.raises(FgpeError):
            ctx.set_variant(XSTAGE, 0, 3)


@pytest.mark.parametrize("nx", [128, 256])
def test_x_stage_third_plan_1d(nx):
    """1D grids (one x row, fewer rows than a block's pencils) with the x
    stage's third plan and the 1D soft-core kernel: RK4 against the oracle
    (FP64 1e-12, mixed 1e-6)."""
    from fgpe_amd import configs
    c = Case(nx, 1, 1, L=(12.0, 1.0, 1.0))
    h = dict(a=(0.3,), a_dd=0.15, dip=(1.0, 0.0, 0.0), sc_U=2.0, sc_R=1.1, sc_dim=1)
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 1.0))
    psi = configs.gaussian_noisy(c.spec, 63, width=1.5, noise=0.02)
    vt = c.vtilde(h["a_dd"], h["dip"], sc=(1, h["sc_R"], h["sc_U"]))
    kmax = float(np.max(np.abs(c.kx)))
    dt, n = min(1e-3, 0.5 / (19.739208802 * kmax ** 2)), 12
    ref = orc.rk4(c.grid, orc.Ham(1, h["a"], vtilde=vt), psi, V, dt, n)
    for prec, tol in ((0, 1e-12), (1, 1e-6)):
        with c.ctx(prec) as ctx:
            ctx.set_variant(XSTAGE, 0, 2)
            ctx.set_hamiltonian(**h)
            ctx.upload_psi(psi)
            ctx.upload_vext(V)
            ctx.rk4_steps(dt, n)
            err = rel_max(ctx.download_psi(), ref)
        assert err < tol, (prec, err)
