"""Adjoint gradients of real-time RK4 (fgpe_adjoint_*): against central
finite differences of the HIP forward, against finite differences of the CPU
FP64 oracle, through torch.autograd, and mixed against FP64."""
import numpy as np
import pytest

from fgpe_amd import configs
from oracle import oracle as orc
from tests.cases import Case

pytestmark = pytest.mark.gpu


def setup_case():
    c = Case(32, 32, 64, L=(10.0, 10.0, 20.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.25)).real
    psi0 = configs.normalise(np.exp(-0.5 * configs.harmonic_trap(c.spec, (1, 1, 0.5)).real) +
                             0.01 * c.rand_psi(41), c.dv)
    target = configs.normalise(np.exp(-0.6 * configs.harmonic_trap(c.spec, (1, 1, 0.4)).real) *
                               np.exp(0.2j * c.rand_psi(42).real), c.dv)
    a, a_dd = 0.6, 0.5
    return c, V, psi0, target, a, a_dd


def smooth_field(c, seed):
    u = c.rand_psi(seed).real
    F = np.fft.fftn(u)
    KZ, KY, KX = np.meshgrid(c.kz, c.ky, c.kx, indexing="ij")
    F *= np.exp(-4.0 * (KX ** 2 + KY ** 2 + KZ ** 2))
    f = np.fft.ifftn(F).real
    return f / np.abs(f).max()


def J_of(psiT, target, dv):
    ov = np.sum(np.conj(target) * psiT) * dv
    return float(abs(ov) ** 2), ov


SC = (3, 1.2)   # soft-core dimension and radius of the soft-core cases


def hamiltonian(ctx, a, s, softcore):
    """a_dd = s (dipolar along z) or, with softcore, the soft-core U = s."""
    if softcore:
        ctx.set_hamiltonian(a=(a,), sc_U=s, sc_R=SC[1], sc_dim=SC[0])
    else:
        ctx.set_hamiltonian(a=(a,), a_dd=s, dip=(0.0, 0.0, 1.0))


def gpu_forward(c, V, a, a_dd, psi0, dt, n, prec=0, softcore=False):
    with c.ctx(prec) as ctx:
        hamiltonian(ctx, a, a_dd, softcore)
        ctx.upload_vext(V.astype(np.complex128))
        ctx.upload_psi(psi0)
        ctx.rk4_steps(dt, n)
        return ctx.download_psi()


def gpu_gradient(torch, c, V, a, a_dd, psi0, target, dt, n, every, prec=0, softcore=False):
    with c.ctx(prec) as ctx:
        hamiltonian(ctx, a, a_dd, softcore)
        ctx.upload_vext(V.astype(np.complex128))
        ctx.upload_psi(psi0)
        ctx.adjoint_forward(dt, n, every)
        psiT = ctx.download_psi()
        J, ov = J_of(psiT, target, c.dv)
        lamT = torch.from_numpy(2 * ov * c.dv * target).cuda()
        gV = torch.empty(c.shape, dtype=torch.float64, device="cuda")
        lam0 = torch.empty_like(lamT)
        g = ctx.adjoint_backward(lamT.data_ptr(), lam0.data_ptr(), gV.data_ptr())
        ctx.sync()
        return J, gV.cpu().numpy(), g, lam0.cpu().numpy()


def test_adjoint_against_finite_differences(torch_cuda):
    c, V, psi0, target, a, a_dd = setup_case()
    dt, n, every = 5e-4, 40, 7       # uneven last segment on purpose
    J, gV, g, lam0 = gpu_gradient(torch_cuda, c, V, a, a_dd, psi0, target, dt, n, every)
    h = 1e-4
    for j in range(4):
        dV = smooth_field(c, 100 + j)
        Jp = J_of(gpu_forward(c, V + h * dV, a, a_dd, psi0, dt, n), target, c.dv)[0]
        Jm = J_of(gpu_forward(c, V - h * dV, a, a_dd, psi0, dt, n), target, c.dv)[0]
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.sum(gV * dV))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (j, ad, fd)
    for k, (pa, pd) in enumerate(((h, 0.0), (0.0, h))):
        Jp = J_of(gpu_forward(c, V, a + pa, a_dd + pd, psi0, dt, n), target, c.dv)[0]
        Jm = J_of(gpu_forward(c, V, a - pa, a_dd - pd, psi0, dt, n), target, c.dv)[0]
        fd = (Jp - Jm) / (2 * h)
        assert abs(g[k] - fd) <= 1e-6 * abs(fd), (k, g[k], fd)
    # initial-state gradient: directional derivative along a complex perturbation
    dpsi = c.rand_psi(77, 1e-2)
    Jp = J_of(gpu_forward(c, V, a, a_dd, psi0 + h * dpsi, dt, n), target, c.dv)[0]
    Jm = J_of(gpu_forward(c, V, a, a_dd, psi0 - h * dpsi, dt, n), target, c.dv)[0]
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(np.real(np.conj(lam0) * dpsi)))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_adjoint_against_oracle_finite_differences(torch_cuda):
    """The FP64 reference gradient: central differences of the CPU oracle."""
    c, V, psi0, target, a, a_dd = setup_case()
    dt, n, every = 5e-4, 20, 5
    J, gV, g, _ = gpu_gradient(torch_cuda, c, V, a, a_dd, psi0, target, dt, n, every)
    h = 1e-4
    for j in range(2):
        dV = smooth_field(c, 200 + j)
        Js = []
        for sgn in (1, -1):
            hm = orc.Ham(1, (a,), vtilde=orc.vdd(c.grid, a_dd, (0, 0, 1)))
            out = orc.rk4(c.grid, hm, psi0, (V + sgn * h * dV).astype(np.complex128), dt, n)
            Js.append(J_of(out, target, c.dv)[0])
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float(np.sum(gV * dV))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (j, ad, fd)


def test_adjoint_through_autograd(torch_cuda):
    torch = torch_cuda
    from fgpe_amd.adjoint import AdjointSolver, GPEEvolve, fidelity
    c, V, psi0, target, a, a_dd = setup_case()
    dt, n = 5e-4, 30
    # on torch's stream: device tensors pass straight in; no autotuning, so that the
    # plans match the plain context's below bit for bit
    solver = AdjointSolver(c.spec, dt, n, torch=torch, autotune=False)
    Vt = torch.tensor(V, dtype=torch.float64, device="cuda", requires_grad=True)
    at = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    addt = torch.tensor(a_dd, dtype=torch.float64, requires_grad=True)
    p0 = torch.tensor(psi0, dtype=torch.complex128, device="cuda", requires_grad=True)
    tg = torch.tensor(target, dtype=torch.complex128, device="cuda")
    psiT = GPEEvolve().apply(Vt, at, addt, p0, solver)
    loss = fidelity(psiT, tg, c.dv)
    loss.backward()
    J, gV, g, lam0 = gpu_gradient(torch, c, V, a, a_dd, psi0, target, dt, n, solver.every)
    assert abs(loss.item() - J) < 1e-14
    assert np.abs(Vt.grad.cpu().numpy() - gV).max() <= 1e-13 * np.abs(gV).max()
    assert abs(at.grad.item() - g[0]) <= 1e-13 * abs(g[0])
    assert abs(addt.grad.item() - g[1]) <= 1e-13 * abs(g[1])
    assert np.abs(p0.grad.cpu().numpy() - lam0).max() <= 1e-13 * np.abs(lam0).max()
    solver.ctx.close()


def test_adjoint_mixed_precision(torch_cuda):
    c, V, psi0, target, a, a_dd = setup_case()
    dt, n, every = 5e-4, 40, 7
    _, g64, s64, _ = gpu_gradient(torch_cuda, c, V, a, a_dd, psi0, target, dt, n, every, prec=0)
    _, g32, s32, _ = gpu_gradient(torch_cuda, c, V, a, a_dd, psi0, target, dt, n, every, prec=1)
    rel = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    print(f"mixed vs f64 adjoint: rel L2 dJ/dV {rel:.3e}, da {abs(s32[0]-s64[0])/abs(s64[0]):.3e}, "
          f"da_dd {abs(s32[1]-s64[1])/abs(s64[1]):.3e}")
    assert rel < 1e-4
    assert abs(s32[0] - s64[0]) <= 1e-4 * abs(s64[0])
    assert abs(s32[1] - s64[1]) <= 1e-4 * abs(s64[1])


def test_softcore_adjoint_against_finite_differences(torch_cuda):
    """The soft-core kernel (a_dd = 0): dJ/dV, dJ/da and dJ/dU against central
    differences of the HIP forward, dJ/dV also against the CPU oracle's."""
    c, V, psi0, target, a, _ = setup_case()
    U = 1.5
    dt, n, every = 5e-4, 30, 6
    J, gV, g, _ = gpu_gradient(torch_cuda, c, V, a, U, psi0, target, dt, n, every, softcore=True)
    assert g[1] == 0.0
    h = 1e-4
    for j in range(2):
        dV = smooth_field(c, 300 + j)
        Jp = J_of(gpu_forward(c, V + h * dV, a, U, psi0, dt, n, softcore=True), target, c.dv)[0]
        Jm = J_of(gpu_forward(c, V - h * dV, a, U, psi0, dt, n, softcore=True), target, c.dv)[0]
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.sum(gV * dV))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (j, ad, fd)
    for k, (pa, pu) in ((0, (h, 0.0)), (2, (0.0, h))):
        Jp = J_of(gpu_forward(c, V, a + pa, U + pu, psi0, dt, n, softcore=True), target, c.dv)[0]
        Jm = J_of(gpu_forward(c, V, a - pa, U - pu, psi0, dt, n, softcore=True), target, c.dv)[0]
        fd = (Jp - Jm) / (2 * h)
        assert abs(g[k] - fd) <= 1e-6 * abs(fd), (k, g[k], fd)
    dV = smooth_field(c, 310)
    Js = []
    for sgn in (1, -1):
        hm = orc.Ham(1, (a,), vtilde=orc.vsoftcore(c.grid, SC[0], SC[1], U))
        out = orc.rk4(c.grid, hm, psi0, (V + sgn * h * dV).astype(np.complex128), dt, n)
        Js.append(J_of(out, target, c.dv)[0])
    fd = (Js[0] - Js[1]) / (2 * h)
    ad = float(np.sum(gV * dV))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_softcore_adjoint_through_autograd(torch_cuda):
    torch = torch_cuda
    from fgpe_amd.adjoint import AdjointSolver, GPEEvolve, fidelity
    c, V, psi0, target, a, _ = setup_case()
    U, dt, n = 1.5, 5e-4, 20
    solver = AdjointSolver(c.spec, dt, n, torch=torch, softcore=SC, autotune=False)
    Vt = torch.tensor(V, dtype=torch.float64, device="cuda", requires_grad=True)
    at = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    ut = torch.tensor(U, dtype=torch.float64, requires_grad=True)
    p0 = torch.tensor(psi0, dtype=torch.complex128, device="cuda", requires_grad=True)
    tg = torch.tensor(target, dtype=torch.complex128, device="cuda")
    loss = fidelity(GPEEvolve().apply(Vt, at, ut, p0, solver), tg, c.dv)
    loss.backward()
    J, gV, g, lam0 = gpu_gradient(torch, c, V, a, U, psi0, target, dt, n, solver.every, softcore=True)
    assert abs(loss.item() - J) < 1e-14
    assert np.abs(Vt.grad.cpu().numpy() - gV).max() <= 1e-13 * np.abs(gV).max()
    assert abs(at.grad.item() - g[0]) <= 1e-13 * abs(g[0])
    assert abs(ut.grad.item() - g[2]) <= 1e-13 * abs(g[2])
    solver.ctx.close()


@pytest.mark.parametrize("dims", ["3d", "1d"])
def test_adjoint_with_both_kernels(torch_cuda, dims):
    """Dipolar and soft-core kernels on together: dJ/dV, dJ/da, dJ/da_dd and
    dJ/dU against central differences of the HIP forward (the two strengths'
    sums separated by the dipolar kernel alone on a third channel), one V
    direction against the oracle's; on a 3D grid and on a 1D one."""
    if dims == "3d":
        c, V, psi0, target, a, a_dd = setup_case()
        dip, sc = (0.0, 0.0, 1.0), (3, 1.2, 1.5)
    else:
        c = Case(128, 1, 1, L=(12.0, 1.0, 1.0))
        V = configs.harmonic_trap(c.spec, (1.0, 1.0, 1.0)).real
        psi0 = configs.gaussian_noisy(c.spec, 95, width=1.5, noise=0.02)
        target = configs.normalise(np.exp(-0.4 * V) * np.exp(0.5j * c.rand_psi(96).real), c.dv)
        a, a_dd, dip, sc = 0.3, 0.15, (1.0, 0.0, 0.0), (1, 1.1, 2.0)
    ham = dict(a=(a,), a_dd=a_dd, dip=dip, sc_U=sc[2], sc_R=sc[1], sc_dim=sc[0])
    dt, n, every = 5e-4, 30, 7

    def forward(Vx, hx):
        with c.ctx() as ctx:
            ctx.set_hamiltonian(**hx)
            ctx.upload_vext(Vx.astype(np.complex128))
            ctx.upload_psi(psi0)
            ctx.rk4_steps(dt, n)
            return J_of(ctx.download_psi(), target, c.dv)[0]

    with c.ctx() as ctx:
        ctx.set_hamiltonian(**ham)
        ctx.upload_vext(V.astype(np.complex128))
        ctx.upload_psi(psi0)
        ctx.adjoint_forward(dt, n, every)
        J, ov = J_of(ctx.download_psi(), target, c.dv)
        lamT = torch_cuda.from_numpy(2 * ov * c.dv * target).cuda()
        gV = torch_cuda.empty(c.shape, dtype=torch_cuda.float64, device="cuda")
        g = ctx.adjoint_backward(lamT.data_ptr(), 0, gV.data_ptr())
        ctx.sync()
        gV = gV.cpu().numpy()
    assert abs(J - forward(V, ham)) < 1e-14
    h = 1e-4
    for j in range(2):
        dV = smooth_field(c, 320 + j)
        f = (forward(V + h * dV, ham) - forward(V - h * dV, ham)) / (2 * h)
        ad = float(np.sum(gV * dV))
        assert abs(ad - f) <= 1e-6 * abs(f), (j, ad, f)
    for k, key in ((0, "a"), (1, "a_dd"), (2, "sc_U")):
        hp, hm = dict(ham), dict(ham)
        if key == "a":
            hp["a"], hm["a"] = (a + h,), (a - h,)
        else:
            hp[key], hm[key] = ham[key] + h, ham[key] - h
        f = (forward(V, hp) - forward(V, hm)) / (2 * h)
        assert abs(g[k] - f) <= 1e-6 * abs(f), (key, g[k], f)
    dV = smooth_field(c, 330)
    vt = c.vtilde(a_dd, dip, sc=sc)
    Js = []
    for sgn in (1, -1):
        out = orc.rk4(c.grid, orc.Ham(1, (a,), vtilde=vt), psi0, (V + sgn * h * dV).astype(np.complex128), dt, n)
        Js.append(J_of(out, target, c.dv)[0])
    f = (Js[0] - Js[1]) / (2 * h)
    ad = float(np.sum(gV * dV))
    assert abs(ad - f) <= 1e-6 * abs(f), (ad, f)


@pytest.mark.parametrize("nx", [128, 256])
def test_adjoint_x_stage_third_plan(torch_cuda, nx):
    """The adjoint's x kernels with the x class's third plan (four pencils
    per block: half-wave pencils at nx = 128, one wave each at nx = 256): the
    same gradient as with variant 1 (to rounding) and central differences."""
    c = Case(nx, 16, 32, L=(16.0, 8.0, 10.0))
    V = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.25)).real
    psi0 = configs.normalise(np.exp(-0.5 * configs.harmonic_trap(c.spec, (1, 1, 0.5)).real) +
                             0.01 * c.rand_psi(43), c.dv)
    target = configs.normalise(np.exp(-0.6 * configs.harmonic_trap(c.spec, (1, 1, 0.4)).real) *
                               np.exp(0.2j * c.rand_psi(44).real), c.dv)
    a, a_dd, dt, n, every = 0.6, 0.5, 5e-4, 40, 7
    grads = []
    for variant in (1, 2):
        with c.ctx() as ctx:
            ctx.set_variant(0, 0, variant)
            ctx.set_hamiltonian(a=(a,), a_dd=a_dd, dip=(0.0, 0.0, 1.0))
            ctx.upload_vext(V.astype(np.complex128))
            ctx.upload_psi(psi0)
            ctx.adjoint_forward(dt, n, every)
            J, ov = J_of(ctx.download_psi(), target, c.dv)
            lamT = torch_cuda.from_numpy(2 * ov * c.dv * target).cuda()
            gV = torch_cuda.empty(c.shape, dtype=torch_cuda.float64, device="cuda")
            g = ctx.adjoint_backward(lamT.data_ptr(), 0, gV.data_ptr())
            ctx.sync()
            grads.append((gV.cpu().numpy(), g))
    (g1, s1), (g2, s2) = grads
    assert np.abs(g2 - g1).max() <= 1e-10 * np.abs(g1).max()
    assert abs(s2[0] - s1[0]) <= 1e-10 * abs(s1[0]) and abs(s2[1] - s1[1]) <= 1e-10 * abs(s1[1])
    h = 1e-4
    dV = smooth_field(c, 400)
    Jp = J_of(gpu_forward(c, V + h * dV, a, a_dd, psi0, dt, n), target, c.dv)[0]
    Jm = J_of(gpu_forward(c, V - h * dV, a, a_dd, psi0, dt, n), target, c.dv)[0]
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(g2 * dV))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_adjoint_split_pass_plans(torch_cuda):
    """The adjoint's y and last-axis passes on 256-long axes, where plan 0 of
    the FP64 passes exchanges real and imaginary parts in turn: the same
    gradient as with plan 1 (to rounding) and central differences."""
    c = Case(16, 256, 256, L=(8.0, 16.0, 24.0))
    V = configs.harmonic_trap(c.spec, (1.0, 0.8, 0.5)).real
    psi0 = configs.normalise(np.exp(-0.5 * configs.harmonic_trap(c.spec, (1, 1, 0.5)).real) +
                             0.01 * c.rand_psi(45), c.dv)
    target = configs.normalise(np.exp(-0.6 * configs.harmonic_trap(c.spec, (1, 1, 0.4)).real) *
                               np.exp(0.2j * c.rand_psi(46).real), c.dv)
    a, a_dd, dt, n, every = 0.6, 0.5, 5e-4, 24, 5
    grads = []
    for variant in (0, 1):
        with c.ctx() as ctx:
            for k in (1, 2):   # strided and last-axis classes, every axis
                for axis in range(3):
                    ctx.set_variant(k, axis, variant)
            ctx.set_hamiltonian(a=(a,), a_dd=a_dd, dip=(0.0, 0.0, 1.0))
            ctx.upload_vext(V.astype(np.complex128))
            ctx.upload_psi(psi0)
            ctx.adjoint_forward(dt, n, every)
            J, ov = J_of(ctx.download_psi(), target, c.dv)
            lamT = torch_cuda.from_numpy(2 * ov * c.dv * target).cuda()
            gV = torch_cuda.empty(c.shape, dtype=torch_cuda.float64, device="cuda")
            g = ctx.adjoint_backward(lamT.data_ptr(), 0, gV.data_ptr())
            ctx.sync()
            grads.append((gV.cpu().numpy(), g))
    (g0, s0), (g1, s1) = grads
    assert np.abs(g0 - g1).max() <= 1e-10 * np.abs(g1).max()
    assert abs(s0[0] - s1[0]) <= 1e-10 * abs(s1[0]) and abs(s0[1] - s1[1]) <= 1e-10 * abs(s1[1])
    # central differences at h and h/2, Richardson-extrapolated, along the
    # gradient itself: on this grid a smooth random direction's derivative is
    # about 1e-6 of J's scale, where the forward's rounding decides the
    # differences' last digits
    dV = g0 / np.abs(g0).max()

    def fd(h):
        Jp = J_of(gpu_forward(c, V + h * dV, a, a_dd, psi0, dt, n), target, c.dv)[0]
        Jm = J_of(gpu_forward(c, V - h * dV, a, a_dd, psi0, dt, n), target, c.dv)[0]
        return (Jp - Jm) / (2 * h)

    f1, f2 = fd(2e-4), fd(1e-4)
    fr = (4 * f2 - f1) / 3
    ad = float(np.sum(g0 * dV))
    assert abs(ad - fr) <= 1e-6 * abs(fr), (ad, fr, f1, f2)


def test_adjoint_solver_autotuned(torch_cuda):
    """AdjointSolver's default: the plan autotuner runs at the first forward;
    the gradient then agrees with the untuned one to rounding."""
    torch = torch_cuda
    from fgpe_amd.adjoint import AdjointSolver, GPEEvolve, fidelity
    c, V, psi0, target, a, a_dd = setup_case()
    dt, n = 5e-4, 20
    grads = []
    for tune in (False, True):
        solver = AdjointSolver(c.spec, dt, n, torch=torch, autotune=tune)
        Vt = torch.tensor(V, dtype=torch.float64, device="cuda", requires_grad=True)
        p0 = torch.tensor(psi0, dtype=torch.complex128, device="cuda")
        tg = torch.tensor(target, dtype=torch.complex128, device="cuda")
        at = torch.tensor(a, dtype=torch.float64)
        addt = torch.tensor(a_dd, dtype=torch.float64)
        loss = fidelity(GPEEvolve().apply(Vt, at, addt, p0, solver), tg, c.dv)
        loss.backward()
        grads.append(Vt.grad.cpu().numpy())
        solver.ctx.close()
    assert np.abs(grads[1] - grads[0]).max() <= 1e-10 * np.abs(grads[0]).max()


def test_adjoint_with_a_ramp(torch_cuda):
    """The a(t) quench ramp in the differentiated dynamics: dJ/dV and the
    derivative along a uniform shift of a(t) (a0, a1, a2 moved together)
    against central differences of the HIP forward."""
    c, V, psi0, target, a, a_dd = setup_case()
    dt, n, every = 5e-4, 30, 6
    # a(t): 0.6 -> 0.9 over (0, 0.004], held to 0.008, then to 0.45 at 0.012
    q = (0.6, 0.9, 0.45, 0.004, 0.008, 0.012)

    def ham(ctx, qq):
        ctx.set_hamiltonian(a=(qq[0],), a_dd=a_dd, dip=(0.0, 0.0, 1.0), quench=qq)

    def forward(Vx, qq):
        with c.ctx() as ctx:
            ham(ctx, qq)
            ctx.upload_vext(Vx.astype(np.complex128))
            ctx.upload_psi(psi0)
            ctx.rk4_steps(dt, n)
            return ctx.download_psi()

    with c.ctx() as ctx:
        ham(ctx, q)
        ctx.upload_vext(V.astype(np.complex128))
        ctx.upload_psi(psi0)
        ctx.adjoint_forward(dt, n, every)
        J, ov = J_of(ctx.download_psi(), target, c.dv)
        lamT = torch_cuda.from_numpy(2 * ov * c.dv * target).cuda()
        gV = torch_cuda.empty(c.shape, dtype=torch_cuda.float64, device="cuda")
        g = ctx.adjoint_backward(lamT.data_ptr(), 0, gV.data_ptr())
        ctx.sync()
        gV = gV.cpu().numpy()
    assert abs(J - J_of(forward(V, q), target, c.dv)[0]) < 1e-14
    h = 1e-4
    dV = smooth_field(c, 500)
    fd = (J_of(forward(V + h * dV, q), target, c.dv)[0] - J_of(forward(V - h * dV, q), target, c.dv)[0]) / (2 * h)
    ad = float(np.sum(gV * dV))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    qp = tuple(v + h if i < 3 else v for i, v in enumerate(q))
    qm = tuple(v - h if i < 3 else v for i, v in enumerate(q))
    fd = (J_of(forward(V, qp), target, c.dv)[0] - J_of(forward(V, qm), target, c.dv)[0]) / (2 * h)
    assert abs(g[0] - fd) <= 1e-6 * abs(fd), (g[0], fd)


def two_comp_setup():
    c = Case(32, 32, 64, L=(12.0, 12.0, 16.0))
    V1 = configs.harmonic_trap(c.spec, (1.0, 1.0, 0.5)).real
    V = np.stack([V1, V1 + 0.1])
    p1 = configs.gaussian_noisy(c.spec, 7, offset=(1.0, 0, 0), noise=0.02)
    p2 = configs.gaussian_noisy(c.spec, 8, offset=(-1.0, 0, 0), noise=0.02)
    psi0 = np.stack([p1, p2 * np.exp(0.3j)])
    t1 = configs.normalise(np.exp(-0.6 * configs.harmonic_trap(c.spec, (1, 1, 0.4)).real) *
                           np.exp(0.2j * c.rand_psi(45).real), c.dv)
    t2 = configs.normalise(np.exp(-0.5 * configs.harmonic_trap(c.spec, (1, 1, 0.6)).real) *
                           np.exp(-0.3j * c.rand_psi(46).real), c.dv)
    f = 1.0 / (4 * np.pi)
    a = (60 * f, 50 * f, 40 * f)   # a_11, a_22, a_12
    return c, V, psi0, np.stack([t1, t2]), a


def J2(psiT, target, dv):
    ov = np.sum(np.conj(target) * psiT, axis=(1, 2, 3)) * dv
    return float(np.sum(np.abs(ov) ** 2)), ov


def ham2(ctx, a, U):
    if U:
        ctx.set_hamiltonian(ncomp=2, a=a, sc_U=U, sc_R=1.5, sc_dim=3)
    else:
        ctx.set_hamiltonian(ncomp=2, a=a)


def fwd2(c, V, a, U, psi0, dt, n, prec=0):
    with c.ctx(prec) as ctx:
        ham2(ctx, a, U)
        ctx.upload_vext(V.astype(np.complex128))
        ctx.upload_psi(psi0)
        ctx.rk4_steps(dt, n)
        return ctx.download_psi()


def grad2(torch, c, V, a, U, psi0, target, dt, n, every, prec=0):
    with c.ctx(prec) as ctx:
        ham2(ctx, a, U)
        ctx.upload_vext(V.astype(np.complex128))
        ctx.upload_psi(psi0)
        ctx.adjoint_forward(dt, n, every)
        J, ov = J2(ctx.download_psi(), target, c.dv)
        lamT = torch.from_numpy(np.ascontiguousarray(2 * ov[:, None, None, None] * c.dv * target)).cuda()
        gV = torch.empty((2,) + c.shape, dtype=torch.float64, device="cuda")
        lam0 = torch.empty_like(lamT)
        g = ctx.adjoint_backward(lamT.data_ptr(), lam0.data_ptr(), gV.data_ptr())
        ctx.sync()
        return J, gV.cpu().numpy(), g, lam0.cpu().numpy()


@pytest.mark.parametrize("U", [0.0, 5.0])
def test_two_component_adjoint_against_finite_differences(torch_cuda, U):
    """Two coupled components (contact, and with the soft-core kernel on the
    total density): dJ/dV_c, dJ/da_11, dJ/da_22, dJ/da_12, dJ/dU and dJ/dpsi_0
    against central differences of the HIP forward, one V direction against
    the CPU oracle's."""
    c, V, psi0, target, a = two_comp_setup()
    dt, n, every = 5e-4, 24, 5
    J, gV, g, lam0 = grad2(torch_cuda, c, V, a, U, psi0, target, dt, n, every)
    assert abs(J - J2(fwd2(c, V, a, U, psi0, dt, n), target, c.dv)[0]) < 1e-14
    h = 1e-4

    def fd(**kw):
        Vx, ax, Ux, px = kw.get("V", V), kw.get("a", a), kw.get("U", U), kw.get("psi", psi0)
        return J2(fwd2(c, Vx, ax, Ux, px, dt, n), target, c.dv)[0]

    for comp in range(2):
        dV = np.zeros_like(V)
        dV[comp] = smooth_field(c, 600 + comp)
        f = (fd(V=V + h * dV) - fd(V=V - h * dV)) / (2 * h)
        ad = float(np.sum(gV * dV))
        assert abs(ad - f) <= 1e-6 * abs(f), (comp, ad, f)
    for k, idx in ((0, 0), (3, 1), (4, 2)):
        # Richardson-extrapolated central differences (steps 2h and h): the
        # a_12 derivative is small here, and the plain difference's h^2 term is
        # not negligible against it
        fs = []
        for hh in (2 * h, h):
            ap = list(a); am = list(a)
            ap[idx] += hh; am[idx] -= hh
            fs.append((fd(a=tuple(ap)) - fd(a=tuple(am))) / (2 * hh))
        f = (4 * fs[1] - fs[0]) / 3
        assert abs(g[k] - f) <= 1e-6 * abs(f), (k, g[k], f, fs)
    if U:
        f = (fd(U=U + h) - fd(U=U - h)) / (2 * h)
        assert abs(g[2] - f) <= 1e-6 * abs(f), (g[2], f)
    else:
        assert g[1] == 0.0 and g[2] == 0.0
    dpsi = np.stack([c.rand_psi(78, 1e-2), c.rand_psi(79, 1e-2)])
    f = (fd(psi=psi0 + h * dpsi) - fd(psi=psi0 - h * dpsi)) / (2 * h)
    ad = float(np.sum(np.real(np.conj(lam0) * dpsi)))
    assert abs(ad - f) <= 1e-6 * abs(f), (ad, f)
    # the FP64 reference: central differences of the CPU oracle along one direction
    dV = np.stack([smooth_field(c, 610), smooth_field(c, 611)])
    vt = c.vtilde(0.0, sc=(3, 1.5, U)) if U else None
    Js = []
    for sgn in (1, -1):
        out = orc.rk4(c.grid, orc.Ham(2, a, vtilde=vt), psi0, (V + sgn * h * dV).astype(np.complex128), dt, n)
        Js.append(J2(out, target, c.dv)[0])
    f = (Js[0] - Js[1]) / (2 * h)
    ad = float(np.sum(gV * dV))
    assert abs(ad - f) <= 1e-6 * abs(f), (ad, f)


def test_two_component_adjoint_mixed_precision(torch_cuda):
    c, V, psi0, target, a = two_comp_setup()
    dt, n, every = 5e-4, 24, 5
    _, g64, s64, _ = grad2(torch_cuda, c, V, a, 5.0, psi0, target, dt, n, every, prec=0)
    _, g32, s32, _ = grad2(torch_cuda, c, V, a, 5.0, psi0, target, dt, n, every, prec=1)
    rel = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    assert rel < 1e-4, rel
    for k in (0, 2, 3, 4):
        assert abs(s32[k] - s64[k]) <= 1e-4 * abs(s64[k]), (k, s32[k], s64[k])


@pytest.mark.parametrize("dims", ["1d", "2d"])
def test_adjoint_low_dimensional_grids(torch_cuda, dims):
    """A 1D grid (one x row, fewer rows than a block's pencils) with the 1D
    soft-core kernel, and a 2D grid with an in-plane dipole: dJ/dV, dJ/da, the
    kernel's strength and dJ/dpsi_0 against central differences of the HIP
    forward, one V direction against the oracle's."""
    if dims == "1d":
        c = Case(128, 1, 1, L=(12.0, 1.0, 1.0))
        ham = dict(a=(0.3,), sc_U=2.0, sc_R=1.1, sc_dim=1)
        kname, gk = "sc_U", 2
        vt_of = lambda h: c.vtilde(sc=(1, h["sc_R"], h["sc_U"]))
    else:
        c = Case(64, 32, 1, L=(12.0, 10.0, 1.0))
        ham = dict(a=(0.4,), a_dd=0.3, dip=(1.0, 0.0, 0.0))
        kname, gk = "a_dd", 1
        vt_of = lambda h: c.vtilde(h["a_dd"], h["dip"])
    V = configs.harmonic_trap(c.spec, (1.0, 1.2, 1.0)).real
    psi0 = configs.gaussian_noisy(c.spec, 91, width=1.5, noise=0.02)
    target = configs.normalise(np.exp(-0.4 * V) * np.exp(0.2j * c.rand_psi(92).real), c.dv)
    dt, n, every = 5e-4, 30, 7

    def forward(Vx, hx, px=psi0):
        with c.ctx() as ctx:
            ctx.set_hamiltonian(**hx)
            ctx.upload_vext(Vx.astype(np.complex128))
            ctx.upload_psi(px)
            ctx.rk4_steps(dt, n)
            return J_of(ctx.download_psi(), target, c.dv)[0]

    with c.ctx() as ctx:
        ctx.set_hamiltonian(**ham)
        ctx.upload_vext(V.astype(np.complex128))
        ctx.upload_psi(psi0)
        ctx.adjoint_forward(dt, n, every)
        J, ov = J_of(ctx.download_psi(), target, c.dv)
        lamT = torch_cuda.from_numpy(2 * ov * c.dv * target).cuda()
        gV = torch_cuda.empty(c.shape, dtype=torch_cuda.float64, device="cuda")
        lam0 = torch_cuda.empty_like(lamT)
        g = ctx.adjoint_backward(lamT.data_ptr(), lam0.data_ptr(), gV.data_ptr())
        ctx.sync()
        gV, lam0 = gV.cpu().numpy(), lam0.cpu().numpy()
    assert abs(J - forward(V, ham)) < 1e-14
    h = 1e-4
    for j in range(2):
        dV = smooth_field(c, 300 + j)
        f = (forward(V + h * dV, ham) - forward(V - h * dV, ham)) / (2 * h)
        ad = float(np.sum(gV * dV))
        assert abs(ad - f) <= 1e-6 * abs(f), (j, ad, f)
    for k, key in ((0, "a"), (gk, kname)):
        hp, hm = dict(ham), dict(ham)
        if key == "a":
            hp["a"], hm["a"] = (ham["a"][0] + h,), (ham["a"][0] - h,)
        else:
            hp[key], hm[key] = ham[key] + h, ham[key] - h
        f = (forward(V, hp) - forward(V, hm)) / (2 * h)
        assert abs(g[k] - f) <= 1e-6 * abs(f), (key, g[k], f)
    dpsi = c.rand_psi(93, 1e-2)
    f = (forward(V, ham, psi0 + h * dpsi) - forward(V, ham, psi0 - h * dpsi)) / (2 * h)
    ad = float(np.sum(np.real(np.conj(lam0) * dpsi)))
    assert abs(ad - f) <= 1e-6 * abs(f), (ad, f)
    dV = smooth_field(c, 310)
    Js = []
    for sgn in (1, -1):
        out = orc.rk4(c.grid, orc.Ham(1, ham["a"], vtilde=vt_of(ham)), psi0,
                      (V + sgn * h * dV).astype(np.complex128), dt, n)
        Js.append(J_of(out, target, c.dv)[0])
    f = (Js[0] - Js[1]) / (2 * h)
    ad = float(np.sum(gV * dV))
    assert abs(ad - f) <= 1e-6 * abs(f), (ad, f)


def test_two_component_adjoint_through_autograd(torch_cuda):
    """GPEEvolve with a two-component solver (V, psi0 with a leading axis of
    2, a = (a_11, a_22, a_12)) against the C ABI's gradient, soft-core kernel on."""
    torch = torch_cuda
    from fgpe_amd.adjoint import AdjointSolver, GPEEvolve, fidelity
    c, V, psi0, target, a = two_comp_setup()
    dt, n, U = 5e-4, 24, 5.0
    solver = AdjointSolver(c.spec, dt, n, 5, torch=torch, softcore=(3, 1.5), autotune=False, ncomp=2)
    Vt = torch.tensor(V, dtype=torch.float64, device="cuda", requires_grad=True)
    at = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    Ut = torch.tensor(U, dtype=torch.float64, requires_grad=True)
    p0 = torch.tensor(psi0, dtype=torch.complex128, device="cuda", requires_grad=True)
    tg = torch.tensor(target, dtype=torch.complex128, device="cuda")
    loss = fidelity(GPEEvolve().apply(Vt, at, Ut, p0, solver), tg, c.dv)
    loss.backward()
    J, gV, g, lam0 = grad2(torch, c, V, a, U, psi0, target, dt, n, 5)
    assert abs(loss.item() - J) < 1e-14
    assert np.abs(Vt.grad.cpu().numpy() - gV).max() <= 1e-13 * np.abs(gV).max()
    ga = at.grad.numpy()
    for i, k in enumerate((0, 3, 4)):
        assert abs(ga[i] - g[k]) <= 1e-13 * abs(g[k]), (i, ga[i], g[k])
    assert abs(Ut.grad.item() - g[2]) <= 1e-13 * abs(g[2])
    assert np.abs(p0.grad.cpu().numpy() - lam0).max() <= 1e-13 * np.abs(lam0).max()
