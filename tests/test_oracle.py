"""Pins the CPU oracle (oracle/fgpe_oracle.c) against golden vectors and
against what the mathematics fixes: brute-force DFT, numpy/scipy golden
values, closed forms, eigenfunctions and scalar reductions.  No GPU."""
import math
import os

import mpmath
import numpy as np
import pytest
from scipy import integrate, special

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
LIT = 19.739208802  # the reference's Laplacian literal (src/parallel.cu:199)


def grid(nx, ny, nz, L=(8.0, 8.0, 8.0)):
    ks = []
    for n, l in zip((nx, ny, nz), L):
        dx = l / n
        ks.append(orc.kaxis(n, 1.0 / (n * dx), 1.0 / (2 * dx), n > 1))
    dv = np.prod([l / n for n, l in zip((nx, ny, nz), L) if n > 1])
    return orc.Grid(nx, ny, nz, *ks, dv)


def coords(g, L):
    x = [(-l / 2 + (np.arange(n) + 0.5) * l / n) if n > 1 else np.zeros(1)
         for n, l in zip((g.nx, g.ny, g.nz), L)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    return X, Y, Z


def test_twiddle_against_mpmath():
    mpmath.mp.dps = 40
    for n in (3, 5, 7, 12, 96, 192, 256, 1000):
        for m in range(0, n, max(1, n // 17)):
            c, s = orc.twiddle(n, m)
            th = 2 * mpmath.pi * m / n
            assert abs(c - float(mpmath.cos(th))) <= 1.2e-16
            assert abs(s - float(mpmath.sin(th))) <= 1.2e-16


@pytest.mark.parametrize("n", list(range(1, 21)) + [24, 27, 32, 45, 49, 96])
def test_fft1d_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    ref_f = W @ x
    ref_b = W.conj() @ x
    scale = np.abs(x).sum()
    assert np.abs(orc.fft1d(x, -1) - ref_f).max() <= 1e-14 * scale
    assert np.abs(orc.fft1d(x, +1) - ref_b).max() <= 1e-14 * scale


def test_fft3d_golden_vectors():
    d = np.load(os.path.join(GOLD, "fft_golden.npz"))
    i = 0
    while f"x{i}" in d:
        x = d[f"x{i}"]
        for sign, key in ((-1, "f"), (+1, "b")):
            ref = d[f"{key}{i}"]
            got = orc.fft3d(x, sign)
            assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), (i, sign)
        i += 1
    assert i >= 5


def test_fft3d_parseval_and_inverse():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((6, 10, 12)) + 1j * rng.standard_normal((6, 10, 12))
    f = orc.fft3d(x, -1)
    assert abs(np.vdot(f, f).real - x.size * np.vdot(x, x).real) < 1e-12 * x.size * np.vdot(x, x).real
    back = orc.fft3d(f, +1)
    assert np.abs(back - x.size * x).max() < 1e-12 * x.size
    imp = np.zeros((4, 4, 4), complex)
    imp[0, 0, 0] = 1.0
    assert np.abs(orc.fft3d(imp, -1) - 1.0).max() == 0.0


@pytest.mark.parametrize("n", [1, 2, 5, 8, 9, 64])
def test_kaxis_integer_closed_form(n):
    L = 16.0
    dx = L / n
    k = orc.kaxis(n, 1.0 / (n * dx), 1.0 / (2 * dx), n > 1)
    if n == 1:
        assert k[0] == 0.0
        return
    if n % 2 == 0:
        # frequencies are integers / L, Nyquist bin positive (i = n/2 -> +n/2)
        m = np.array([i if i <= n // 2 else i - n for i in range(n)], float)
    else:
        # odd n: the reference's -KMAX + (i - n/2) DK (src/baseHamiltonian.cu:62)
        # lands on half-integer frequencies above n/2
        m = np.array([i if i <= n // 2 else i - n + 0.5 for i in range(n)], float)
    assert np.abs(k * L - m).max() < 1e-12
    assert np.all(k[: n // 2 + 1] >= 0)


@pytest.mark.parametrize("mode", [(1, 0, 0), (2, -3, 1), (0, 4, -2), (-3, 3, 3)])
def test_laplacian_plane_wave_uses_literal(mode):
    L = (8.0, 8.0, 8.0)
    g = grid(8, 8, 8, L)
    X, Y, Z = coords(g, L)
    m = np.array(mode, float)
    psi = np.exp(2j * np.pi * (m[0] * X / L[0] + m[1] * Y / L[1] + m[2] * Z / L[2]))
    lam = LIT * float(np.sum((m / np.array(L)) ** 2))
    out = orc.laplace(g, psi)
    # 1e-13 relative separates the literal from 2*pi^2 (9e-12 apart)
    assert np.abs(out - lam * psi).max() <= 1e-13 * lam


def test_vdd_closed_form():
    L = (8.0, 8.0, 16.0)
    g = grid(8, 8, 16, L)
    a_dd = 0.37
    d = np.array([0.0, 0.6, 0.8])
    vt = orc.vdd(g, a_dd, d)
    KZ, KY, KX = np.meshgrid(g.kz, g.ky, g.kx, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        cos2 = ((KX * d[0] + KY * d[1] + KZ * d[2]) / kk) ** 2
        ref = 4 * np.pi * a_dd * (3 * cos2 - 1)
    ref[0, 0, 0] = 0.0
    assert vt[0, 0, 0] == 0.0
    assert np.abs(vt - ref).max() <= 1e-13 * 4 * np.pi * a_dd


def test_vsoftcore_against_quadrature_and_j1_golden():
    L = (8.0, 8.0, 8.0)
    R, U = 1.5, 5.0
    g3 = grid(8, 8, 8, L)
    v3 = orc.vsoftcore(g3, 3, R, U)
    assert abs(v3[0, 0, 0] - U * 4.0 / 3.0 * np.pi * R ** 3) < 1e-12
    for (iz, iy, ix) in [(0, 0, 1), (0, 2, 3), (1, 1, 1), (4, 0, 0)]:
        k = 2 * np.pi * math.sqrt(g3.kx[ix] ** 2 + g3.ky[iy] ** 2 + g3.kz[iz] ** 2)
        # 3D Fourier transform of U*theta(R - r): 4 pi U int_0^R r sin(k r)/k dr
        ref = 4 * np.pi * U * integrate.quad(lambda r: r * math.sin(k * r) / k, 0, R, epsabs=1e-15)[0]
        assert abs(v3[iz, iy, ix] - ref) <= 1e-11 * abs(v3[0, 0, 0])
    g2 = grid(16, 16, 1, (8.0, 8.0, 1.0))
    v2 = orc.vsoftcore(g2, 2, R, U)
    assert abs(v2[0, 0, 0] - U * np.pi * R ** 2) < 1e-12
    for (iy, ix) in [(0, 1), (3, 2), (8, 8)]:
        k = 2 * np.pi * math.hypot(g2.kx[ix], g2.ky[iy])
        # 2D Fourier transform of a disk: 2 pi U int_0^R r J0(k r) dr
        ref = 2 * np.pi * U * integrate.quad(lambda r: r * special.j0(k * r), 0, R, epsabs=1e-15)[0]
        assert abs(v2[0, iy, ix] - ref) <= 1e-11 * abs(v2[0, 0, 0])
    gold = np.load(os.path.join(GOLD, "j1_golden.npz"))
    KY, KX = np.meshgrid(g2.ky, g2.kx, indexing="ij")
    kk = 2 * np.pi * np.sqrt(KX ** 2 + KY ** 2)
    mask = kk > 1e-6
    j1_from_oracle = v2[0][mask] * kk[mask] / (2 * np.pi * R * U)
    ref = np.interp(np.round(kk[mask] * R, 15), gold["args"], gold["j1"])
    assert np.abs(j1_from_oracle - ref).max() < 1e-14
    g1 = grid(16, 1, 1, (8.0, 1.0, 1.0))
    v1 = orc.vsoftcore(g1, 1, R, U)
    for ix in (1, 3, 7):
        k = 2 * np.pi * abs(g1.kx[ix])
        ref = integrate.quad(lambda x: U * math.cos(k * x), -R, R, epsabs=1e-15)[0]
        assert abs(v1[0, 0, ix] - ref) <= 1e-11 * 2 * U * R


def test_convolve_single_mode():
    L = (8.0, 8.0, 16.0)
    g = grid(8, 8, 16, L)
    X, Y, Z = coords(g, L)
    vt = orc.vdd(g, 0.5, (0.0, 0.0, 1.0))
    q = (1, 0, 2)
    arg = 2 * np.pi * (q[0] * X / L[0] + q[2] * Z / L[2])
    n0, eps = 0.7, 0.05
    phi = orc.convolve(g, vt, n0 + eps * np.cos(arg))
    vq = vt[q[2], 0, q[0]]
    assert np.abs(phi - eps * vq * np.cos(arg)).max() < 1e-14


def test_hpsi_plane_wave_closed_form():
    L = (8.0, 8.0, 8.0)
    g = grid(8, 8, 8, L)
    X, Y, Z = coords(g, L)
    m = (1, -2, 0)
    A = 0.3
    psi = A * np.exp(2j * np.pi * (m[0] * X + m[1] * Y) / 8.0)
    v0 = 0.4 + 0.1j
    V = np.full(g.shape, v0)
    a = 0.8
    h = orc.Ham(1, (a,))
    hmu, hen = orc.hpsi(g, h, 0.0, psi, V)
    ke = LIT * (1 + 4) / 64.0
    assert np.abs(hmu - (ke + v0 + 4 * np.pi * a * A * A) * psi).max() < 1e-14
    assert np.abs(hen - (ke + v0 + 2 * np.pi * a * A * A) * psi).max() < 1e-14


def test_hpsi_moving_frame_plane_wave():
    """-vpx F^-1[k_x F psi]/N on a plane wave e^{2 pi i m.x/L} is -vpx (m_x/L) psi
    in the full form (alg_calcHpsi) and absent from the MU form."""
    L = (8.0, 8.0, 8.0)
    g = grid(8, 8, 8, L)
    X, Y, Z = coords(g, L)
    m = (3, -1, 2)
    psi = 0.5 * np.exp(2j * np.pi * (m[0] * X + m[1] * Y + m[2] * Z) / 8.0)
    V = np.zeros(g.shape, complex)
    lam = 1.7
    h = orc.Ham(1, (0.0,), vpx=(lam,))
    hmu, hen = orc.hpsi(g, h, 0.0, psi, V)
    ke = LIT * (9 + 1 + 4) / 64.0
    assert np.abs(hmu - (ke - lam * 3 / 8.0) * psi).max() < 1e-13
    assert np.abs(hen - (ke - lam * 3 / 8.0) * psi).max() < 1e-13
    mu_only = orc.hpsi(g, h, 0.0, psi, V, energy=False)
    assert np.abs(mu_only - ke * psi).max() < 1e-13
    # two components: each its own factor
    h2 = orc.Ham(2, (0.0, 0.0, 0.0), vpx=(lam, -0.4))
    two = np.stack([psi, 2.0 * psi])
    hm2, he2 = orc.hpsi(g, h2, 0.0, two, np.stack([V, V]))
    assert np.abs(hm2[0] - (ke - lam * 3 / 8.0) * psi).max() < 1e-13
    assert np.abs(hm2[1] - (ke + 0.4 * 3 / 8.0) * 2.0 * psi).max() < 1e-13


def test_rk4_moving_frame_enters_stage_four_only():
    """One component: stages 1-3 apply the MU form (eps0 = kinetic), stage 4
    the full form (eps1 = eps0 - vpx k_x); on a plane wave the step multiplies
    by 1 + (-i dt / 6)(eps0 (1 + 2 a1 + 2 a2) + eps1 a3), a1 = 1 + z/2,
    a2 = 1 + z a1 / 2, a3 = 1 + z a2, z = -i dt eps0.  Two components never
    apply it in RK4."""
    L = (8.0, 8.0, 8.0)
    g = grid(8, 8, 8, L)
    X, Y, Z = coords(g, L)
    psi = np.exp(2j * np.pi * (2 * X - Y) / 8.0)
    V = np.zeros(g.shape, complex)
    dt, n, lam = 0.01, 4, 3.0
    e0 = LIT * 5 / 64.0
    e1 = e0 - lam * 2 / 8.0
    z = -1j * dt * e0
    a1 = 1 + z / 2
    a2 = 1 + z * a1 / 2
    a3 = 1 + z * a2
    A = 1 + (-1j * dt / 6) * (e0 * (1 + 2 * a1 + 2 * a2) + e1 * a3)
    out = orc.rk4(g, orc.Ham(1, (0.0,), vpx=(lam,)), psi, V, dt, n)
    assert np.abs(out - A ** n * psi).max() < 1e-13
    # without the quirk the amplification would be the full-form RK4 polynomial
    zf = -1j * dt * e1
    full = 1 + zf + zf ** 2 / 2 + zf ** 3 / 6 + zf ** 4 / 24
    assert np.abs(out - full ** n * psi).max() > 1e-6
    two = orc.rk4(g, orc.Ham(2, (0.0, 0.0, 0.0), vpx=(lam, lam)), np.stack([psi, psi]),
                  np.stack([V, V]), dt, n)
    R0 = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.abs(two[0] - R0 ** n * psi).max() < 1e-13


def test_quench_ramp():
    prms = (1.0, 2.0, 3.0, 1.0, 2.0, 4.0)
    assert orc.quench_a(0.0, prms) == 3.0      # t <= 0 falls through to a2
    assert orc.quench_a(0.5, prms) == 1.5
    assert orc.quench_a(1.5, prms) == 2.0
    assert orc.quench_a(3.0, prms) == 2.5
    assert orc.quench_a(9.0, prms) == 3.0


def test_rk4_mode_amplification():
    L = (8.0, 8.0, 8.0)
    g = grid(8, 8, 8, L)
    X, Y, Z = coords(g, L)
    m = (2, 1, -1)
    psi = np.exp(2j * np.pi * (m[0] * X + m[1] * Y + m[2] * Z) / 8.0)
    dt, n = 0.01, 5
    E = LIT * 6 / 64.0
    h = orc.Ham(1, (0.0,))
    V = np.zeros(g.shape, complex)
    for imag in (False, True):
        z = -dt * E * (1.0 if imag else 1j)
        R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        out = orc.rk4(g, h, psi, V, dt, n, imaginary=imag, npart=float(np.sum(abs(psi) ** 2) * g.dv))
        ref = R ** n * psi
        if imag:  # renormalised each step to the initial norm
            ref = psi * (R / abs(R)) ** n
        assert np.abs(out - ref).max() < 1e-13


def test_rk4_uniform_nonlinear_scalar():
    g = grid(4, 4, 4)
    A0 = 0.5 + 0.2j
    psi = np.full(g.shape, A0)
    v0, a, dt, n = 0.3, 0.9, 0.02, 7
    V = np.full(g.shape, v0 + 0j)
    out = orc.rk4(g, orc.Ham(1, (a,)), psi, V, dt, n)

    def f(y):  # i dy/dt = (v0 + 4 pi a |y|^2) y
        return -1j * (v0 + 4 * np.pi * a * abs(y) ** 2) * y

    y = A0
    for _ in range(n):  # textbook RK4 on one complex number
        k1 = f(y); k2 = f(y + dt / 2 * k1); k3 = f(y + dt / 2 * k2); k4 = f(y + dt * k3)
        y = y + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert np.abs(out - y).max() < 1e-14


def test_heavy_ball_uniform_state():
    g = grid(4, 4, 4)
    psi = np.full(g.shape, 0.3 + 0j)
    v0, a = 0.25, 0.6
    V = np.full(g.shape, v0 + 0j)
    npart = 2.0
    out, n, hist = orc.imag_heavy_ball(g, orc.Ham(1, (a,)), psi, V, 0.01, 0.9, npart, 1e-12, 50)
    A2 = npart / (g.N * g.dv)
    assert np.abs(np.abs(out) ** 2 - A2).max() < 1e-14
    assert abs(hist[-1, 1] - (v0 + 4 * np.pi * a * A2)) < 1e-13
    assert abs(hist[-1, 2] - (v0 + 2 * np.pi * a * A2)) < 1e-13
    # E of step k pairs psi_k with H(psi_{k-1}) (src/oneComponentGPSolver.cu:72-73):
    # exact from step 2 on, so |E3 - E2| = 0 stops the loop at step 3
    assert abs(hist[0, 2] - (v0 + 2 * np.pi * a * A2)) > 1e-3
    assert n == 3


def test_norm_gaussian():
    L = (12.0, 12.0, 12.0)
    g = grid(24, 24, 24, L)
    X, Y, Z = coords(g, L)
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 2).astype(complex)
    assert abs(orc.norm(g, psi) - np.pi ** 1.5) < 1e-12
    hmu = 2.0 * psi
    mu, en = orc.obs2(g, np.pi ** 1.5, psi, hmu, 3.0 * psi)
    assert abs(mu - 2.0) < 1e-13 and abs(en - 3.0) < 1e-13


def test_harmonic_ground_state_energy():
    """g = 0 isotropic trap: imaginary time converges to E = 3/2 (hbar w = 1)."""
    L = (10.0, 10.0, 10.0)
    g = grid(20, 20, 20, L)
    X, Y, Z = coords(g, L)
    V = (0.5 * (X ** 2 + Y ** 2 + Z ** 2)).astype(complex)
    psi = np.exp(-((X - 0.3) ** 2 + Y ** 2 + 2 * Z ** 2) / 3).astype(complex)
    h = orc.Ham(1, (0.0,))
    # dt * (E_kin,max + V_max) = 0.02 * (59.2 + 37.5) < 2.78: RK4 stays stable
    out = orc.rk4(g, h, psi, V, 0.02, 1000, imaginary=True, npart=1.0)
    hmu, hen = orc.hpsi(g, h, 0.0, out, V)
    mu, en = orc.obs2(g, orc.norm(g, out), out, hmu, hen)
    assert abs(en - 1.5) < 1e-7 and abs(mu - 1.5) < 1e-7


def test_heavy_ball_two_component_uniform_scalar():
    """Two uniform components reduce the 3D loop to a recursion on four
    numbers (amplitudes and previous amplitudes), run here by hand."""
    g = grid(4, 4, 4)
    Vol = g.N * g.dv
    a11, a22, a12 = 0.5, 0.7, 0.3
    v = (0.2, 0.35)
    dt, beta, npart, niter = 0.01, 0.6, 1.5, 6
    A = [0.3, 0.5]
    psi = np.stack([np.full(g.shape, A[0] + 0j), np.full(g.shape, A[1] + 0j)])
    V = np.stack([np.full(g.shape, v[0] + 0j), np.full(g.shape, v[1] + 0j)])
    out, n, hist = orc.imag_heavy_ball(g, orc.Ham(2, (a11, a22, a12)), psi, V, dt, beta, npart,
                                       0.0, niter)
    assert n == niter
    old = list(A)
    mu = [0.0, 0.0]
    aa = [[a11, a12], [a12, a22]]
    for it in range(niter):
        n_ = [A[0] ** 2, A[1] ** 2]
        h = [(v[c] + 4 * np.pi * (aa[c][c] * n_[c] + aa[c][1 - c] * n_[1 - c])) * A[c] for c in range(2)]
        he = [(v[c] + 2 * np.pi * (aa[c][c] * n_[c] + aa[c][1 - c] * n_[1 - c])) * A[c] for c in range(2)]
        B = [(1 + mu[c] * dt) * A[c] - dt * h[c] + beta * (A[c] - old[c]) for c in range(2)]
        old = list(A)
        for c in range(2):
            nrm = B[c] ** 2 * Vol
            s = (npart / nrm) ** 0.5
            A[c] = s * B[c]
            mu[c] = A[c] * h[c] * Vol / nrm
            en = A[c] * he[c] * Vol / nrm
            assert abs(hist[it, 3 * c] - nrm) < 1e-12 * nrm
            assert abs(hist[it, 3 * c + 1] - mu[c]) < 1e-12 * abs(mu[c])
            assert abs(hist[it, 3 * c + 2] - en) < 1e-12 * abs(en)
    assert abs(out[0, 0, 0, 0] - A[0]) < 1e-14 and abs(out[1, 3, 2, 1] - A[1]) < 1e-14


def test_two_component_hpsi_uniform():
    g = grid(4, 4, 4)
    A = (0.4 + 0.1j, 0.2 - 0.3j)
    a11, a22, a12 = 0.5, 0.7, 0.3
    psi = np.stack([np.full(g.shape, A[0]), np.full(g.shape, A[1])])
    V = np.stack([np.full(g.shape, 0.1 + 0.05j), np.full(g.shape, -0.2 + 0.0j)])
    hmu, hen = orc.hpsi(g, orc.Ham(2, (a11, a22, a12)), 0.0, psi, V)
    n1, n2 = abs(A[0]) ** 2, abs(A[1]) ** 2
    ref1 = (0.1 + 0.05j + 4 * np.pi * (a11 * n1 + a12 * n2)) * A[0]
    ref2 = (-0.2 + 4 * np.pi * (a22 * n2 + a12 * n1)) * A[1]
    assert np.abs(hmu[0] - ref1).max() < 1e-14 and np.abs(hmu[1] - ref2).max() < 1e-14
    e1 = (0.1 + 0.05j + 2 * np.pi * (a11 * n1 + a12 * n2)) * A[0]
    assert np.abs(hen[0] - e1).max() < 1e-14


# ---- the example user pointwise term (fgpe_amd/csrc/user_term.hpp) --------
def test_user_term_loss_decay_closed_form():
    """Three-body loss alone on a uniform state: dn/dt = -K3 n^3, so
    n(t) = n0 / sqrt(1 + 2 K3 n0^2 t); the phase stays put."""
    g = grid(4, 4, 4)
    A0 = 0.8 + 0.3j
    K3, dt, n = 0.6, 1e-3, 400
    out = orc.rk4(g, orc.Ham(1, (0.0,), user=(0.0, 0.0, K3, 0.0)), np.full(g.shape, A0),
                  np.zeros(g.shape, complex), dt, n)
    n0, t = abs(A0) ** 2, n * dt
    assert np.abs(np.abs(out) ** 2 - n0 / np.sqrt(1 + 2 * K3 * n0 ** 2 * t)).max() < 1e-11
    assert np.abs(np.angle(out) - np.angle(A0)).max() < 1e-13


def test_user_term_modulated_coupling_phase_closed_form():
    """dg(t) = u0 sin(u1 t) alone on a uniform state: |psi| is kept and the
    phase is -n u0 (1 - cos(u1 t)) / u1 (the stage times t, t + dt/2, t + dt
    enter through RK4)."""
    g = grid(4, 4, 4)
    A0 = 0.6 - 0.2j
    u0, u1, dt, n = 1.3, 2.1, 1e-3, 500
    out = orc.rk4(g, orc.Ham(1, (0.0,), user=(u0, u1)), np.full(g.shape, A0), np.zeros(g.shape, complex),
                  dt, n)
    nn, t = abs(A0) ** 2, n * dt
    ref = A0 * np.exp(-1j * nn * u0 * (1 - np.cos(u1 * t)) / u1)
    assert np.abs(out - ref).max() < 1e-11


def test_user_term_ramp_and_energy_halves():
    """On a uniform state (no kinetic term) the ramp adds u3 iz psi to both
    operators, the modulated coupling dg n psi to H and dg n psi / 2 to H_en."""
    g = grid(4, 6, 8)
    A0 = 0.7 + 0.1j
    psi = np.full(g.shape, A0)
    V = np.zeros(g.shape, complex)
    u = (0.9, 1.7, 0.0, 0.25)
    t = 0.4
    h0, e0 = orc.hpsi(g, orc.Ham(1, (0.0,)), t, psi, V)
    h1, e1 = orc.hpsi(g, orc.Ham(1, (0.0,), user=u), t, psi, V)
    iz = np.arange(g.nz)[:, None, None] * np.ones(g.shape)
    dg = u[0] * np.sin(u[1] * t) * abs(A0) ** 2
    assert np.abs((h1 - h0) - (dg + u[3] * iz) * psi).max() < 1e-14
    assert np.abs((e1 - e0) - (0.5 * dg + u[3] * iz) * psi).max() < 1e-14


def test_user_term_drive_pointwise_and_component_sign():
    """The second test term on a plane wave (so the kinetic part is not
    uniform) and two components: H and H_en each gain
    sgn_c A exp(-gamma t) cos(kx ix + ky iy) psi_c, + for component 0 and -
    for component 1, with ix, iy the grid indices (x fastest)."""
    g = grid(6, 4, 3)
    x = np.arange(g.nx)[None, None, :] * np.ones(g.shape)
    y = np.arange(g.ny)[None, :, None] * np.ones(g.shape)
    psi = np.stack([np.exp(2j * np.pi * x / g.nx) * (0.5 + 0.1j), np.exp(-2j * np.pi * y / g.ny) * (0.3 - 0.2j)])
    V = np.zeros(psi.shape, complex)
    u = (0.0, 0.0, 0.0, 0.0, 1.4, 0.7, 0.35, 0.9)
    t = 0.3
    h0, e0 = orc.hpsi(g, orc.Ham(2, (0.01, 0.02, 0.005)), t, psi, V)
    h1, e1 = orc.hpsi(g, orc.Ham(2, (0.01, 0.02, 0.005), user=u, user_id=2), t, psi, V)
    w = u[4] * np.exp(-u[5] * t) * np.cos(u[6] * x + u[7] * y)
    for c, sgn in ((0, 1.0), (1, -1.0)):
        assert np.abs((h1[c] - h0[c]) - sgn * w * psi[c]).max() < 1e-14
        assert np.abs((e1[c] - e0[c]) - sgn * w * psi[c]).max() < 1e-14
    # a transposed index (iy for ix) would fail: the lattice is not symmetric
    wt = u[4] * np.exp(-u[5] * t) * np.cos(u[6] * y + u[7] * x)
    assert np.abs((h1[0] - h0[0]) - wt * psi[0]).max() > 1e-3
