"""Grids, seeded initial states and traps for the benchmark workloads.

Units are oscillator units (hbar = m = 1, lengths in a_ho, energies in
hbar*omega): the code units of the reference after Params::calculateLengthScales
(src/params.cu:223-279).  The contact term of the reference is 4*pi*a*|psi|^2
(hamiltonians/BEC_oneComponent.cu:145), so a coupling g is a = g/(4*pi).

Random numbers come from SplitMix64 (Steele, Lea, Flood 2014) mapped to
[-1, 1), so every consumer that starts from the same seed reads identical
bytes (the reference's std::mt19937 + uniform_real_distribution at
src/psi.cu:7,93 is implementation-defined).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

PI = 3.141592653589793  # hpp/header.hpp:28

_MASK = (1 << 64) - 1


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [-1, 1) from SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed & _MASK) + idx * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def kaxis(n: int, L: float) -> np.ndarray:
    """k axis in cycles per length, as BaseHamiltonian::InitializeKspace
    (src/baseHamiltonian.cu:61-68) with DK = 1/(n DX), KMAX = 1/(2 DX)
    (src/params.cu:206-211): k[i] = i*DK for i <= n/2, else -KMAX + (i-n/2)*DK."""
    dx = L / n
    dk = 1.0 / (n * dx)
    kmax = 1.0 / (2.0 * dx)
    k = np.empty(n)
    for i in range(n):
        k[i] = i * dk if i <= n // 2 else -kmax + (i - n // 2) * dk
    if n == 1:
        k[:] = 0.0
    return k


def xaxis(n: int, L: float) -> np.ndarray:
    """Cell-centred positions -XMAX + i DX + DX/2 (src/baseHamiltonian.cu:43-47)."""
    dx = L / n
    return -L / 2 + np.arange(n) * dx + 0.5 * dx


@dataclasses.dataclass
class GridSpec:
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def npoints(self):
        return self.nx * self.ny * self.nz

    @property
    def dv(self):
        """dx dy dz over the axes the grid has (dx dy in 2D, dx in 1D)."""
        if self.nz > 1:
            return (self.lx / self.nx) * (self.ly / self.ny) * (self.lz / self.nz)
        if self.ny > 1:
            return (self.lx / self.nx) * (self.ly / self.ny)
        return self.lx / self.nx

    def kaxes(self):
        return (kaxis(self.nx, self.lx), kaxis(self.ny, self.ly),
                kaxis(self.nz, self.lz) if self.nz > 1 else np.zeros(1))

    def mesh(self):
        x = xaxis(self.nx, self.lx)
        y = xaxis(self.ny, self.ly)
        z = xaxis(self.nz, self.lz) if self.nz > 1 else np.zeros(1)
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        return X, Y, Z


def harmonic_trap(grid: GridSpec, omega=(1.0, 1.0, 1.0)) -> np.ndarray:
    """V = 1/2 sum_i omega_i^2 x_i^2 as a complex array (imag 0), the form of
    Vext::initHarmonic (src/vext.cu:72-85) in oscillator units."""
    X, Y, Z = grid.mesh()
    V = 0.5 * (omega[0] ** 2 * X ** 2 + omega[1] ** 2 * Y ** 2 + omega[2] ** 2 * Z ** 2)
    return V.astype(np.complex128)


def normalise(psi: np.ndarray, dv: float, npart: float = 1.0) -> np.ndarray:
    """Psi::normalize (src/psi.cu:351-374): scale to sum |psi|^2 DV = npart."""
    n = float(np.sum(np.abs(psi) ** 2)) * dv
    return psi * math.sqrt(npart / n)


def gaussian_noisy(grid: GridSpec, seed: int, width=1.0, noise=0.01, offset=(0.0, 0.0, 0.0)):
    """exp(-r^2/(2 w^2)) * (1 + noise*u), u in U[-1,1) (real), normalised to 1."""
    X, Y, Z = grid.mesh()
    r2 = (X - offset[0]) ** 2 + (Y - offset[1]) ** 2 + (Z - offset[2]) ** 2
    u = splitmix64_uniform(seed, grid.npoints).reshape(grid.shape)
    psi = np.exp(-r2 / (2 * width ** 2)) * (1.0 + noise * u)
    return normalise(psi.astype(np.complex128), grid.dv)


def thomas_fermi_noisy(grid: GridSpec, seed: int, g: float, omega, noise=1e-3):
    """sqrt(max(0, mu_TF - V)/g) + noise*(u1 + i u2), normalised to 1; mu_TF
    from the 3D Thomas-Fermi normalisation for N = 1."""
    wbar3 = omega[0] * omega[1] * omega[2]
    # N = (8 pi / 15) (2 mu)^{5/2} / (g wbar^3)
    mu = 0.5 * (15.0 * g * wbar3 / (8.0 * PI)) ** 0.4
    V = harmonic_trap(grid, omega).real
    base = np.sqrt(np.maximum(0.0, mu - V) / g)
    u = splitmix64_uniform(seed, 2 * grid.npoints).reshape((2,) + grid.shape)
    psi = base + noise * (u[0] + 1j * u[1])
    return normalise(psi.astype(np.complex128), grid.dv)


# ---- the workloads ------------------------------------------------------

@dataclasses.dataclass
class Workload:
    name: str
    grid: GridSpec
    ncomp: int
    a: tuple            # contact lengths (a = g / 4 pi); 2 comp: a11, a22, a12
    dt: float
    nsteps: int
    imaginary: bool
    omega: tuple
    a_dd: float = 0.0
    dip: tuple = (0.0, 0.0, 1.0)
    sc_U: float = 0.0
    sc_R: float = 0.0
    seed: int = 1

    def initial_state(self):
        if self.name.startswith("contact_imag"):
            return gaussian_noisy(self.grid, self.seed)
        if self.ncomp == 2:
            p1 = gaussian_noisy(self.grid, self.seed, offset=(1.0, 0, 0), noise=0.0)
            p2 = gaussian_noisy(self.grid, self.seed + 100, offset=(-1.0, 0, 0), noise=0.0)
            return np.stack([p1, p2])
        g = 4 * PI * self.a[0]
        return thomas_fermi_noisy(self.grid, self.seed, g, self.omega)

    def trap(self):
        V = harmonic_trap(self.grid, self.omega)
        return np.stack([V, V]) if self.ncomp == 2 else V


def contact_imag_64():
    """64^3 one-component contact GPE, imaginary-time RK4 with renormalisation,
    L = 16, omega = (1,1,1), g = 100, dt = 1e-3, 2000 steps."""
    return Workload("contact_imag_64", GridSpec(64, 64, 64, 16.0, 16.0, 16.0), 1,
                    (100.0 / (4 * PI),), 1e-3, 2000, True, (1.0, 1.0, 1.0), seed=1)


def dipolar_real(nx=128, ny=128, nz=256, nsteps=4000):
    """Dipolar real-time RK4: cigar trap omega = (1,1,0.25), g = 200,
    eps_dd = 1.3 (a_dd = 1.3 a), dipoles along z, dt = 5e-4; box 16 x 16 x 32."""
    a = 200.0 / (4 * PI)
    return Workload(f"dipolar_real_{nx}x{ny}x{nz}", GridSpec(nx, ny, nz, 16.0, 16.0, 32.0), 1,
                    (a,), 5e-4, nsteps, False, (1.0, 1.0, 0.25), a_dd=1.3 * a, seed=2)


def twocomp_softcore_256(n=256, nsteps=2000):
    """Two-component coupled GPE, g11 = g22 = 100, g12 = 80, omega = (1,1,1),
    soft-core U0 = 5, R = 1.5 on n1 + n2, dt = 5e-4; box 16^3."""
    f = 1.0 / (4 * PI)
    return Workload(f"twocomp_softcore_{n}", GridSpec(n, n, n, 16.0, 16.0, 16.0), 2,
                    (100.0 * f, 100.0 * f, 80.0 * f), 5e-4, nsteps, False, (1.0, 1.0, 1.0),
                    sc_U=5.0, sc_R=1.5, seed=3)


# ---- batched ground-state sweep ------------------------------------------

SWEEP_SHAPES = ((64, 64, 64), (96, 96, 96), (128, 128, 128), (64, 64, 192), (512, 512, 1))


@dataclasses.dataclass
class SweepInstance:
    """One imaginary-time ground-state problem of the sweep."""
    index: int
    grid: GridSpec
    a: float            # contact length, g = 4 pi a
    omega: tuple
    a_dd: float = 0.0
    dip: tuple = (0.0, 0.0, 1.0)
    sc_U: float = 0.0
    sc_R: float = 0.0
    sc_dim: int = 3
    seed: int = 0

    @property
    def kind(self):
        return "dipolar" if self.a_dd != 0.0 else "softcore"

    def initial_state(self):
        return gaussian_noisy(self.grid, self.seed, width=1.0, noise=0.01)

    def trap(self):
        return harmonic_trap(self.grid, self.omega)


def sweep_instances(n=64, seed=5, shapes=SWEEP_SHAPES):
    """n instances: shapes cycled through `shapes`, grid spacing 0.25 in 3D
    (box 16 per 64 points) and 1/8 in 2D (box 64); g ~ U[50, 300]; trap
    omega = (l, l, 1) in 3D (pancake, aspect l ~ U[0.2, 1]) and (1, l) in 2D;
    even instances dipolar with eps_dd ~ U[0.5, 1.5] (dipoles along z in 3D,
    along x in 2D, where a z dipole gives a constant -4 pi a_dd), odd ones a
    soft-core term U0 ~ U[1, 10], R ~ U[1, 2] (3D or 2D transform)."""
    u = 0.5 * (splitmix64_uniform(seed, 4 * n).reshape(n, 4) + 1.0)   # U[0, 1)
    out = []
    for i in range(n):
        nx, ny, nz = shapes[i % len(shapes)]
        if nz > 1:
            grid = GridSpec(nx, ny, nz, nx * 0.25, ny * 0.25, nz * 0.25)
        else:
            grid = GridSpec(nx, ny, 1, nx / 8.0, ny / 8.0, 1.0)
        g = 50.0 + 250.0 * u[i, 0]
        lam = 0.2 + 0.8 * u[i, 1]
        omega = (lam, lam, 1.0) if nz > 1 else (1.0, lam, 0.0)
        a = g / (4 * PI)
        inst = SweepInstance(i, grid, a, omega, seed=1000 + i)
        if i % 2 == 0:
            inst.a_dd = (0.5 + 1.0 * u[i, 2]) * a
            inst.dip = (0.0, 0.0, 1.0) if nz > 1 else (1.0, 0.0, 0.0)
        else:
            inst.sc_U = 1.0 + 9.0 * u[i, 2]
            inst.sc_R = 1.0 + 1.0 * u[i, 3]
            inst.sc_dim = 3 if nz > 1 else 2
        out.append(inst)
    return out
