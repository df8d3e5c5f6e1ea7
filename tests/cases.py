"""Shared set-up for the parity tests: one grid description feeds both the
HIP context (through the C ABI) and the CPU oracle."""
import numpy as np

from fgpe_amd import configs
from oracle import oracle as orc


class Case:
    def __init__(self, nx, ny, nz, L=(16.0, 16.0, 16.0)):
        self.spec = configs.GridSpec(nx, ny, nz, *L)
        self.kx, self.ky, self.kz = self.spec.kaxes()
        if nz == 1:
            self.kz = np.zeros(1)
        self.dv = self.spec.dv
        self.grid = orc.Grid(nx, ny, nz, self.kx, self.ky, self.kz, self.dv)
        self.shape = self.spec.shape

    def ctx(self, precision=0):
        from fgpe_amd import Context
        c = Context(self.spec.nx, self.spec.ny, self.spec.nz, precision)
        c.set_kspace(self.kx, self.ky, self.kz, self.dv)
        return c

    def rand_psi(self, seed, amp=1.0):
        u = configs.splitmix64_uniform(seed, 2 * self.spec.npoints).reshape((2,) + self.shape)
        return amp * (u[0] + 1j * u[1])

    def vtilde(self, a_dd=0.0, dip=(0, 0, 1), sc=None):
        vt = np.zeros(self.shape)
        if a_dd:
            vt = vt + orc.vdd(self.grid, a_dd, dip)
        if sc is not None:
            dim, R, U = sc
            vt = vt + orc.vsoftcore(self.grid, dim, R, U)
        return vt


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())
