"""Writes the oracle's long-horizon results as small fixtures, for the GPU
parity tests of the benchmark configurations over their full horizons (the
oracle takes minutes there, too long to run inside a GPU test):

  config1_64cubed_2000.npz
      64^3 contact GPE, imaginary-time RK4 renormalised to N = 1, dt = 1e-3,
      2000 steps (the full horizon).
  config2_<nx>x<ny>x<nz>_<nsteps>.npz
      the dipolar real-time RK4 physics (eps_dd = 1.3, cigar trap,
      dt = 5e-4, box 16 x 16 x 32) on the given grid over nsteps steps.

At every checkpoint the fixture keeps psi on every 4th (config 2: every
(nx/8)th) grid point per axis (`psi_s`), max |psi| over the whole grid (`psimax`, the scale of the
relative error), sum |psi| dv (`absum`), and the norm, mu and E of the
full state from the oracle's H (`obs`).  The computation is the oracle's
(oracle/fgpe_oracle.c, built with contraction off), on this machine's cores.

Run:  python tests/golden/make_horizon.py config1
      python tests/golden/make_horizon.py config2 NX NY NZ NSTEPS EVERY [SNAPSHOT_DIR]
SNAPSHOT_DIR: read psi_<step>.npy files that an identical oracle run already
wrote there instead of recomputing them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fgpe_amd import configs  # noqa: E402
from oracle import oracle as orc  # noqa: E402

SUB = 4   # sampling stride per axis (config 2: nx / 8, so its fixtures stay small)


def summary(og, ham, psi, V, t, sub=SUB):
    hmu, hen = orc.hpsi(og, ham, t, psi, V)
    nrm = orc.norm(og, psi)
    mu, en = orc.obs2(og, nrm, psi, hmu, hen)
    return dict(psi_s=psi[::sub, ::sub, ::sub].copy(), psimax=float(np.abs(psi).max()),
                absum=float(np.abs(psi).sum() * og.dv), obs=np.array([nrm, mu, en]))


def write(name, steps, sums, sub=SUB, **meta):
    out = dict(steps=np.array(steps), sub=sub, **meta)
    for key in ("psi_s", "psimax", "absum", "obs"):
        out[key] = np.array([s[key] for s in sums])
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **out)
    print("wrote", path, flush=True)


def oracle_grid(w):
    g = w.grid
    kx, ky, kz = g.kaxes()
    return orc.Grid(g.nx, g.ny, g.nz, kx, ky, kz, g.dv)


def config1():
    w = configs.contact_imag_64()
    og = oracle_grid(w)
    ham = orc.Ham(1, w.a)
    psi, V = w.initial_state(), w.trap()
    steps, sums = [], []
    for s in range(0, w.nsteps, 500):
        psi = orc.rk4(og, ham, psi, V, w.dt, 500, t0=s * w.dt, imaginary=True, npart=1.0)
        steps.append(s + 500)
        sums.append(summary(og, ham, psi, V, (s + 500) * w.dt))
        print(s + 500, sums[-1]["obs"], flush=True)
    write("config1_64cubed_2000.npz", steps, sums, dt=w.dt)


def config2(nx, ny, nz, nsteps, every, snapdir=None):
    w = configs.dipolar_real(nx, ny, nz, nsteps=nsteps)
    og = oracle_grid(w)
    ham = orc.Ham(1, w.a, vtilde=orc.vdd(og, w.a_dd, w.dip))
    psi, V = w.initial_state(), w.trap()
    steps, sums = [], []
    for s in range(0, nsteps, every):
        if snapdir:
            psi = np.load(os.path.join(snapdir, f"psi_{s + every}.npy"))
        else:
            psi = orc.rk4(og, ham, psi, V, w.dt, every, t0=s * w.dt)
        steps.append(s + every)
        sums.append(summary(og, ham, psi, V, (s + every) * w.dt, sub=nx // 8))
        print(s + every, sums[-1]["obs"], flush=True)
    write(f"config2_{nx}x{ny}x{nz}_{nsteps}.npz", steps, sums, sub=nx // 8, dt=w.dt)


def main():
    orc.set_threads(int(os.environ.get("OMP_NUM_THREADS", "8")))
    if sys.argv[1] == "config1":
        config1()
    else:
        nx, ny, nz, nsteps, every = map(int, sys.argv[2:7])
        config2(nx, ny, nz, nsteps, every, sys.argv[7] if len(sys.argv) > 7 else None)


if __name__ == "__main__":
    main()
