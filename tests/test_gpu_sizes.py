"""Grid lengths beyond the power-of-two set and 1D grids (cuFFT plans any
length and 1D grids, src/parallel.cu:67-110): lengths 16, 48, 384 and 1024
on every axis role (x pencils, strided y, last axis), grids with fewer x
rows than a block has pencils, and 1D grids with the 1D soft-core kernel
(src/parallel.cu:371-397), against the oracle.  Also the reference's sample
input at 384 x 64 x 64 through `gpe`."""
import json
import os
import subprocess

import numpy as np
import pytest
import torch  # noqa: F401  (torch's HIP runtime before the library's)

from fgpe_amd import FGPE_IMAG_HEAVY_BALL, FGPE_PREC_MIXED, FgpeError, configs
from oracle import oracle as orc
from tests.cases import Case, rel_max

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (nx, ny, nz): x / strided / last-axis roles of each new length
SHAPES = [(16, 16, 16), (48, 32, 16), (16, 48, 48), (384, 32, 16), (32, 16, 384), (1024, 16, 16),
          (16, 16, 1024), (48, 16, 1), (16, 1024, 1), (1024, 1, 1), (48, 1, 1), (16, 1, 1)]


def stable_dt(c, cap):
    """An explicit step the finest axis allows: RK4 (and the heavy ball) need
    dt times the largest kinetic energy, 19.739208802 |k|^2 at the Nyquist
    k, well under 2.8; long axes in a fixed box have large k."""
    kmax2 = sum(float(np.max(np.abs(k))) ** 2 for k in (c.kx, c.ky, c.kz))
    return min(cap, 0.5 / (19.739208802 * kmax2))


def hamiltonian(shape):
    dim = 3 if shape[2] > 1 else (2 if shape[1] > 1 else 1)
    dip = (0.0, 0.0, 1.0) if dim == 3 else (1.0, 0.0, 0.0)
    return dict(a=(0.3,), a_dd=0.15, dip=dip, sc_U=2.0, sc_R=1.1, sc_dim=dim), dim


def set_all_variants(ctx, v):
    for k, axis in ((0, 0), (1, 1), (1, 2), (2, 1), (2, 2)):
        ctx.set_variant(k, axis, v)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_rk4_new_lengths_against_oracle(shape):
    """12 RK4 steps of the dipolar + soft-core GPE with either plan variant
    (variant 1's wider tiles leave partial tiles on short rows): FP64 within
    1e-12 of the oracle, the mixed form within 1e-6."""
    L = tuple(12.0 if n > 1 else 1.0 for n in shape)
    c = Case(*shape, L=L)
    h, dim = hamiltonian(shape)
    print(f"{shape}: {dim}D", flush=True)
    V = configs.harmonic_trap(c.spec, (1.0, 1.1, 0.9))
    psi = configs.gaussian_noisy(c.spec, 61, width=1.5, noise=0.02)
    vt = c.vtilde(h["a_dd"], h["dip"], sc=(dim, h["sc_R"], h["sc_U"]))
    dt, n = stable_dt(c, 1e-3), 12
    ref = orc.rk4(c.grid, orc.Ham(1, h["a"], vtilde=vt), psi, V, dt, n)
    for prec, tol in ((0, 1e-12), (FGPE_PREC_MIXED, 1e-6)):
        for var in (0, 1):
            with c.ctx(prec) as ctx:
                set_all_variants(ctx, var)
                ctx.set_hamiltonian(**h)
                ctx.upload_psi(psi)
                ctx.upload_vext(V)
                ctx.rk4_steps(dt, n)
                err = rel_max(ctx.download_psi(), ref)
            print(f"  precision {prec} variant {var}: rel err {err:.2e}", flush=True)
            assert err < tol, (shape, prec, var, err)


@pytest.mark.parametrize("shape", [(48, 32, 16), (16, 1024, 1), (1024, 1, 1)], ids=lambda s: "x".join(map(str, s)))
def test_heavy_ball_new_lengths_against_oracle(shape):
    """The reference's imaginary-time scheme on the new lengths and in 1D:
    same state and norm / mu / E history as the oracle."""
    L = tuple(12.0 if n > 1 else 1.0 for n in shape)
    c = Case(*shape, L=L)
    h, dim = hamiltonian(shape)
    V = configs.harmonic_trap(c.spec, (1.0, 1.1, 0.9))
    psi = configs.gaussian_noisy(c.spec, 62, width=1.5, noise=0.02)
    vt = c.vtilde(h["a_dd"], h["dip"], sc=(dim, h["sc_R"], h["sc_U"]))
    dt = stable_dt(c, 2e-3)
    ref, nref, href = orc.imag_heavy_ball(c.grid, orc.Ham(1, h["a"], vtilde=vt), psi, V, dt, 0.9, 1.0, 1e-30, 40)
    with c.ctx() as ctx:
        ctx.set_hamiltonian(**h)
        ctx.upload_psi(psi)
        ctx.upload_vext(V)
        done, hist = ctx.imag_steps(dt, 40, scheme=FGPE_IMAG_HEAVY_BALL, beta=0.9, npart=1.0, eps=1e-30,
                                    history=True)
        got = ctx.download_psi()
    assert done == nref == 40
    assert rel_max(got, ref) < 1e-11
    assert np.abs(hist - href).max() <= 1e-11 * np.abs(href).max()


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fft_new_lengths_against_numpy(torch_cuda, shape):
    """The plain unnormalised 3D transform against numpy.fft.fftn."""
    torch = torch_cuda
    c = Case(*shape)
    x = c.rand_psi(63)
    with c.ctx() as ctx:
        d = torch.from_numpy(x).cuda()
        out = torch.empty_like(d)
        ctx.fft3d_f64(d.data_ptr(), out.data_ptr(), -1)
        ctx.sync()
        got = out.cpu().numpy()
    ref = np.fft.fftn(x)
    assert rel_max(got, ref) < 1e-13


def test_unsupported_length_is_refused():
    from fgpe_amd import Context
    with pytest.raises(FgpeError, match="one of 16, 32, 48"):
        Context(40, 32, 32)
    with pytest.raises(FgpeError, match="ny = 1 must have nz = 1"):
        Context(32, 1, 32)


# The reference's sample input (.ipynb_checkpoints/input-checkpoint.txt:
# grid :35-37, box :31-33, trap :25-27, atoms :9-10,22-23, steps :39,
# noise :47-48, rotation :51-53, vortex :62-63, beta / seed :91-93) with the
# control keys it lacks (hamiltonian, output prefix, ...) and a short run.
SAMPLE = """
hamiltonian     "dbec"
outprefix       "{out}"
inprefix        "{out}"
vprefix         "{out}"
treal           false
read_vext       false
read_psi        false
t2component     false
npart           160000
mass            163.929170
tvseed          false
nseed           1
aseed           0.01
a               94.4
add             132.0
dipole_x        0.0
dipole_y        0.0
dipole_z        1.0
omx             0.0
omy             100.0
omz             100.0
xmax            48.0
ymax            18.0
zmax            20.0
nx              384
ny              64
nz              64
dt              0.0020
niter           20
epsilon_e       1E-34
itmod           1500000
itStatMod       10
time0           0.0
rand_pct        0.010
crand_pct       0.20
trotate         true
omega           2.570
tpi_vortex      true
api_prc         0.0
beta            0.90
rndseed         367
"""


def test_gpe_reference_sample_input_384(tmp_path):
    """`gpe` on the sample input's 384 x 64 x 64 tube: 20 heavy-ball
    iterations (rotating frame, dipolar term, seeded noise) against the
    oracle started from the files the run wrote."""
    out = str(tmp_path / "s_")
    path = tmp_path / "sample.txt"
    path.write_text(SAMPLE.format(out=out))
    r = subprocess.run([os.path.join(ROOT, "fgpe_amd", "gpe"), str(path)], capture_output=True, text=True,
                       cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    d = json.loads(subprocess.run([os.path.join(ROOT, "fgpe_amd", "gpe_params"), str(path)], capture_output=True,
                                  text=True, check=True).stdout.strip().splitlines()[-1])
    nx, ny, nz = d["NX"]
    assert (nx, ny, nz) == (384, 64, 64)
    ks = [orc.kaxis(n, d["DKX"][i], d["KXMAX"][i], n > 1) for i, n in enumerate((nx, ny, nz))]
    g = orc.Grid(nx, ny, nz, *ks, d["DV"])
    shape = (nz, ny, nx)
    psi0 = np.fromfile(out + "psi_init.wdat", dtype=np.complex128).reshape(shape)
    V = np.fromfile(out + "vext.wdat", dtype=np.complex128).reshape(shape)
    vpx = 2 * 3.141592653589793 * 2.570 / (d["omho"] * d["aho"])
    h = orc.Ham(1, (d["a"],), vtilde=orc.vdd(g, d["add"], (0.0, 0.0, 1.0)), vpx=(vpx,))
    orc.set_threads(16)
    dt = d["dt"] * d["omho"] / 1000
    ref, n, hist = orc.imag_heavy_ball(g, h, psi0, V, dt, 0.9, d["npart"], d["epsilon_e"], d["niter"])
    got = np.fromfile(out + "psi_final.wdat", dtype=np.complex128).reshape(shape)
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"sample input 384x64x64: {n} iterations, rel err {err:.2e}", flush=True)
    assert n == 20 and err < 1e-11
