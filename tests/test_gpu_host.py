"""End-to-end runs of the `gpe` executable (the reference-shaped C++ host API
over the C ABI) against the CPU oracle fed with the files the run wrote
(psi_init.wdat, vext.wdat) and the code-unit parameters of its input."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPE = os.path.join(ROOT, "fgpe_amd", "gpe")
GPE_PARAMS = os.path.join(ROOT, "fgpe_amd", "gpe_params")

ONE_COMP = """
hamiltonian     "dbec"
outprefix       "{out}"
inprefix        "{inp}"
vprefix         "{out}"
treal           {treal}
read_vext       false
read_psi        {read}
t2component     false
npart           1000
mass            86.909
tvseed          false
nseed           1
aseed           0.0
a               100.0
add             50.0
dipole_x        0.0
dipole_y        0.0
dipole_z        1.0
omx             100.0
omy             100.0
omz             140.0
xmax            8.0
ymax            8.0
zmax            8.0
nx              32
ny              32
nz              32
dt              {dt}
niter           {niter}
epsilon_e       1E-12
itmod           1000000
itStatMod       {stat}
time0           0.0
rand_pct        0.01
crand_pct       0.0
trotate         false
omega           0.0
tpi_vortex      false
api_prc         0.0
beta            0.9
rndseed         7
"""


def run_gpe(tmp_path, text, name):
    path = tmp_path / name
    path.write_text(text)
    print(f"running gpe {name}", flush=True)
    r = subprocess.run([GPE, str(path)], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    d = subprocess.run([GPE_PARAMS, str(path)], capture_output=True, text=True, check=True)
    return json.loads(d.stdout.strip().splitlines()[-1])


def wdat(path, shape):
    a = np.fromfile(path, dtype=np.complex128)
    return a.reshape((-1,) + shape) if a.size > np.prod(shape) else a.reshape(shape)


def oracle_grid(d):
    nx, ny, nz = d["NX"]
    ks = [orc.kaxis(n, d["DKX"][i], d["KXMAX"][i], n > 1) for i, n in enumerate((nx, ny, nz))]
    return orc.Grid(nx, ny, nz, *ks, d["DV"])


def status_rows(path):
    rows = [l.split() for l in open(path).read().splitlines()[1:] if l.strip()]
    return [[float(v) for v in r] for r in rows]


def test_gpe_dipolar_imaginary_then_real_time(tmp_path):
    im = str(tmp_path / "im_")
    d = run_gpe(tmp_path, ONE_COMP.format(out=im, inp=im, treal="false", read="false", dt=0.01, niter=301,
                                          stat=100), "im.txt")
    g = oracle_grid(d)
    shape = (g.nz, g.ny, g.nx)
    psi0 = wdat(im + "psi_init.wdat", shape)
    V = wdat(im + "vext.wdat", shape)
    h = orc.Ham(1, (d["a"],), vtilde=orc.vdd(g, d["add"], (0.0, 0.0, 1.0)))
    dt = d["dt"] * d["omho"] / 1000
    ref, n, hist = orc.imag_heavy_ball(g, h, psi0, V, dt, 0.9, d["npart"], d["epsilon_e"], d["niter"])
    got = wdat(im + "psi_final.wdat", shape)
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    assert err < 1e-11, err
    rows = status_rows(im + "status.out")
    assert [int(r[0]) for r in rows] == [0, 100, 200, 300]
    for r in rows:
        it = int(r[0])
        assert abs(r[3] - hist[it, 1]) < 1e-11 * max(1.0, abs(hist[it, 1]))   # mu
        assert abs(r[4] - hist[it, 2]) < 1e-11 * max(1.0, abs(hist[it, 2]))   # E

    re = str(tmp_path / "re_")
    d2 = run_gpe(tmp_path, ONE_COMP.format(out=re, inp=im, treal="true", read="true", dt=0.002, niter=30,
                                           stat=10), "re.txt")
    dt2 = d2["dt"] * d2["omho"] / 1000
    ref2 = orc.rk4(g, h, got, V, dt2, 30)
    got2 = wdat(re + "psi_final.wdat", shape)
    err2 = float(np.abs(got2 - ref2).max() / np.abs(ref2).max())
    assert err2 < 1e-12, err2
    rows = status_rows(re + "status.out")
    assert len(rows) == 4   # t = 0 and every 10 steps
    hmu, hen = orc.hpsi(g, h, 30 * dt2, ref2, V)
    nrm = orc.norm(g, ref2)
    mu, en = orc.obs2(g, nrm, ref2, hmu, hen)
    assert abs(rows[-1][2] - nrm) < 0.051                      # printed with %.1f
    assert abs(rows[-1][4] - en) < 1e-11 * abs(en)


TWO_COMP = """
hamiltonian     "2bec"
outprefix       "{out}"
inprefix        "{out}"
vprefix         "{out}"
treal           true
read_vext       false
read_psi        false
t2component     true
npart1          600
npart2          400
mass1           86.909
mass2           86.909
a11             100.0
a22             95.0
a12             80.0
omx1            100.0
omy1            100.0
omz1            100.0
omx2            100.0
omy2            100.0
omz2            100.0
xmax            8.0
ymax            8.0
zmax            8.0
nx              32
ny              32
nz              32
dt              0.002
niter           21
epsilon_e       1E-12
itmod           1000000
itStatMod       10
time0           0.0
rand_pct        0.05
crand_pct       0.02
tvseed          false
trotate         false
omega           0.0
tpi_vortex      false
api_prc         0.0
beta            0.9
rndseed         11
"""


def test_gpe_two_component_real_time(tmp_path):
    out = str(tmp_path / "tc_")
    d = run_gpe(tmp_path, TWO_COMP.format(out=out), "tc.txt")
    g = oracle_grid(d)
    shape = (g.nz, g.ny, g.nx)
    psi = np.stack([wdat(out + "1psi_init.wdat", shape), wdat(out + "2psi_init.wdat", shape)])
    V = np.stack([wdat(out + "1vext.wdat", shape), wdat(out + "2vext.wdat", shape)])
    h = orc.Ham(2, (d["a11"], d["a22"], d["a12"]))
    dt = d["dt"] * d["omho"] / 1000
    # iterations 1 .. niter-1 (src/twoComponentGPSolver.cu:142)
    ref = orc.rk4(g, h, psi, V, dt, d["niter"] - 1)
    got = np.stack([wdat(out + "1psi_final.wdat", shape), wdat(out + "2psi_final.wdat", shape)])
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    assert err < 1e-12, err
