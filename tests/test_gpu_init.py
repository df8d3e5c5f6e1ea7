"""The device initialisers (fgpe_psi_*, fgpe_vext_*) against the oracle's
restatement of src/psi.cu and src/vext.cu (oracle/fgpe_oracle_init.c), and
the `gpe` executable with `device_init true` against its host path.

Bar: the kernels keep the reference's operation order without contraction
and take every factor that depends on one axis from the host math library,
so the traps are bit for bit the oracle's; the states differ only by the
last-ulp differences of the device exp / cos in the Gaussian and the
lattices and the order of the normalisation sum: max |device - oracle| <=
1e-14 max |oracle| (1e-13 after a normalisation and for the lattices' sums
of up to nine cosines)."""
import json
import os
import subprocess

import numpy as np
import pytest

from fgpe_amd import FGPE_PREC_F64, FGPE_PREC_MIXED, Context, Geom, LATTICES
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPE = os.path.join(ROOT, "fgpe_amd", "gpe")

SHAPES = [((32, 32, 64), (6.0, 5.0, 9.0)), ((64, 32, 32), (10.0, 4.0, 4.5))]


def setup(shape, box):
    dx = [2 * b / n for b, n in zip(box, shape)]
    return Geom.of(box, dx), orc.geom(shape, box, dx)


def rel(a, b):
    a, b = np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("prec", [FGPE_PREC_F64, FGPE_PREC_MIXED])
@pytest.mark.parametrize("shape,box", SHAPES)
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_psi_init_and_normalize(shape, box, kind, prec):
    g, og = setup(shape, box)
    ax = (1.3, 0.9, 1.7)
    print(f"psi_init kind {kind} {shape} prec {prec}", flush=True)
    with Context(*shape, precision=prec) as ctx:
        ctx.init_psi(g, kind, ax, rand=(0.05, 0.02), seed=367)
        raw = ctx.download_psi()
        nrm = ctx.psi_normalize(g, 800.0)
        got = ctx.download_psi()
    ref = orc.psi_init(og, kind, ax, rand=(0.05, 0.02), seed=367)
    assert rel(raw, ref) <= 1e-14
    ref_n, ref_nrm = orc.psi_normalize(og, ref, 800.0)
    assert rel(got, ref_n) <= 1e-13
    assert nrm == pytest.approx(ref_nrm, rel=1e-13)
    assert nrm == pytest.approx(800.0, rel=1e-12)


@pytest.mark.parametrize("shape,box", SHAPES)
def test_noise_vortex_and_lattices(shape, box):
    g, og = setup(shape, box)
    ax = (1.0, 1.1, 0.8)
    with Context(*shape) as ctx:
        ctx.init_psi(g, 0, ax)
        ctx.psi_add_noise(amp=(0.1, 0.03), seed=5, stream=2)
        ref = orc.psi_add_noise(og, orc.psi_init(og, 0, ax), (0.1, 0.03), 5, 2)
        assert rel(ctx.download_psi(), ref) <= 1e-14
        ctx.psi_imprint_vortex_tube(g, 3.0)
        ref = orc.psi_vortex(og, ref, 3.0)
        assert rel(ctx.download_psi(), ref) <= 1e-14
        for lat in LATTICES:
            print(f"lattice {lat} {shape}", flush=True)
            off = (0.3, -0.2, 0.45)
            ctx.psi_force_lattice(g, lat, 0.4, 1.9, off, npart=500.0)
            ref = orc.psi_force(og, ref, lat, 0.4, 1.9, off, 500.0)
            assert rel(ctx.download_psi(), ref) <= 1e-13, lat
            ctx.psi_normalize(g, 500.0)
            ref, _ = orc.psi_normalize(og, ref, 500.0)
            assert rel(ctx.download_psi(), ref) <= 1e-13, lat


@pytest.mark.parametrize("prec", [FGPE_PREC_F64, FGPE_PREC_MIXED])
@pytest.mark.parametrize("shape,box", SHAPES)
def test_vext_builders(shape, box, prec):
    g, og = setup(shape, box)
    ax = (1.2, 0.7, 1.05)
    with Context(*shape, precision=prec) as ctx:
        for kind in (0, 1, 2):
            print(f"vext kind {kind} {shape} prec {prec}", flush=True)
            ctx.vext_init(g, kind, ax)
            ctx.vext_add_cos(g, 0.02, 3.0)
            ctx.vext_add_cos(g, -0.01, 1.0)
            got = ctx.vext_commit(download=True).reshape(-1)
            ref = orc.vext_add_cos(og, orc.vext_add_cos(og, orc.vext_init(og, kind, ax), 0.02, 3.0), -0.01, 1.0)
            assert np.all(got.imag == 0)
            # per-axis factors from the host math library, combined in the
            # reference's order without contraction: bit for bit
            assert np.array_equal(got.real, ref)
        ctx.vext_init_box(g, 2.5, 1.3, (2.0, 1.5, 1.8))
        got = ctx.vext_commit(download=True).reshape(-1)
        ref = orc.vext_box(og, 2.5, 1.3, (2.0, 1.5, 1.8))
        assert np.array_equal(got.real, ref)


def test_device_built_state_steps_like_uploaded_one():
    """Device-built psi and V, then RK4 steps, against the oracle's arrays
    uploaded and stepped: the trap enters the Hamiltonian at commit."""
    from fgpe_amd import configs
    shape, box = (32, 32, 64), (6.0, 5.0, 9.0)
    g, og = setup(shape, box)
    w = configs.dipolar_real(*shape, nsteps=4)
    kx, ky, kz = w.grid.kaxes()
    ax = (1.0, 1.0, 1.4)
    outs = []
    for dev in (True, False):
        with Context(*shape) as ctx:
            ctx.set_kspace(kx, ky, kz, w.grid.dv)
            ctx.set_hamiltonian(a=w.a, a_dd=w.a_dd, dip=w.dip)
            if dev:
                ctx.init_psi(g, 0, ax)
                ctx.psi_normalize(g, 300.0)
                ctx.vext_init(g, 0, ax)
                ctx.vext_commit()
            else:
                psi, _ = orc.psi_normalize(og, orc.psi_init(og, 0, ax), 300.0)
                ctx.upload_psi(psi)
                ctx.upload_vext(orc.vext_init(og, 0, ax).astype(np.complex128))
            ctx.rk4_steps(w.dt, 4)
            outs.append(ctx.download_psi())
    assert rel(outs[0], outs[1]) <= 1e-12


INPUT = """
hamiltonian     "bec"
outprefix       "{out}"
inprefix        "{out}"
vprefix         "{out}"
treal           true
read_vext       false
read_psi        false
t2component     false
npart           1000
mass            86.909
tvseed          true
nseed           2
aseed           0.05
a               100.0
omx             100.0
omy             100.0
omz             140.0
xmax            8.0
ymax            8.0
zmax            8.0
nx              32
ny              32
nz              32
dt              0.002
niter           8
epsilon_e       1E-12
itmod           1000000
itStatMod       4
time0           0.0
rand_pct        0.0
crand_pct       0.0
trotate         false
omega           0.0
tpi_vortex      true
api_prc         1.0
beta            0.9
rndseed         7
{extra}
"""


def test_gpe_device_init_matches_host_init(tmp_path):
    res = {}
    for mode in ("host", "device"):
        out = str(tmp_path / f"{mode}_")
        inp = tmp_path / f"{mode}.txt"
        inp.write_text(INPUT.format(out=out, extra="device_init     true" if mode == "device" else ""))
        print(f"gpe run {mode}", flush=True)
        r = subprocess.run([GPE, str(inp)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        if mode == "device":
            assert "(device)" in r.stdout
        res[mode] = {f: np.fromfile(out + f, dtype=np.complex128) for f in ("psi_init.wdat", "vext.wdat")}
        res[mode]["final"] = np.fromfile(out + "psi_final.wdat", dtype=np.complex128)
    for f in ("psi_init.wdat", "vext.wdat", "final"):
        assert rel(res["device"][f], res["host"][f]) <= 1e-13, f
