"""Pins of the oracle's initial-state and trap builders (oracle/fgpe_oracle_init.c).

They are checked against
  - the published Philox4x32-10 known-answer vectors (Random123 kat_vectors),
  - the reference-shaped C++ host classes Psi / Vext (fgpe_amd/host, a
    line-by-line mirror of src/psi.cu and src/vext.cu, run through
    `gpe_params --init`), with zero noise, where the two must agree,
  - invariants the formulas fix: normalisation, |psi| under a phase imprint,
    the phase step of the vortex, lattice values at the origin, periods and
    symmetries, mirror symmetry of the cell centres, uniform statistics of
    the noise.
CPU only."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPE_PARAMS = os.path.join(ROOT, "fgpe_amd", "gpe_params")

ONE = """
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
tvseed          false
nseed           3
aseed           0.02
a               100.0
omx             {omx}
omy             120.0
omz             90.0
xmax            9.0
ymax            7.0
zmax            8.0
nx              24
ny              16
nz              20
dt              0.002
niter           1
epsilon_e       1E-12
itmod           1000000
itStatMod       10
time0           0.0
rand_pct        0.0
crand_pct       0.0
trotate         false
omega           0.0
tpi_vortex      false
api_prc         0.0
beta            0.9
rndseed         5
Vbox            3.5
xbox            4.0
ybox            3.0
zbox            2.5
Bbox            1.7
"""

TWO = """
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
mass2           40.0
a11             100.0
a22             95.0
a12             80.0
omx1            {omx}
omy1            110.0
omz1            120.0
omx2            {omx}
omy2            110.0
omz2            120.0
xmax            8.0
ymax            7.0
zmax            6.0
nx              20
ny              16
nz              12
dt              0.002
niter           1
epsilon_e       1E-12
itmod           1000000
itStatMod       10
time0           0.0
rand_pct        0.0
crand_pct       0.0
tvseed          false
trotate         false
omega           0.0
tpi_vortex      false
api_prc         0.0
beta            0.9
rndseed         11
"""


def host_init(tmp_path, text, *extra):
    """Run the host classes; returns (params dict, {name: array})."""
    if not os.path.exists(GPE_PARAMS):
        pytest.fail("gpe_params is not built (run __graft_entry__.build())")
    out = str(tmp_path / "h_")
    inp = tmp_path / "in.txt"
    inp.write_text(text.format(out=out))
    d = json.loads(subprocess.run([GPE_PARAMS, str(inp)], capture_output=True, text=True,
                                  check=True).stdout.strip().splitlines()[-1])
    subprocess.run([GPE_PARAMS, str(inp), "--init", out, *map(str, extra)], capture_output=True, text=True,
                   check=True)
    arrs = {}
    for name in ("psi1", "psi2", "vext1", "vext2"):
        f = out + name + ".bin"
        if os.path.exists(f):
            arrs[name] = np.fromfile(f, dtype=np.complex128)
    return d, arrs


def geom_of(d):
    return orc.geom(d["NX"], d["XMAX"], d["DX"])


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# ---------------------------------------------------------------- Philox ---
# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key -> output
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert orc.philox(ctr, key) == out


def test_noise_word_layout():
    """u1 from words (0, 1), u2 from (2, 3) of block (i_lo, i_hi, stream, 0)."""
    seed, i, stream = 0x0123456789ABCDEF, (7 << 32) + 12345, 3
    o = orc.philox((i & 0xFFFFFFFF, i >> 32, stream, 0), (seed & 0xFFFFFFFF, seed >> 32))
    u1 = (((o[1] << 32) | o[0]) >> 11) / 2.0 ** 52 - 1.0
    u2 = (((o[3] << 32) | o[2]) >> 11) / 2.0 ** 52 - 1.0
    assert orc.noise(i, stream, seed) == (u1, u2)


def test_noise_uniform_statistics():
    g = orc.geom((64, 64, 64), (1, 1, 1), (1, 1, 1))
    p = orc.psi_init(g, 2, rand=(1.0, 1.0), seed=367)
    u1, u2 = p.real - 1.0, p.imag
    n = u1.size
    for u in (u1, u2):
        assert u.min() >= -1.0 and u.max() < 1.0
        assert abs(u.mean()) < 5 * math.sqrt(1 / 3 / n)
        assert abs(u.var() - 1 / 3) < 0.01
        assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 5 / math.sqrt(n)
    assert abs(np.corrcoef(u1, u2)[0, 1]) < 5 / math.sqrt(n)
    q = orc.psi_init(g, 2, rand=(1.0, 1.0), seed=368)
    assert not np.any(q == p)


# ------------------------------------------------- against the host classes ---
@pytest.mark.parametrize("omx,kind", [(100.0, 0), (0.0, 1)])
def test_one_component_psi_and_vext_match_host_classes(tmp_path, omx, kind):
    d, h = host_init(tmp_path, ONE.replace("{omx}", str(omx)))
    g = geom_of(d)
    psi, nrm = orc.psi_normalize(g, orc.psi_init(g, kind, d["ax"]), d["npart"])
    assert rel(psi, h["psi1"]) <= 1e-15
    assert nrm == pytest.approx(d["npart"], rel=1e-13)
    v = orc.vext_init(g, kind, d["ax"])
    assert np.all(h["vext1"].imag == 0)
    assert rel(v, h["vext1"].real) <= 1e-15


@pytest.mark.parametrize("omx,kind", [(100.0, 0), (0.0, 1)])
def test_two_component_psi_and_vext_match_host_classes(tmp_path, omx, kind):
    d, h = host_init(tmp_path, TWO.replace("{omx}", str(omx)))
    g = geom_of(d)
    assert d["ax1"] != d["ax2"]   # the masses differ
    for c in (1, 2):
        ax, npart = d[f"ax{c}"], d[f"npart{c}"]
        psi, _ = orc.psi_normalize(g, orc.psi_init(g, kind, ax), npart)
        assert rel(psi, h[f"psi{c}"]) <= 1e-15, c
        assert rel(orc.vext_init(g, kind, ax), h[f"vext{c}"].real) <= 1e-15, c


@pytest.mark.parametrize("lattice", orc.LATTICES)
def test_lattices_match_host_classes(tmp_path, lattice):
    amp, k, off = 0.3, 1.7, (0.25, -0.4, 0.6)
    d, h = host_init(tmp_path, ONE.replace("{omx}", "100.0"), lattice, amp, k, *off)
    g = geom_of(d)
    psi, _ = orc.psi_normalize(g, orc.psi_init(g, 0, d["ax"]), d["npart"])
    psi = orc.psi_force(g, psi, lattice, amp, k, off, d["npart"])
    psi, _ = orc.psi_normalize(g, psi, d["npart"])
    assert rel(psi, h["psi1"]) <= 1e-14


def test_box_matches_host_class(tmp_path):
    d, h = host_init(tmp_path, ONE.replace("{omx}", "100.0"), "box")
    g = geom_of(d)
    w = [x / d["aho"] for x in (4.0, 3.0, 2.5)]
    v = orc.vext_box(g, 3.5, 1.7, w)
    assert rel(v, h["vext1"].real) <= 1e-15


# --------------------------------------------------------------- invariants ---
G = orc.geom((24, 16, 20), (9.0, 7.0, 8.0), (0.75, 0.875, 0.8))
N3 = (20, 16, 24)   # numpy shape (z, y, x)


def test_cell_centres_mirror_symmetric_and_widths_per_axis():
    ax = (1.3, 0.8, 1.1)
    p = orc.psi_init(G, 0, ax).real.reshape(N3)
    for axis in range(3):
        np.testing.assert_allclose(p, np.flip(p, axis=axis), rtol=1e-14, atol=0)
    # ratio of neighbouring cells along one axis through the centre column
    # fixes the Gaussian width on that axis: exp(-0.05 ((x2/ax)^2 - (x1/ax)^2))
    xs = [-G.xmax[a] + (np.arange(G.n[a]) + 0.5) * G.dx[a] for a in range(3)]
    line = p[N3[0] // 2, N3[1] // 2, :]
    x = xs[0] / ax[0]
    np.testing.assert_allclose(np.log(line[1:] / line[:-1]), -0.05 * (x[1:] ** 2 - x[:-1] ** 2), rtol=1e-9,
                               atol=1e-13)


def test_tube_has_no_x_dependence():
    p = orc.psi_init(G, 1, (0.0, 0.9, 1.2)).reshape(N3)
    assert np.all(p == p[:, :, :1])


def test_normalize_fixes_norm_and_keeps_shape():
    p0 = orc.psi_init(G, 0, (1.0, 1.0, 1.0), rand=(0.1, 0.05), seed=9)
    p, nrm = orc.psi_normalize(G, p0, 1234.5)
    dv = G.dx[0] * G.dx[1] * G.dx[2]
    assert np.sum(np.abs(p) ** 2) * dv == pytest.approx(1234.5, rel=1e-13)
    assert nrm == pytest.approx(1234.5, rel=1e-13)
    r = p / p0
    assert np.allclose(r, r[0], rtol=1e-14, atol=0)


def test_vortex_keeps_modulus_and_steps_phase():
    p0 = orc.psi_init(G, 0, (1.0, 1.0, 1.0))
    p = orc.psi_vortex(G, p0, 2.0)
    np.testing.assert_allclose(np.abs(p), np.abs(p0), rtol=1e-15)
    step = 2.0 * 2 * math.pi / (2 * G.xmax[0]) * G.dx[0]
    ph = np.angle(p.reshape(N3) / p0.reshape(N3))
    d = np.angle(np.exp(1j * (ph[:, :, 1:] - ph[:, :, :-1])))
    np.testing.assert_allclose(d, step, atol=1e-12)
    # two full turns over the box: back to the start one box length on
    assert abs(math.remainder(step * G.n[0], 2 * math.pi)) < 1e-12


@pytest.mark.parametrize("lattice", [l for l in orc.LATTICES if l != "1d"])
def test_lattice_profiles_equal_amp_at_origin(lattice):
    assert orc.force(lattice, 0.0, 0.0, 0.0, 0.7, 2.3) == pytest.approx(0.7, rel=1e-15)


def test_lattice_profile_shapes():
    amp, k = 0.7, 2.3
    assert orc.force("1d", 0.0, 5.0, -3.0, amp, k) == 0.0
    assert orc.force("1d", math.pi / k, 0.0, 0.0, amp, k) == pytest.approx(amp, rel=1e-15)
    rng = np.random.default_rng(4)
    per = 2 * math.pi / k
    for x, y, z in rng.uniform(-3, 3, size=(20, 3)):
        f = orc.force("3d_sc", x, y, z, amp, k)
        for s in ((per, 0, 0), (0, per, 0), (0, 0, per)):
            assert orc.force("3d_sc", x + s[0], y + s[1], z + s[2], amp, k) == pytest.approx(f, abs=1e-12)
        # the triangular profile has the symmetry of the triangle: 120 degree turns
        c, s_ = math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)
        f2 = orc.force("2d_triangular", x, y, 0.0, amp, k)
        assert orc.force("2d_triangular", c * x - s_ * y, s_ * x + c * y, 0.0, amp, k) == pytest.approx(f2, abs=1e-12)
        # square: a 90 degree turn
        f4 = orc.force("2d_square", x, y, z, amp, k)
        assert orc.force("2d_square", -y, x, z, amp, k) == pytest.approx(f4, abs=1e-12)
        # FCC and BCC sums are symmetric under any permutation of the axes
        for lat in ("3d_fcc", "3d_bcc"):
            f3 = orc.force(lat, x, y, z, amp, k)
            assert orc.force(lat, y, z, x, amp, k) == pytest.approx(f3, abs=1e-12)
            assert orc.force(lat, y, x, z, amp, k) == pytest.approx(f3, abs=1e-12)
        # HEX: the in-plane directions are 120 degrees apart, z separate
        fh = orc.force("3d_hex", x, y, z, amp, k)
        assert orc.force("3d_hex", c * x - s_ * y, s_ * x + c * y, z, amp, k) == pytest.approx(fh, abs=1e-12)


def test_vext_harmonic_mirror_symmetric_and_cos_seed_period():
    v = orc.vext_init(G, 0, (1.2, 0.9, 1.1)).reshape(N3)
    for axis in range(3):
        np.testing.assert_allclose(v, np.flip(v, axis=axis), rtol=1e-14, atol=0)
    w = orc.vext_add_cos(G, np.zeros(24 * 16 * 20), 0.3, 2.0).reshape(N3)
    x = -G.xmax[0] + (np.arange(24) + 0.5) * G.dx[0]
    np.testing.assert_allclose(w[3, 5, :], 0.3 * np.cos(2.0 * 2 * np.pi * x / (2 * G.xmax[0])), atol=1e-15)
    # n = 2 over the box length: two periods, values repeat every 12 cells
    np.testing.assert_allclose(w[..., :12], w[..., 12:], atol=1e-15)
