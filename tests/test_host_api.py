"""The reference-shaped C++ host API (fgpe_amd/host): input parsing, unit
conversion and the CLI's exit codes, run on the CPU (no device work)."""
import json
import math
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPE = os.path.join(ROOT, "fgpe_amd", "gpe")
GPE_PARAMS = os.path.join(ROOT, "fgpe_amd", "gpe_params")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")

HBAR, ABOHR, PI = 63507.799, 5.2917721E-5, 3.141592653589793


def dump(path):
    if not os.path.exists(GPE_PARAMS):
        pytest.fail("gpe_params is not built (run __graft_entry__.build())")
    r = subprocess.run([GPE_PARAMS, path], capture_output=True, text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def parse(path):
    """Key/value reading of an input file (first two tokens, '#' comments)."""
    kv = {}
    for line in open(path):
        line = line.split("#", 1)[0].split()
        if len(line) >= 2:
            kv[line[0]] = line[1]
    return kv


@pytest.mark.parametrize("name", ["im_test_input.txt", "re_test_input.txt", "input.txt"])
def test_params_of_reference_inputs(name):
    path = os.path.join(INPUTS, name)
    kv = parse(path)
    d = dump(path)
    om = [2 * PI * float(kv[k]) for k in ("omx", "omy", "omz")]
    m = float(kv["mass"])
    if abs(om[0]) < 1e-6 and abs(om[1]) < 1e-6:
        aho, omho = math.sqrt(HBAR / (m * om[2])), om[2]
    elif abs(om[0]) < 1e-6:
        a1, a2 = math.sqrt(HBAR / (m * om[1])), math.sqrt(HBAR / (m * om[2]))
        aho, omho = math.sqrt(a1 * a2), math.sqrt(om[2] * om[1])
    else:
        ax = [math.sqrt(HBAR / (m * w)) for w in om]
        aho, omho = (ax[0] * ax[1] * ax[2]) ** (1 / 3), (om[0] * om[1] * om[2]) ** (1 / 3)
    assert d["aho"] == pytest.approx(aho, rel=1e-14)
    assert d["omho"] == pytest.approx(omho, rel=1e-14)
    assert d["a"] == pytest.approx(float(kv["a"]) * ABOHR / aho, rel=1e-14)
    nx = [int(kv[k]) for k in ("nx", "ny", "nz")]
    xmax = [float(kv[k]) / aho for k in ("xmax", "ymax", "zmax")]
    dx = [2 * float(kv[k]) / n / aho for k, n in zip(("xmax", "ymax", "zmax"), nx)]
    assert d["NX"] == nx
    for i in range(3):
        assert d["XMAX"][i] == pytest.approx(xmax[i], rel=1e-14)
        assert d["DX"][i] == pytest.approx(dx[i], rel=1e-14)
        assert d["DKX"][i] == pytest.approx(1.0 / (nx[i] * dx[i]), rel=1e-14)
        assert d["KXMAX"][i] == pytest.approx(1.0 / (2 * dx[i]), rel=1e-14)
    assert d["DV"] == pytest.approx(dx[0] * dx[1] * dx[2], rel=1e-14)
    assert d["npart"] == float(kv["npart"])
    assert d["DIM"] == 3


def write_input(tmp_path, text):
    p = tmp_path / "in.txt"
    p.write_text(text)
    return str(p)


def test_missing_typed_key_exits_3(tmp_path):
    src = open(os.path.join(INPUTS, "im_test_input.txt")).read()
    text = "\n".join(l for l in src.splitlines() if not l.startswith("nx"))
    r = subprocess.run([GPE_PARAMS, write_input(tmp_path, text)], capture_output=True, text=True)
    assert r.returncode == 3 and "No such variable nx" in r.stderr


def test_value_typing_rules(tmp_path):
    # "100" is an int and a double; a quoted value is a string; '#' starts a comment
    src = open(os.path.join(INPUTS, "im_test_input.txt")).read()
    text = src.replace("npart           200000", "npart 1e3   # scientific notation").replace(
        "nx              256", "nx 256 trailing tokens ignored")
    d = dump(write_input(tmp_path, text))
    assert d["npart"] == 1000.0 and d["NX"][0] == 256


def test_cli_exit_codes(tmp_path):
    if not os.path.exists(GPE):
        pytest.fail("gpe is not built (run __graft_entry__.build())")
    assert subprocess.run([GPE], capture_output=True).returncode == 1
    assert subprocess.run([GPE, str(tmp_path / "missing.txt")], capture_output=True).returncode == 2
    src = open(os.path.join(INPUTS, "im_test_input.txt")).read()
    r = subprocess.run([GPE, write_input(tmp_path, src.replace('"bec"', '"nosuch"'))], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "No hamiltonian available for: nosuch" in r.stderr


@pytest.mark.parametrize("lattice,fx,fy,fz", [("triangular", 1, 3 ** 0.5, 1), ("HEX", 1, 3 ** 0.5, 1),
                                              ("BCC", 2 ** 0.5, 2 ** 0.5, 2 ** 0.5),
                                              ("SC", 2 ** -0.5, 2 ** -0.5, 2 ** -0.5),
                                              ("FCC", 3 ** 0.5, 3 ** 0.5, 3 ** 0.5), ("stripes", 1, 1, 1)])
def test_soft_lattice_box_redefinition(tmp_path, lattice, fx, fy, fz):
    """Params::RedefineLatticeSoft (src/params.cu:35-131): with hamiltonian
    "soft" each box length becomes the whole number of inter-droplet
    distances d = a_soft / (akrot / 2 pi) (stretched per lattice) nearest to
    it; the box is then converted to code units like any other."""
    src = open(os.path.join(INPUTS, "im_test_input.txt")).read().replace('"bec"', '"soft"')
    src += "\nakrot           3.0\na_soft          1.7\nLatticetype     \"%s\"\n" % lattice
    d = dump(write_input(tmp_path, src))
    kv = parse(os.path.join(INPUTS, "im_test_input.txt"))
    idd = 1.7 / (3.0 / (2 * PI))   # in the input's length unit, like xmax
    for i, (key, f) in enumerate(zip(("xmax", "ymax", "zmax"), (fx, fy, fz))):
        step = idd * f
        L = 2 * float(kv[key])
        n = round(L / step)
        assert abs(n * step - L) <= 0.5 * step + 1e-12
        assert d["XMAX"][i] == pytest.approx(0.5 * n * step / d["aho"], rel=1e-14)
        assert d["DX"][i] == pytest.approx(2 * d["XMAX"][i] / d["NX"][i], rel=1e-14)


def test_plugin_overriding_hamiltonian_does_not_compile(tmp_path):
    """H is the library's kernel configured by hparams, so a reference plugin
    that defines its own alg_calcHpsi / alg_calcHpsiMU would lose its terms:
    the solver classes make such a definition a compile error."""
    host = os.path.join(ROOT, "fgpe_amd", "host")
    for solver, base in (("oneComponentGPSolver", "OneComponentGPSolver"),
                         ("twoComponentGPSolver", "TwoComponentGPSolver")):
        for member in ("alg_calcHpsi", "alg_calcHpsiMU"):
            src = tmp_path / f"{member}.cpp"
            src.write_text(f'#include "hpp/{solver}.hpp"\n'
                           f"class Mine : public {base} {{\n"
                           f"  public:\n    void {member}();\n}};\n")
            r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", host, str(src)],
                               capture_output=True, text=True)
            assert r.returncode != 0 and "final" in r.stderr, (solver, member, r.stderr[-500:])
