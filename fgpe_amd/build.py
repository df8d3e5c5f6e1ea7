"""Build recipe of the HIP library libfgpe.so (gfx950) and the CPU oracle.

    python -m fgpe_amd.build          # or __graft_entry__.build()

Every call recompiles.  The library is built in-tree so that it travels with
the source tree; it is git-ignored.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfgpe.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 arithmetic stays off (v_pk_* FP32 forms are avoided)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")

# Library variants: name -> extra flags.  "user_example" compiles the two
# example user pointwise terms of csrc/user_term.hpp (the tests' custom
# Hamiltonians) for the pencil lengths those tests use (FGPE_TEST_SIZES).
VARIANTS = {"": [], "user_example": ["-DFGPE_USER_TERM_EXAMPLE", "-DFGPE_TEST_SIZES"]}


def lengths(variant: str = ""):
    """The pencil lengths a variant compiles (FGPE_FOR_EACH_N in csrc/ctx.hpp)."""
    text = open(os.path.join(CSRC, "ctx.hpp")).read()
    blocks = re.findall(r"#define FGPE_FOR_EACH_N\(X\)([^\n]*)", text)
    small, full = (list(map(int, re.findall(r"X\((\d+)\)", b))) for b in blocks)
    return small if "-DFGPE_TEST_SIZES" in VARIANTS[variant] else full


def units(variant: str = ""):
    """(source, object name, extra flags) of every translation unit of a variant:
    the C ABI, the initialisers, the two dispatch units and one launcher unit
    per precision and pencil length."""
    out = [(os.path.join(CSRC, f), f + ".o", []) for f in
           ("fgpe_capi.hip", "fgpe_init.hip", "launch_f64.hip", "launch_f32.hip")]
    for prec, t in (("f64", "double"), ("f32", "float")):
        for n in lengths(variant):
            out.append((os.path.join(CSRC, "launch_inst.hip"), f"launch_{prec}_{n}.o",
                        [f"-DFGPE_INST_T={t}", f"-DFGPE_INST_N={n}"]))
    return out


def lib_path(variant: str = "") -> str:
    return LIB if not variant else os.path.join(HERE, f"libfgpe_{variant}.so")


def obj_dir(variant: str) -> str:
    return os.path.join(BUILD_DIR, variant) if variant else BUILD_DIR


def _compile(unit, verbose: bool, variant: str = "") -> str:
    src, name, extra = unit
    bdir = obj_dir(variant)
    os.makedirs(bdir, exist_ok=True)
    obj = os.path.join(bdir, name)
    cmd = [HIPCC, *HIP_FLAGS, *VARIANTS[variant], *extra, "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return obj


def link(variant: str, objs, verbose: bool = False) -> str:
    out = lib_path(variant)
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_lib(verbose: bool = False, variants=("", "user_example"), jobs: int = 0) -> str:
    """Compile libfgpe.so and its variants (all units in parallel, the largest first)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    work = [(v, u) for v in variants for u in units(v)]
    # the long lengths take longest: start them first
    work.sort(key=lambda j: -int(re.sub(r"\D", "", j[1][1]) or 0))
    jobs = jobs or min(len(work), os.cpu_count() or 8, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda j: _compile(j[1], verbose, j[0]), work))
    for v in variants:
        link(v, [o for (vv, _), o in zip(work, objs) if vv == v], verbose)
    return LIB


def build_probe(verbose: bool = False) -> str:
    """fgpe_amd/hbm_probe: the achievable HBM bandwidth of the step's access
    patterns (contiguous streams, the FFT passes' tile pattern) on this GPU."""
    out = os.path.join(HERE, "hbm_probe")
    cmd = [HIPCC, *[f for f in HIP_FLAGS if f != "-fPIC"], os.path.join(CSRC, "hbm_probe.hip"), "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], check=True)


def build_host() -> None:
    """The reference-shaped C++ host API, `gpe` and `gpe_params` (g++)."""
    subprocess.run(["make", "-s", "-B", "-j8", "-C", os.path.join(HERE, "host")], check=True)


def main() -> None:
    build_lib(verbose=True)
    build_probe(verbose=True)
    build_host()
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())
