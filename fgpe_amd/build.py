"""Build recipe of the HIP library libfgpe.so (gfx950) and the CPU oracle.

    python -m fgpe_amd.build          # or __graft_entry__.build()

Every call recompiles.  The library is built in-tree so that it travels with
the source tree; it is git-ignored.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
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
SOURCES = [os.path.join(HERE, "csrc", f) for f in
           ("launch_f64.hip", "launch_f32.hip", "fgpe_capi.hip", "fgpe_init.hip")]
BUILD_DIR = os.path.join(HERE, "_build")


# Library variants: name -> extra flags.  "user_example" compiles the example
# user pointwise term of csrc/user_term.hpp (the tests' custom Hamiltonian).
VARIANTS = {"": [], "user_example": ["-DFGPE_USER_TERM_EXAMPLE"]}


def lib_path(variant: str = "") -> str:
    return LIB if not variant else os.path.join(HERE, f"libfgpe_{variant}.so")


def _compile(src: str, verbose: bool, variant: str = "") -> str:
    bdir = os.path.join(BUILD_DIR, variant) if variant else BUILD_DIR
    os.makedirs(bdir, exist_ok=True)
    obj = os.path.join(bdir, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIP_FLAGS, *VARIANTS[variant], "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return obj


def build_lib(verbose: bool = False, variants=("", "user_example")) -> str:
    """Compile libfgpe.so and its variants (all sources in parallel)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = [(v, src) for v in variants for src in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: _compile(j[1], verbose, j[0]), jobs))
    for v in variants:
        out = lib_path(v)
        tmp = out + ".tmp"
        mine = [o for (vv, _), o in zip(jobs, objs) if vv == v]
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *mine, "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
    return LIB


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], check=True)


def build_host() -> None:
    """The reference-shaped C++ host API, `gpe` and `gpe_params` (g++)."""
    subprocess.run(["make", "-s", "-B", "-j8", "-C", os.path.join(HERE, "host")], check=True)


def main() -> None:
    build_lib(verbose=True)
    build_host()
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())
