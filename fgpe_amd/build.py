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


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return obj


def build_lib(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
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
