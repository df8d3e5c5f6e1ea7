"""The custom-Hamiltonian path: the library built with the example user
pointwise term (libfgpe_user_example.so) against the oracle's restatement of
that term, in a child process (tests/user_term_checks.py) because a process
loads one build of the library; and the shipped library reports none."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VARIANT = os.path.join(ROOT, "fgpe_amd", "libfgpe_user_example.so")


def test_shipped_library_has_no_user_term():
    from fgpe_amd._lib import user_term_enabled
    assert not user_term_enabled()


def test_user_term_build_matches_oracle():
    if not os.path.exists(VARIANT):
        pytest.fail("libfgpe_user_example.so is not built (run __graft_entry__.build())")
    env = dict(os.environ, FGPE_LIBRARY=VARIANT)
    print("running tests/user_term_checks.py with the user-term build", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "user_term_checks.py")], env=env,
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    print(r.stdout[-4000:], r.stderr[-4000:])
    assert r.returncode == 0
