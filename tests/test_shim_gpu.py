"""The C++ front end's shims of ShapeFunction, ScatterGatherImpl and the
caller geometry / stream of the engine (csrst_amd/host/csrst.hpp), run by
the check program tests/shim_check.cpp against the oracle, bit for bit."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shim_check")


def test_cpp_shims_against_oracle(gpu):
    assert os.path.exists(EXE), "tests/shim_check is not built (__graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("all ok")
    assert r.stdout.count(": ok") >= 3 * (2 + 2 + 6 + 3) + 1
