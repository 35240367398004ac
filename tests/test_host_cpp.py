"""The C++ front end (csrst_amd/host): its command line and bunch generator
on the CPU, and the driver binary's error handling without a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "csrst_amd", "csrst")


def test_front_end_unit(tmp_path):
    exe = str(tmp_path / "test_host")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_host.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_driver_rejects_unknown_flag():
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    r = subprocess.run([EXE, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown argument: --bogus" in r.stderr
