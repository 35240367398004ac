"""Mutation check of the oracle pins: each pin of tests/test_oracle_pins.py
must fail on a copy of the oracle with one plausible mistake planted (a sign
flip, an off-by-one, a dropped factor, a swapped operand, a wrong exchange
side).  The broken copies are compiled into a temporary directory on the
CPU and loaded through CSRST_ORACLE_LIB in a child process; the committed
oracle is never changed.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "csrst_oracle.cpp")

# (name, text in the oracle, replacement, pin that must fail)
MUTANTS = [
    ("H half step: sign of the dEx/dy term",
     "H.at(1, i, j) = H.at(1, i, j) + sy * (E.at(0, i, j + 1) - E.at(0, i, j));",
     "H.at(1, i, j) = H.at(1, i, j) - sy * (E.at(0, i, j + 1) - E.at(0, i, j));", "check_free_space_modes"),
    ("Ey sweep: H(i-1, j) taken as H(i, j-1)",
     "(s * (((H.at(0, i, j) - H.at(0, i - 1, j)) + H.at(1, i, j)) - H.at(1, i - 1, j)) + dteps * J.at(1, i, j));",
     "(s * (((H.at(0, i, j) - H.at(0, i, j - 1)) + H.at(1, i, j)) - H.at(1, i, j - 1)) + dteps * J.at(1, i, j));",
     "check_free_space_modes"),
    ("Ex sweep: sign of the current",
     " - dteps * J.at(0, i, j));", " + dteps * J.at(0, i, j));", "check_uniform_current"),
    ("cut-cell Hz: lambda and lambdaN swapped",
     "b.lamN[1] * E.at(1, i + 1, j) - b.lam[1] * E.at(1, i, j)) / b.area;",
     "b.lam[1] * E.at(1, i + 1, j) - b.lamN[1] * E.at(1, i, j)) / b.area;", "check_walls_reduce_to_pec"),
    ("cut-cell Hz: area not divided",
     "b.lamN[1] * E.at(1, i + 1, j) - b.lam[1] * E.at(1, i, j)) / b.area;",
     "b.lamN[1] * E.at(1, i + 1, j) - b.lam[1] * E.at(1, i, j));", "check_fractional_wall"),
    ("cut-cell Hz: hy in place of hx",
     "const double etax = dt / (hr[0] * MU_0);", "const double etax = dt / (hr[1] * MU_0);",
     "check_fractional_wall"),
    ("push: external field not halved",
     "const double B = MU_0 * hf + 0.5 * bext;", "const double B = MU_0 * hf + bext;", "check_push_rotation"),
    ("push: rotation operand sign", "um.y + a * -um.x", "um.y + a * um.x", "check_push_rotation"),
    ("push: external field not halved (orbit)",
     "const double B = MU_0 * hf + 0.5 * bext;", "const double B = MU_0 * hf + bext;", "check_bend_orbit"),
    ("exchange: Ex global-to-window ghost row dropped", "(1u << 7) | (1u << 1)", "(1u << 1)",
     "check_two_grid_invariant"),
    ("moveMesh: leading H column refilled one column short",
     "L[3].push_back(rect(l0 + pos[0], l0 + pos[0], f1 + pos[1], l1 + pos[1]));",
     "L[3].push_back(rect(l0 + pos[0] - 1, l0 + pos[0] - 1, f1 + pos[1], l1 + pos[1]));",
     "check_two_grid_invariant"),
]


@pytest.fixture(scope="module")
def mutant_libs():
    src = open(SRC).read()
    tmp = tempfile.mkdtemp(prefix="oracle_mutants_")
    procs, libs = [], []
    for k, (name, old, new, _) in enumerate(MUTANTS):
        assert old in src, name
        path = os.path.join(tmp, "m%d.cpp" % k)
        with open(path, "w") as f:
            f.write(src.replace(old, new))
        so = os.path.join(tmp, "m%d.so" % k)
        cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
               "-I" + os.path.join(ROOT, "oracle"), path, "-o", so]
        procs.append(subprocess.Popen(cmd))
        libs.append(so)
        if len(procs) >= 6:
            assert procs.pop(0).wait() == 0
    for p in procs:
        assert p.wait() == 0
    yield libs
    shutil.rmtree(tmp, ignore_errors=True)


def run_pin(lib, pin):
    env = dict(os.environ, CSRST_ORACLE_LIB=lib)
    code = ("import sys; sys.path.insert(0, %r); sys.path.insert(0, %r); import test_oracle_pins as t; t.%s()"
            % (ROOT, os.path.join(ROOT, "tests"), pin))
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)


def test_pins_pass_on_the_oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    for pin in sorted({m[3] for m in MUTANTS}):
        r = run_pin(lib, pin)
        assert r.returncode == 0, (pin, r.stderr[-2000:])


@pytest.mark.parametrize("k", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pin_catches_mutant(mutant_libs, k):
    name, _, _, pin = MUTANTS[k]
    r = run_pin(mutant_libs[k], pin)
    assert r.returncode != 0, "%s: %s passed on the broken oracle" % (name, pin)
    # a failed check, not a crash
    assert "AssertionError" in r.stderr or "RuntimeError" in r.stderr, r.stderr[-2000:]
