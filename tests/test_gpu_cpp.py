"""Runs the C++ stage tests (tests/cpp/test_solver.cpp) built against
libnbfmm_amd.so through the C++ nbfmm::Solver mirror."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_solver")


def test_cpp_stage_suite():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_build/test_solver is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
