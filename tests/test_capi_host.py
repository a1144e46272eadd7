"""CPU checks of the C ABI library: it loads, exports every function that
include/nbfmm_amd.h declares, and rejects bad arguments before touching a
GPU.  No compute calls are made here."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nbfmm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbfmm_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from nbfmm_amd import _lib
    if not os.path.exists(_lib.lib_path()):
        pytest.fail("libnbfmm_amd.so is not built (python -m nbfmm_amd.build)")
    return _lib.lib()


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_binding_table_matches_header():
    from nbfmm_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_version(lib):
    assert lib.nbfmm_version() == 10200


def _create(lib, L, n, lim, eps):
    arr = (ctypes.c_float * 4)(*lim)
    h = ctypes.c_void_p()
    rc = lib.nbfmm_solver_create(L, n, arr, eps, ctypes.byref(h))
    return rc, h


@pytest.mark.parametrize("L,n,lim,eps", [
    (9, 100, (0, 0, 1, 1), 0.1),      # level beyond 8 (base_dim 1024, solver.cu:25)
    (-1, 100, (0, 0, 1, 1), 0.1),
    (3, 0, (0, 0, 1, 1), 0.1),        # no capacity
    (3, 100, (1, 0, 0, 1), 0.1),      # x_min >= x_max (solver.cu:27)
    (3, 100, (0, 1, 1, 1), 0.1),
    (3, 100, (0, 0, 1, 1), 0.0),      # epsilon must be > 0
    (3, 100, (0, 0, 1, 1), float("nan")),
])
def test_create_rejects_bad_arguments(lib, L, n, lim, eps):
    rc, h = _create(lib, L, n, lim, eps)
    assert rc == 1
    assert not h.value
    assert lib.nbfmm_last_error()


def test_null_arguments_rejected(lib):
    assert lib.nbfmm_solver_destroy(None) == 1
    assert lib.nbfmm_solve(None, 10, None, None, None, None) == 1
    assert lib.nbfmm_update(-1, 0.1, None, None, None, None) == 1
    assert lib.nbfmm_update(5, 0.1, None, None, None, None) == 1
    assert lib.nbfmm_solver_create(3, 10, None, 0.1, None) == 1
    kept = ctypes.c_int(7)
    assert lib.nbfmm_prune(None, 10, None, None, None, ctypes.byref(kept), None) == 1


def test_video_arguments_rejected(lib, tmp_path):
    lim = (ctypes.c_float * 4)(0, 0, 1, 1)
    bad_lim = (ctypes.c_float * 4)(1, 0, 0, 1)
    frame = ctypes.c_void_p(16)  # never dereferenced: every call below fails its argument checks
    assert lib.nbfmm_display(1, 15, 10, 1.0, lim, frame, frame, frame, None) == 1    # odd width
    assert lib.nbfmm_display(1, 16, 0, 1.0, lim, frame, frame, frame, None) == 1
    assert lib.nbfmm_display(1, 16, 10, 0.0, lim, frame, frame, frame, None) == 1    # size_scale
    assert lib.nbfmm_display(1, 16, 10, 1.0, bad_lim, frame, frame, frame, None) == 1
    assert lib.nbfmm_display(1, 16, 10, 1.0, lim, None, frame, frame, None) == 1
    assert lib.nbfmm_display(0, 16, 10, 1.0, lim, None, None, None, None) == 1      # no frame
    assert b"nbfmm_display" in lib.nbfmm_last_error()
    h = ctypes.c_void_p()
    path = str(tmp_path / "x.y4m").encode()
    assert lib.nbfmm_y4m_open(path, 15, 10, 30, ctypes.byref(h)) == 1
    assert lib.nbfmm_y4m_open(path, 16, 10, 0, ctypes.byref(h)) == 1
    assert lib.nbfmm_y4m_open(None, 16, 10, 30, ctypes.byref(h)) == 1
    assert lib.nbfmm_y4m_open(path, 16, 10, 30, None) == 1
    assert not h.value
    assert lib.nbfmm_y4m_write_frame(None, frame, None) == 1
    assert lib.nbfmm_y4m_close(None, None) == 1


def test_main_usage_without_arguments():
    import subprocess
    from nbfmm_amd import build
    if not os.path.exists(build.MAIN):
        pytest.fail("nbfmm_main is not built (python -m nbfmm_amd.build)")
    r = subprocess.run([build.MAIN, "out.y4m", "64"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Usage:" in r.stdout and "<model> [model parameters]" in r.stdout
