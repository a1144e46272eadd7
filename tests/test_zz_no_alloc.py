"""nbfmm_solve_noalloc (no reference counterpart): nbfmm_solve without a
solver and without device allocation, all scratch in a static per-GPU array.
Called through the C ABI, it must give nbfmm_solve's bits (same kernels, cold
sort) on two calls in a row, the second smaller (fewer particles, a coarser
grid) so that the array holds the first call's leftovers; the forces also
match the oracle.  Sizes whose scratch exceeds the array are rejected."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

NBFMM_ERR_CAPACITY = 4


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def noalloc(L, x, w, lim, eps, out, stream=None):
    from nbfmm_amd._lib import lib
    st = torch.cuda.current_stream() if stream is None else stream
    arr = (ctypes.c_float * 4)(*[float(v) for v in lim])
    return lib().nbfmm_solve_noalloc(int(L), int(x.shape[0]), arr, float(eps), x.data_ptr(), w.data_ptr(),
                                     out.data_ptr(), int(st.cuda_stream))


def solver_solve(L, x, w, lim, eps):
    from nbfmm_amd import Solver
    n = x.shape[0]
    s = Solver(L, n, lim, eps)
    a = torch.zeros((n, 2), device="cuda")
    s.solve(n, x, w, a)
    torch.cuda.synchronize()
    s.close()
    return a


def test_noalloc_equals_solve_twice_second_smaller():
    from nbfmm_amd.workloads import gaussian_disk_system, disk
    big = gaussian_disk_system(n=65536, num_level=6)      # 128 x 128 leaves
    small = disk(n=5000)                                  # 32 x 32 leaves
    for wl in (big, small):
        x, w = dev(wl.position), dev(wl.weight)
        out = torch.full((wl.n, 2), float("nan"), device="cuda")
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            assert noalloc(wl.num_level, x, w, wl.limits, wl.eps, out, st) == 0
        st.synchronize()
        ref = solver_solve(wl.num_level, x, w, wl.limits, wl.eps)
        assert torch.equal(out, ref), wl.name
        want, _ = orc.solve(wl.num_level, wl.position, wl.weight, np.array(wl.limits, np.float32), wl.eps)
        got = out.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5


def test_noalloc_rejects_oversize():
    from nbfmm_amd._lib import lib
    x = torch.zeros((1 << 20, 2), device="cuda")
    w = torch.ones(1 << 20, device="cuda")
    out = torch.zeros_like(x)
    rc = noalloc(8, x, w, (-1, -1, 1, 1), 1e-3, out)
    assert rc == NBFMM_ERR_CAPACITY
    assert b"static scratch" in lib().nbfmm_last_error()
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0  # nothing ran
