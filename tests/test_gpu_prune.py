"""Stars::prune on the GPU (prune.cu:28-62): stars outside the solver's box
move behind the others, both groups keeping their order, and the count of
kept stars is returned.  Bit-exact against the oracle on the same arrays,
alone and inside the Stars::update loop (solve, Verlet, prune) where N
shrinks from step to step."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def _check(s, lim, xc, xp, w):
    n = w.shape[0]
    Xc, Xp, W = dev(xc), dev(xp), dev(w)
    kept = s.prune(n, Xc, Xp, W)
    rc, rp, rw, rk = orc.prune(lim, xc, xp, w)
    assert kept == rk
    np.testing.assert_array_equal(Xc.cpu().numpy(), rc)
    np.testing.assert_array_equal(Xp.cpu().numpy(), rp)
    np.testing.assert_array_equal(W.cpu().numpy(), rw)
    return kept


@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 100000, 1 << 20])
def test_prune_matches_oracle(n):
    from nbfmm_amd import Solver
    rng = np.random.default_rng(n)
    lim = np.array([-1.0, -1.5, 2.0, 1.0], np.float32)
    s = Solver(4, n, lim, 1e-3)
    xc = rng.uniform(-1.3, 2.3, (n, 2)).astype(np.float32)
    xp = rng.standard_normal((n, 2)).astype(np.float32)
    w = rng.random(n).astype(np.float32)
    if n > 8:
        xc[0] = (2.0, 1.0)                                            # on x_max, y_max: kept
        xc[1] = (-1.0, -1.5)                                          # on x_min, y_min: kept
        xc[2] = (np.nextafter(np.float32(2.0), np.float32(3.0)), 0)   # just beyond x_max: removed
        xc[3] = (0, np.nextafter(np.float32(-1.5), np.float32(-2.0)))  # just below y_min: removed
        xc[4] = (np.nan, 0.0)                                         # NaN compares false: kept
    _check(s, lim, xc, xp, w)
    s.close()


def test_prune_nothing_and_everything():
    from nbfmm_amd import Solver
    lim = np.array([0, 0, 1, 1], np.float32)
    rng = np.random.default_rng(3)
    n = 5000
    s = Solver(3, n, lim, 1e-3)
    xc = rng.uniform(0.1, 0.9, (n, 2)).astype(np.float32)
    xp, w = xc.copy(), rng.random(n).astype(np.float32)
    assert _check(s, lim, xc, xp, w) == n
    assert _check(s, lim, xc + 2, xp, w) == 0
    s.close()


def test_stars_update_loop_shrinks_n():
    """An expanding disk leaves the box: every step is solve + Verlet + prune
    (Stars::update).  Each prune equals the oracle's on the same post-step
    arrays, and the forces at each N match the oracle solve."""
    from nbfmm_amd import Solver
    rng = np.random.default_rng(8)
    n0 = 20000
    L = 5
    lim = np.array([-1, -1, 1, 1], np.float32)
    r = 0.9 * np.sqrt(rng.random(n0))
    t = 2 * np.pi * rng.random(n0)
    x = np.stack([r * np.cos(t), r * np.sin(t)], 1).astype(np.float32)
    v = 3.0 * x                                                      # radial outflow, faster than escape
    tick = 0.05
    xpre = (x - v * tick).astype(np.float32)
    w = (1.0 / n0 * (0.5 + rng.random(n0))).astype(np.float32)
    s = Solver(L, n0, lim, 1e-3)
    X, P, W = dev(x), dev(xpre), dev(w)
    A = torch.zeros_like(X)
    n = n0
    for step in range(6):
        s.step(n, tick, X, P, W, A)
        torch.cuda.synchronize()
        xc, xp, ww = X[:n].cpu().numpy(), P[:n].cpu().numpy(), W[:n].cpu().numpy()
        kept = s.prune(n, X, P, W)
        rc, rp, rw, rk = orc.prune(lim, xc, xp, ww)
        assert kept == rk, step
        np.testing.assert_array_equal(X[:n].cpu().numpy(), rc)
        np.testing.assert_array_equal(P[:n].cpu().numpy(), rp)
        np.testing.assert_array_equal(W[:n].cpu().numpy(), rw)
        n = kept
        a = torch.zeros((n, 2), device="cuda")
        s.solve(n, X[:n], W[:n], a)
        ref, _ = orc.solve(L, rc[:n], rw[:n], lim, 1e-3)
        got = a.cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5, step
    assert n < n0                                                    # the disk did leave the box
    s.close()
