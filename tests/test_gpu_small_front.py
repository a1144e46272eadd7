"""One-block front for small trees (num_level 1..4, N <= 4096): the sort,
P2M and every M2M level in one kernel of one block.  Forces, Verlet states,
the sorted buffers and the multipoles of every level must equal, bit for
bit, the general kernels' (the same solver with the front switched off),
for ragged N, the N limit, particles outside the box (clamped and flagged),
leaves above the P2M wave threshold of 256 particles, and L = 1..4; the
forces also match the oracle's solve."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def pair(num_level, n, limits, eps):
    from nbfmm_amd import Solver
    front = Solver(num_level, n, limits, eps)
    plain = Solver(num_level, n, limits, eps)
    plain.small_front(False)
    return front, plain


def state(s, n):
    out = [s.read_buffer(b, n).cpu() for b in ("position", "weight", "perm", "index")]
    out.append(s.read_buffer("head").cpu())
    out += [s.read_buffer(b).cpu() for b in ("cell_position", "cell_weight", "cell_effect")]
    return out


def assert_same_state(a, b, n):
    for k, (x, y) in enumerate(zip(state(a, n), state(b, n))):
        assert torch.equal(x, y), k


def system(n, num_level, seed, spread=1.0, crowd=0, outside=0):
    """n particles: Gaussian blob (sigma 0.4 * spread) in [-1, 1]^2, `crowd`
    of them packed into one leaf, `outside` of them outside the box."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 0.4 * spread, (n, 2))
    x = np.clip(x, -0.99, 0.99)
    B = 2 << num_level
    if crowd:
        cell = 2.0 / B
        x[:crowd] = -1.0 + cell * (B // 2 + rng.uniform(0.1, 0.9, (crowd, 2)))
    if outside:
        k = rng.choice(n, outside, replace=False)
        x[k] = rng.uniform(1.05, 1.5, (outside, 2)) * rng.choice([-1, 1], (outside, 2))
    w = rng.uniform(0.5, 1.5, n)
    v = rng.normal(0, 0.05, (n, 2))
    return x.astype(np.float32), w.astype(np.float32), v.astype(np.float32)


LIMITS = (-1.0, -1.0, 1.0, 1.0)


@pytest.mark.parametrize("num_level,n,crowd,outside", [
    (4, 4096, 0, 0), (4, 1, 0, 0), (4, 63, 0, 0), (4, 1000, 0, 0), (4, 3000, 700, 0),
    (3, 4000, 300, 13), (2, 777, 0, 5), (1, 100, 0, 0), (1, 4096, 2000, 0),
])
def test_solve_bitwise_equal_general(num_level, n, crowd, outside):
    x, w, _ = system(n, num_level, 7 + n, crowd=crowd, outside=outside)
    front, plain = pair(num_level, n, LIMITS, 1e-3)
    X, W = dev(x), dev(w)
    Af, Ap = torch.zeros_like(X), torch.zeros_like(X)
    front.solve(n, X, W, Af)
    plain.solve(n, X, W, Ap)
    torch.cuda.synchronize()
    assert torch.equal(Af, Ap)
    assert_same_state(front, plain, n)
    assert front.flags() == plain.flags() == (1 if outside else 0)
    if not outside:
        ref, _ = orc.solve(num_level, x, w, np.array(LIMITS, np.float32), 1e-3)
        got = Af.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    front.close()
    plain.close()


def test_steps_and_graph_runs_bitwise_equal_general():
    """A 4096-particle disk: single steps, then graph runs (single-step
    graphs and the 8-step one), against the general kernels with and without
    the warm start."""
    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import disk
    wl = disk(n=4096, steps=30)
    n = wl.n
    front = Solver(wl.num_level, n, wl.limits, wl.eps)
    plain = Solver(wl.num_level, n, wl.limits, wl.eps)
    plain.small_front(False)
    cold = Solver(wl.num_level, n, wl.limits, wl.eps)
    cold.small_front(False)
    cold.warm_start(False)
    X0, P0, W = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    runs = []
    for s in (front, plain, cold):
        X, P = X0.clone(), P0.clone()
        A = torch.zeros_like(X)
        for _ in range(3):
            s.step(n, wl.tick, X, P, W, A)
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            s.run(n, wl.tick, 4, X, P, W, A, use_graph=True)
            s.run(n, wl.tick, 12, X, P, W, A, use_graph=True)
        st.synchronize()
        s.run(n, wl.tick, 11, X, P, W, A, use_graph=False)
        torch.cuda.synchronize()
        runs.append((X, P, A))
    for k in (1, 2):
        for a, b in zip(runs[0], runs[k]):
            assert torch.equal(a, b), k
    assert_same_state(front, plain, n)
    assert front.flags() == 0 and plain.flags() == 0
    # the trajectory follows the oracle's Stars::update loop
    ox, _, _, clamped = orc.run(wl.num_level, np.array(wl.limits, np.float32), wl.eps, wl.tick, 30,
                                wl.position, wl.position_pre, wl.weight)
    assert clamped == 0
    disp = np.abs(ox - wl.position).max()
    assert np.abs(runs[0][0].cpu().numpy() - ox).max() < 1e-3 * disp + 1e-6
    for s in (front, plain, cold):
        s.close()


def test_above_limit_takes_general_path():
    """N = 4097 (one past the front's limit) and num_level 5 (64 x 64
    leaves) run the general kernels: still equal to the front switched off."""
    for num_level, n in ((4, 4097), (5, 3000)):
        x, w, _ = system(n, num_level, 99)
        front, plain = pair(num_level, n, LIMITS, 1e-3)
        X, W = dev(x), dev(w)
        Af, Ap = torch.zeros_like(X), torch.zeros_like(X)
        front.solve(n, X, W, Af)
        plain.solve(n, X, W, Ap)
        torch.cuda.synchronize()
        assert torch.equal(Af, Ap)
        front.close()
        plain.close()


def test_changing_n_and_switching():
    """Calls of different N on one solver, the front switched on and off
    between calls, and a warm-started general sort right after a front sort
    of the same N."""
    from nbfmm_amd import Solver
    x, w, _ = system(30000, 4, 5)
    s = Solver(4, 30000, LIMITS, 1e-3)
    ref = Solver(4, 30000, LIMITS, 1e-3)
    ref.small_front(False)
    ref.warm_start(False)
    X, W = dev(x), dev(w)
    for n, on in ((4000, True), (4000, False), (2345, True), (2345, True), (2345, False), (30000, True),
                  (4000, True)):
        s.small_front(on)
        A, Ar = torch.zeros_like(X), torch.zeros_like(X)
        s.solve(n, X, W, A)
        ref.solve(n, X, W, Ar)
        torch.cuda.synchronize()
        assert torch.equal(A[:n], Ar[:n]), (n, on)
        assert_same_state(s, ref, n)
    s.close()
    ref.close()
