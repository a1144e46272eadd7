"""The benchmark configurations at their own sizes.

  * two-disk collision, 2^20 particles, 512 x 512 leaves: the timed path
    (a graph replay in which each step's P2P also makes the next step's first
    warm-sort pass) equals, bit for bit, a solver that sorts from scratch
    every step (the stable sort of source/nbfmm/core/solver/predo.cu:122-130
    each time), and its positions follow the oracle's Stars::update loop;
    also at a 10^4 times larger G, where particles change leaf 100 times as
    often;
  * 256 independent disks (N log-uniform in [1024, 131072]): graph-replayed
    batched steps (runs of 3 and 11 steps: single-step graphs and the 8-step
    one) equal each system's own Solver bit for bit, and ten
    sampled systems match the oracle's solve;
  * 64 two-disk collisions (N log-uniform in [4096, 262144]), 100 steps
    checkpointed every 10, observations at steps 25/50/75/100: the batched
    backward's mass gradients equal each system's own trajectory gradients
    bit for bit; translation invariance and degree-1 mass homogeneity of the
    batched solve adjoint, system by system; and a few gradient steps on the
    masses reduce the trajectory loss at the configuration's own tick.
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("g_scale", [1.0, 1e4])
def test_two_disk_collision_graph_chain_equals_cold_sorts(g_scale):
    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import two_disk_collision
    n = 1 << 20
    wl = two_disk_collision(n=n, steps=16, G=g_scale / n)
    assert wl.num_level == 8 and wl.n == n
    warm = Solver(wl.num_level, n, wl.limits, wl.eps)
    cold = Solver(wl.num_level, n, wl.limits, wl.eps)
    cold.warm_start(False)
    Xw, Pw, W = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    Xc, Pc = Xw.clone(), Pw.clone()
    Aw, Ac = torch.zeros_like(Xw), torch.zeros_like(Xw)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        warm.run(n, wl.tick, 4, Xw, Pw, W, Aw, use_graph=True)  # eager step, then chained replays
        warm.run(n, wl.tick, 12, Xw, Pw, W, Aw, use_graph=True)  # the same graphs again, and the 8-step one
        for _ in range(16):
            cold.step(n, wl.tick, Xc, Pc, W, Ac)
    st.synchronize()
    assert torch.equal(Aw, Ac) and torch.equal(Xw, Xc) and torch.equal(Pw, Pc)
    for b in ("perm", "index", "position"):
        assert torch.equal(warm.read_buffer(b, n), cold.read_buffer(b, n)), b
    assert torch.equal(warm.read_buffer("head"), cold.read_buffer("head"))
    assert warm.flags() == 0 and cold.flags() == 0
    lim = np.array(wl.limits, np.float32)
    ox, op, oa, clamped = orc.run(wl.num_level, lim, wl.eps, wl.tick, 16, wl.position, wl.position_pre, wl.weight)
    assert clamped == 0
    xg = Xw.cpu().numpy()
    disp = np.abs(ox - wl.position).max()
    assert np.abs(xg - ox).max() < 1e-3 * disp + 1e-6
    assert rel_l2(Aw.cpu().numpy(), oa) < 1e-5
    warm.close()
    cold.close()


def test_many_disks_batch_equals_single_solvers():
    from nbfmm_amd import Batch, Solver
    from nbfmm_amd.workloads import many_disks
    bw = many_disks()
    assert len(bw.systems) == 256
    x, p, w = bw.concat()
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=[s.tick for s in bw.systems])
    X, P, W = dev(x), dev(p), dev(w)
    A = torch.zeros_like(X)
    A0 = torch.zeros_like(X)
    b.solve(X, W, A0)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        b.run(3, X, P, W, A, use_graph=True)
        b.run(11, X, P, W, A, use_graph=True)  # with the 8-step graph
    st.synchronize()
    assert not b.flags().any()
    off = b.offsets
    for k, sy in enumerate(bw.systems):
        s = Solver(sy.num_level, sy.n, sy.limits, sy.eps)
        xs, ps, ws = dev(sy.position), dev(sy.position_pre), dev(sy.weight)
        a = torch.zeros_like(xs)
        for _ in range(14):
            s.step(sy.n, sy.tick, xs, ps, ws, a)
        torch.cuda.synchronize()
        sl = slice(off[k], off[k + 1])
        assert torch.equal(X[sl], xs) and torch.equal(P[sl], ps) and torch.equal(A[sl], a), k
        s.close()
    rng = np.random.default_rng(44)
    for k in rng.choice(len(bw.systems), 10, replace=False):
        sy = bw.systems[k]
        ref, _ = orc.solve(sy.num_level, sy.position, sy.weight, np.array(sy.limits, np.float32), sy.eps)
        got = A0[off[k]:off[k + 1]].cpu().numpy()
        assert rel_l2(got, ref) < 1e-5, k
    b.close()


@pytest.fixture(scope="module")
def collisions():
    from nbfmm_amd.workloads import many_collisions
    bw = many_collisions()
    assert len(bw.systems) == 64 and bw.steps == 100
    return bw


def _batch(bw):
    from nbfmm_amd import Batch
    return Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
                 [s.eps for s in bw.systems], ticks=[s.tick for s in bw.systems])


OBS = [25, 50, 75, 100]


def test_many_collisions_mass_gradients_equal_single_trajectories(collisions):
    from nbfmm_amd import Solver
    from nbfmm_amd.autograd import fmm_batch_trajectory, fmm_trajectory
    bw = collisions
    x, p, w = bw.concat()
    b = _batch(bw)
    X, P, W = dev(x), dev(p), dev(w)
    rng = np.random.default_rng(41)
    m0 = dev(w * rng.uniform(0.8, 1.2, w.shape[0]).astype(np.float32))
    with torch.no_grad():
        target = fmm_batch_trajectory(b, bw.steps, 10, OBS, X, P, W)
    m = m0.clone().requires_grad_(True)
    out = fmm_batch_trajectory(b, bw.steps, 10, OBS, X, P, m)
    ((out - target) ** 2).sum().backward()
    torch.cuda.synchronize()
    gm = m.grad
    assert bool(torch.isfinite(gm).all()) and float(gm.abs().max()) > 0
    assert not b.flags().any()
    off = b.offsets
    for k, sy in enumerate(bw.systems):
        sl = slice(off[k], off[k + 1])
        s = Solver(sy.num_level, sy.n, sy.limits, sy.eps)
        mk = m0[sl].clone().requires_grad_(True)
        ok = fmm_trajectory(s, sy.tick, bw.steps, 10, OBS, X[sl], P[sl], mk)
        assert torch.equal(ok.detach(), out.detach()[:, sl]), k
        ((ok - target[:, sl]) ** 2).sum().backward()
        torch.cuda.synchronize()
        assert torch.equal(mk.grad, gm[sl]), k
        s.close()
    b.close()


def test_many_collisions_adjoint_identities(collisions):
    """Batched solve adjoint at the initial state, system by system:
    sum_i dloss/dx_i = 0 (forces depend on differences only, cells held
    fixed) and sum_j m_j dloss/dm_j = sum_i g_i . a_i (forces are linear in
    the masses)."""
    bw = collisions
    x, _, w = bw.concat()
    b = _batch(bw)
    rng = np.random.default_rng(5)
    g = rng.normal(0, 1, x.shape).astype(np.float32)
    X, W, G = dev(x), dev(w), dev(g)
    A, GX, GM = torch.zeros_like(X), torch.zeros_like(X), torch.zeros_like(W)
    b.solve(X, W, A)
    b.solve_adjoint(X, W, G, GX, GM)
    torch.cuda.synchronize()
    a, gx, gm = (t.double().cpu().numpy() for t in (A, GX, GM))
    off = b.offsets
    for k in range(len(bw.systems)):
        sl = slice(off[k], off[k + 1])
        assert np.all(np.abs(gx[sl].sum(0)) < 1e-4 * np.abs(gx[sl]).sum(0)), k
        lhs = float((w[sl].astype(np.float64) * gm[sl]).sum())
        rhs = float((g[sl].astype(np.float64) * a[sl]).sum())
        assert abs(lhs - rhs) < 1e-4 * float(np.abs(w[sl] * gm[sl]).sum()), k
    b.close()


def test_many_collisions_mass_fit_reduces_loss(collisions):
    """Mass inference at the configuration's own tick: starting from
    true x U[0.8, 1.2], three gradient steps on the masses of every system
    reduce the position mismatch at steps 25/50/75/100 (their first step
    already)."""
    from nbfmm_amd.autograd import fmm_batch_trajectory
    bw = collisions
    x, p, w = bw.concat()
    b = _batch(bw)
    X, P, W = dev(x), dev(p), dev(w)
    rng = np.random.default_rng(41)
    with torch.no_grad():
        target = fmm_batch_trajectory(b, bw.steps, 10, OBS, X, P, W)
    m = dev(w * rng.uniform(0.8, 1.2, w.shape[0]).astype(np.float32)).requires_grad_(True)
    losses = []
    for _ in range(4):
        out = fmm_batch_trajectory(b, bw.steps, 10, OBS, X, P, m)
        loss = ((out - target) ** 2).sum(-1).mean()
        losses.append(float(loss.detach()))
        m.grad = None
        loss.backward()
        with torch.no_grad():
            # steepest descent with the step that the loss's local quadratic
            # model along the gradient would take at half length
            gn = float((m.grad ** 2).sum())
            m -= 0.5 * losses[-1] / max(gn, 1e-30) * m.grad
    b.close()
    assert losses[0] > 0
    assert losses[1] < losses[0] and losses[-1] < losses[0], losses
