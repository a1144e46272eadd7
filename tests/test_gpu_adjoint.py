"""Adjoint of force evaluation and of the Verlet trajectory (no reference
counterpart; the reference only evaluates forces).

Pins:
  * central finite differences of the CPU oracle's fp64 solve / fp64 Verlet
    loop (same algorithm) with every particle kept inside its cell, to 1e-3
    relative (the GPU adjoint runs in fp32);
  * the exact regimes where the FMM equals the direct sum (num_level 0, and at
    most one particle per cell on every level): torch autograd of an fp64
    direct sum;
  * size-independent identities at 65536 particles: translation invariance
    (sum_i dloss/dx_i = 0) and degree-1 homogeneity in the masses
    (sum_j m_j dloss/dm_j = sum_i g_i . a_i);
  * massless particles (the plain pair) beside massive ones (the pair with
    the target's gradient divided by its mass), against the same differences;
  * the mass-inference configuration's own check: an N = 256 subset of its
    disk, h = 1e-3 central differences of the fp64 oracle's trajectory loss
    against the fp32 GPU gradients (positions, velocities, masses);
  * checkpointing: K = 1, K = 4 and K = steps give the same bits;
  * batches: system by system the same bits as separate solvers.
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def inside_cells(rng, n, L, lim, margin=0.2):
    """n points in distinct random leaves, each at least `margin` of a leaf
    width from the leaf's edges."""
    B = 1 << (L + 1)
    cw = (lim[2] - lim[0]) / B
    ch = (lim[3] - lim[1]) / B
    cells = rng.choice(B * B, size=n, replace=n > B * B)
    ix, iy = cells % B, cells // B
    x = lim[0] + (ix + rng.uniform(margin, 1 - margin, n)) * cw
    y = lim[1] + (iy + rng.uniform(margin, 1 - margin, n)) * ch
    return np.stack([x, y], 1)


def gpu_adjoint(L, pos, w, lim, eps, g):
    from nbfmm_amd import Solver
    n = pos.shape[0]
    s = Solver(L, n, lim, eps)
    gx = torch.zeros((n, 2), device="cuda")
    gm = torch.zeros(n, device="cuda")
    s.solve_adjoint(n, dev(pos), dev(w), dev(g), gx, gm)
    torch.cuda.synchronize()
    out = gx.cpu().numpy(), gm.cpu().numpy()
    s.close()
    return out


def fd_solve(L, pos, w, lim, eps, g, h=1e-5):
    """Central differences of sum_i g_i . a_i with the fp64 oracle."""
    pos = pos.astype(np.float64)
    w = w.astype(np.float64)
    lim64 = np.asarray(lim, np.float64)

    def loss(p, m):
        return float((g * orc.solve_f64(L, p, m, lim64, eps)).sum())

    gx = np.zeros_like(pos)
    for i in range(pos.shape[0]):
        for d in range(2):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += h
            pm[i, d] -= h
            gx[i, d] = (loss(pp, w) - loss(pm, w)) / (2 * h)
    gm = np.zeros_like(w)
    for i in range(w.shape[0]):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        gm[i] = (loss(pos, wp) - loss(pos, wm)) / (2 * h)
    return gx, gm


@pytest.mark.parametrize("L", [1, 3, 4])
def test_solve_adjoint_matches_fd(L):
    rng = np.random.default_rng(300 + L)
    lim = (-1.0, -1.0, 1.0, 1.0)
    n = 256
    pos = inside_cells(rng, n, L, lim).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    g = rng.normal(0, 1, (n, 2))
    eps = 1e-3
    gx, gm = gpu_adjoint(L, pos, w, lim, eps, g)
    fx, fm = fd_solve(L, pos, w, lim, eps, g)
    assert rel(gx, fx) < 1e-3, rel(gx, fx)
    assert rel(gm, fm) < 1e-3, rel(gm, fm)


def test_solve_adjoint_zero_masses_matches_fd():
    """Waves whose targets all have nonzero mass take the pair with the
    target's gradient divided by its mass; a wave with a massless target
    takes the plain pair.  Both against central differences."""
    L = 3
    rng = np.random.default_rng(77)
    lim = (-1.0, -1.0, 1.0, 1.0)
    pos = inside_cells(rng, 256, L, lim)
    w = rng.uniform(0.5, 1.5, 256)
    # massless particles beside massive ones in leaves of the top rows only
    # (a leaf of massless particles alone has no centre of mass)
    top = rng.choice(np.flatnonzero(pos[:, 1] > 0.6), 8, replace=False)
    cw = (lim[2] - lim[0]) / 2 ** (L + 1)
    pos = np.concatenate([pos, pos[top] + 0.05 * cw]).astype(np.float32)
    w = np.concatenate([w, np.zeros(8)]).astype(np.float32)
    n = pos.shape[0]
    g = rng.normal(0, 1, (n, 2))
    eps = 1e-3
    gx, gm = gpu_adjoint(L, pos, w, lim, eps, g)
    fx, fm = fd_solve(L, pos, w, lim, eps, g)
    assert rel(gx, fx) < 1e-3, rel(gx, fx)
    assert rel(gm, fm) < 1e-3, rel(gm, fm)


def direct_torch(pos, w, eps):
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d * d).sum(-1)
    eye = torch.eye(pos.shape[0], dtype=pos.dtype)
    r = torch.sqrt(r2 + eye) * (1 - eye)
    f = w[None, :] / (r ** 3 + eps) * (1 - eye)
    return (d * f[..., None]).sum(1)


@pytest.mark.parametrize("L,n", [(0, 300), (2, 16), (5, 16)])
def test_adjoint_exact_regime_matches_direct_autograd(L, n):
    """num_level 0 is the direct sum; with one particle per top-level cell
    every level has at most one particle per cell and the FMM equals the
    direct sum, so its gradient is the direct sum's."""
    rng = np.random.default_rng(400 + L)
    lim = (-1.0, -1.0, 1.0, 1.0)
    if L == 0:
        pos = rng.uniform(-0.95, 0.95, (n, 2))
    else:
        top = 2.0 / 4
        pos = np.array([[-1 + (i + rng.uniform(0.2, 0.8)) * top, -1 + (j + rng.uniform(0.2, 0.8)) * top]
                        for i in range(4) for j in range(4)])
    pos = pos.astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    g = rng.normal(0, 1, (n, 2))
    eps = 1e-3
    gx, gm = gpu_adjoint(L, pos, w, lim, eps, g)
    P = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    W = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    (direct_torch(P, W, eps) * torch.tensor(g)).sum().backward()
    assert rel(gx, P.grad.numpy()) < 1e-4, rel(gx, P.grad.numpy())
    assert rel(gm, W.grad.numpy()) < 1e-4, rel(gm, W.grad.numpy())


def test_adjoint_identities_full_size():
    """65536 particles, 128x128 leaves: translation invariance and mass
    homogeneity of the FMM hold for its gradient."""
    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import gaussian_disk_system
    wl = gaussian_disk_system()
    n = wl.n
    rng = np.random.default_rng(5)
    g = rng.normal(0, 1, (n, 2)).astype(np.float32)
    s = Solver(wl.num_level, n, wl.limits, wl.eps)
    x, w, G = dev(wl.position), dev(wl.weight), dev(g)
    a = torch.zeros_like(x)
    gx = torch.zeros_like(x)
    gm = torch.zeros_like(w)
    s.solve(n, x, w, a)
    s.solve_adjoint(n, x, w, G, gx, gm)
    torch.cuda.synchronize()
    gx, gm, a = gx.double().cpu().numpy(), gm.double().cpu().numpy(), a.double().cpu().numpy()
    assert np.all(np.abs(gx.sum(0)) < 1e-4 * np.abs(gx).sum(0))
    lhs = float((wl.weight.astype(np.float64) * gm).sum())
    rhs = float((g.astype(np.float64) * a).sum())
    assert abs(lhs - rhs) < 1e-4 * (np.abs(wl.weight * gm).sum())
    s.close()


def _trajectory_case(seed=11):
    """Small system whose fp64 trajectory keeps every particle at least 5% of
    a leaf away from leaf edges at every step (finite differences valid)."""
    L, n, steps = 2, 96, 12
    lim = (-1.0, -1.0, 1.0, 1.0)
    B = 1 << (L + 1)
    for s in range(seed, seed + 50):
        rng = np.random.default_rng(s)
        pos = inside_cells(rng, n, L, lim, margin=0.3)
        vel = rng.normal(0, 0.5, (n, 2))
        w = rng.uniform(0.5, 1.5, n) * 1e-2
        tick = 2e-3
        pre = pos - vel * tick
        ok = True
        x, p = pos.copy(), pre.copy()
        for _ in range(steps + 1):
            fx = (x[:, 0] - lim[0]) / ((lim[2] - lim[0]) / B)
            fy = (x[:, 1] - lim[1]) / ((lim[3] - lim[1]) / B)
            frac = np.minimum(np.minimum(fx % 1, 1 - fx % 1), np.minimum(fy % 1, 1 - fy % 1))
            if frac.min() < 0.05 or np.abs(x).max() > 0.99:
                ok = False
                break
            x, p = orc.run_f64(L, lim, 1e-3, tick, 1, x, p, w)
        if ok:
            return L, lim, pos.astype(np.float32), pre.astype(np.float32), w.astype(np.float32), tick, steps
    pytest.skip("no trajectory without cell crossings found")


def test_trajectory_adjoint_matches_fd_and_checkpoint_invariance():
    from nbfmm_amd import Solver
    from nbfmm_amd.autograd import fmm_trajectory
    L, lim, pos, pre, w, tick, steps = _trajectory_case()
    n = pos.shape[0]
    obs = [6, steps]
    rng = np.random.default_rng(3)
    c = rng.normal(0, 1, (len(obs), n, 2))
    s = Solver(L, n, lim, 1e-3)
    grads = {}
    for K in (1, 4, steps):
        x0 = dev(pos).requires_grad_(True)
        p0 = dev(pre).requires_grad_(True)
        m = dev(w).requires_grad_(True)
        out = fmm_trajectory(s, tick, steps, K, obs, x0, p0, m)
        (out * dev(c)).sum().backward()
        grads[K] = (x0.grad.clone(), p0.grad.clone(), m.grad.clone())
    for K in (4, steps):
        for a, b in zip(grads[1], grads[K]):
            assert torch.equal(a, b), K
    gx, gp, gm = (t.cpu().numpy() for t in grads[4])

    def loss(x, p, m):
        tot = 0.0
        prev = 0
        xc, pc = x.astype(np.float64), p.astype(np.float64)
        for o, k in enumerate(obs):
            xc, pc = orc.run_f64(L, lim, 1e-3, tick, k - prev, xc, pc, m.astype(np.float64))
            prev = k
            tot += float((c[o] * xc).sum())
        return tot

    h = 1e-6
    idx = rng.choice(n, 12, replace=False)
    for i in idx:
        for d in range(2):
            for arr, grad in ((pos, gx), (pre, gp)):
                a_p = arr.astype(np.float64).copy()
                a_m = a_p.copy()
                a_p[i, d] += h
                a_m[i, d] -= h
                if arr is pos:
                    fd = (loss(a_p, pre, w) - loss(a_m, pre, w)) / (2 * h)
                else:
                    fd = (loss(pos, a_p, w) - loss(pos, a_m, w)) / (2 * h)
                assert abs(grad[i, d] - fd) <= 1e-3 * max(abs(fd), 1.0), (i, d, grad[i, d], fd)
        wp = w.astype(np.float64).copy()
        wm = wp.copy()
        wp[i] += h
        wm[i] -= h
        fd = (loss(pos, pre, wp) - loss(pos, pre, wm)) / (2 * h)
        assert abs(gm[i] - fd) <= 1e-3 * max(abs(fd), 1.0), (i, gm[i], fd)
    s.close()


def test_trajectory_checkpoints_chained_recompute():
    """A system above the one-block front's size, whose recomputed steps run
    chained as in a run (each step's P2P making the next sort's first pass)
    with every step's sort kept for its adjoint: K = 1 (nothing recomputed),
    3 and steps give the same bits, with particles changing leaves."""
    from nbfmm_amd import Solver
    from nbfmm_amd.autograd import fmm_trajectory
    L, n, steps, tick = 5, 20000, 9, 2e-3
    lim = (-1.0, -1.0, 1.0, 1.0)
    rng = np.random.default_rng(41)
    pos = rng.uniform(-0.9, 0.9, (n, 2))
    pre = pos - rng.normal(0, 1.0, (n, 2)) * tick
    w = rng.uniform(0.5, 1.5, n) / n
    B = 1 << (L + 1)
    leaf = lambda x: (np.floor((x[:, 0] + 1) * B / 2) + B * np.floor((x[:, 1] + 1) * B / 2)).astype(np.int64)
    assert (leaf(pos) != leaf(2 * pos - pre)).sum() > 100  # movers from the first step on
    obs = [4, steps]
    c = dev(rng.normal(0, 1, (len(obs), n, 2)))
    s = Solver(L, n, lim, 1e-3)
    grads = {}
    for K in (1, 3, steps):
        x0, p0, m = (dev(a).requires_grad_(True) for a in (pos, pre, w))
        out = fmm_trajectory(s, tick, steps, K, obs, x0, p0, m)
        (out * c).sum().backward()
        torch.cuda.synchronize()
        grads[K] = (x0.grad.clone(), p0.grad.clone(), m.grad.clone())
    assert int(s.flags()) == 0
    for K in (3, steps):
        for a, b in zip(grads[1], grads[K]):
            assert torch.equal(a, b), K
    assert all(bool(torch.isfinite(t).all()) for t in grads[1])
    s.close()


def test_forces_autograd_function():
    from nbfmm_amd import Solver
    from nbfmm_amd.autograd import fmm_forces
    rng = np.random.default_rng(8)
    n, L = 3000, 4
    lim = (-1.0, -1.0, 1.0, 1.0)
    pos = rng.uniform(-0.9, 0.9, (n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    g = rng.normal(0, 1, (n, 2)).astype(np.float32)
    s = Solver(L, n, lim, 1e-3)
    x = dev(pos).requires_grad_(True)
    m = dev(w).requires_grad_(True)
    (fmm_forces(s, x, m) * dev(g)).sum().backward()
    gx, gm = gpu_adjoint(L, pos, w, lim, 1e-3, g)
    assert np.array_equal(x.grad.cpu().numpy(), gx)
    assert np.array_equal(m.grad.cpu().numpy(), gm)
    s.close()


def test_mass_inference_loss_decreases():
    """Fit masses to a trajectory made with perturbed masses (small version
    of the mass-inference workload, at its own G and tick): a few gradient
    steps reduce the loss."""
    from nbfmm_amd import Solver
    from nbfmm_amd.autograd import fmm_trajectory
    from nbfmm_amd.workloads import gaussian_disk_system
    wl = gaussian_disk_system(n=4096, num_level=4, steps=20)
    rng = np.random.default_rng(21)
    s = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    true_m = dev(wl.weight)
    obs = [10, 20]
    tick = wl.tick
    with torch.no_grad():
        target = fmm_trajectory(s, tick, 20, 5, obs, dev(wl.position), dev(wl.position_pre), true_m)
    m = (true_m * dev(rng.uniform(0.8, 1.2, wl.n))).requires_grad_(True)
    losses = []
    for _ in range(4):
        out = fmm_trajectory(s, tick, 20, 5, obs, dev(wl.position), dev(wl.position_pre), m)
        loss = ((out - target) ** 2).mean()
        losses.append(float(loss.detach()))
        m.grad = None
        loss.backward()
        with torch.no_grad():
            m -= 0.5 * losses[0] / max(float((m.grad ** 2).sum()), 1e-30) * m.grad
    assert losses[-1] < losses[0]
    s.close()


def test_batch_adjoint_equals_single():
    from nbfmm_amd import Batch, Solver
    from nbfmm_amd.workloads import many_disks
    bw = many_disks(num_systems=6, n_min=300, n_max=20000, seed=17, steps=3)
    x, p, w = bw.concat()
    rng = np.random.default_rng(4)
    g = rng.normal(0, 1, x.shape).astype(np.float32)
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=[s.tick for s in bw.systems])
    GX = torch.zeros_like(dev(x))
    GM = torch.zeros_like(dev(w))
    b.solve_adjoint(dev(x), dev(w), dev(g), GX, GM)
    # one Verlet step adjoint from lambda = g, pi = 0
    LAM, PI, GW = dev(g), torch.zeros_like(dev(x)), torch.zeros_like(dev(w))
    b.step_adjoint(dev(x), dev(w), LAM, PI, GW)
    torch.cuda.synchronize()
    off = b.offsets
    for k, s in enumerate(bw.systems):
        one = Solver(s.num_level, s.n, s.limits, s.eps)
        sl = slice(off[k], off[k + 1])
        gx = torch.zeros((s.n, 2), device="cuda")
        gm = torch.zeros(s.n, device="cuda")
        one.solve_adjoint(s.n, dev(s.position), dev(s.weight), dev(g[sl]), gx, gm)
        lam, pi, gw = dev(g[sl]), torch.zeros((s.n, 2), device="cuda"), torch.zeros(s.n, device="cuda")
        one.step_adjoint(s.n, s.tick, dev(s.position), dev(s.weight), lam, pi, gw)
        torch.cuda.synchronize()
        assert torch.equal(GX[sl], gx) and torch.equal(GM[sl], gm), k
        assert torch.equal(LAM[sl], lam) and torch.equal(PI[sl], pi) and torch.equal(GW[sl], gw), k
        one.close()
    b.close()


def test_batch_trajectory_checkpoints_and_single():
    """fmm_batch_trajectory (its backward keeps the sort of every recomputed
    step for that step's adjoint): gradients for K = 1 (nothing kept), 3 and
    steps are bitwise equal, and system by system equal the single-system
    trajectory's."""
    from nbfmm_amd import Batch, Solver
    from nbfmm_amd.autograd import fmm_batch_trajectory, fmm_trajectory
    from nbfmm_amd.workloads import many_disks
    bw = many_disks(num_systems=3, n_min=500, n_max=4000, seed=23, steps=7)
    x, p, w = bw.concat()
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=[s.tick for s in bw.systems])
    steps, obs = 7, [3, 7]
    rng = np.random.default_rng(8)
    wts = dev(rng.normal(0, 1, (len(obs), x.shape[0], 2)))
    grads = []
    for K in (1, 3, steps):
        X, P, M = (dev(a).requires_grad_(True) for a in (x, p, w))
        out = fmm_batch_trajectory(b, steps, K, obs, X, P, M)
        (out * wts).sum().backward()
        torch.cuda.synchronize()
        grads.append([t.grad.clone() for t in (X, P, M)])
    for g in grads[1:]:
        for a, c in zip(grads[0], g):
            assert torch.equal(a, c)
    off = b.offsets
    for k, sy in enumerate(bw.systems):
        s = Solver(sy.num_level, sy.n, sy.limits, sy.eps)
        sl = slice(off[k], off[k + 1])
        X, P, M = (dev(a[sl]).requires_grad_(True) for a in (x, p, w))
        out = fmm_trajectory(s, sy.tick, steps, 3, obs, X, P, M)
        (out * wts[:, sl]).sum().backward()
        torch.cuda.synchronize()
        for a, t in zip(grads[0], (X, P, M)):
            assert torch.equal(a[sl], t.grad), k
        s.close()
    b.close()


@pytest.mark.parametrize("g_per_n,bar", [(1e4, 2e-3), (3e4, 1e-3)])
def test_mass_inference_fd_check_256_subset(g_per_n, bar):
    """The mass-inference configuration's finite-difference check as stated:
    an N = 256 subset of its 65536-particle Gaussian disk (first 256
    particles, their velocities; its box, 128 x 128 leaves, eps and dt), 50
    Verlet steps checkpointed every 10, loss = mean squared distance of the
    final positions to targets from a run with the masses perturbed by 5%
    Gaussian noise; the GPU (fp32) gradients with respect to initial
    positions (at fixed velocity), velocities and masses against central
    differences with h = 1e-3 of the fp64 oracle's loss (32 particles'
    positions and velocities, 128 particles' masses), per group.

    G: the workload's G N = 1e4 (G = 1e4 / 256; at the full system's G the
    subset's forces are 256 times weaker and the mass signal, rms 1.3e-5, is
    at the fp32 trajectory's rounding, 8e-6), and 3e4.  The GPU gradient is
    seeded with the fp32 trajectory's 2 (x - target) / N, whose rounding
    against the signal (rms 3.4e-6 against 3.9e-3 at 1e4, 2.9e-6 against
    1.2e-2 at 3e4) bounds the agreement at about 8.6e-4 and 2.5e-4: the bar
    is 2e-3 at 1e4 and 1e-3 at 3e4 (measured: positions, velocities, masses
    to 1.06e-3, 6.7e-4, 2.6e-4 at 1e4 and 2.4e-4, 1.1e-4, 7.2e-5 at 3e4).

    The differences are taken with every step's leaf assignment held at the
    base trajectory's (oracle run_f64_cells), the function the adjoint
    differentiates: with free cells the loss jumps wherever a perturbed
    trajectory moves a leaf crossing to another step (leaves are 0.033 wide
    and particles move about 0.01 per step, so no 1e-3 shift of an initial
    position keeps every crossing)."""
    from nbfmm_amd import Solver
    from nbfmm_amd.autograd import fmm_trajectory
    from nbfmm_amd.workloads import FIT_G_PER_N, gaussian_disk_system
    full = gaussian_disk_system(n=65536, steps=50)
    sub = np.arange(256)
    x0, p0, w = full.position[sub], full.position_pre[sub], full.weight[sub]
    L, eps, dt, steps, K = full.num_level, full.eps, full.dt, 50, 10
    assert FIT_G_PER_N == 1e4
    tick = full.dt * np.sqrt(g_per_n / len(sub))
    lim = np.array(full.limits, np.float32)
    n = len(sub)
    rng = np.random.default_rng(31)
    w_pert = (w.astype(np.float64) * (1 + 0.05 * rng.normal(0, 1, n)))
    target, _ = orc.run_f64(L, lim, eps, tick, steps, x0.astype(np.float64), p0.astype(np.float64), w_pert)
    base = (x0.astype(np.float64), p0.astype(np.float64), w.astype(np.float64))
    _, _, cells = orc.run_f64_cells(L, lim, eps, tick, steps, *base)

    def loss64(x, p, m):
        xs, _, _ = orc.run_f64_cells(L, lim, eps, tick, steps, x, p, m, cells)
        return float(((xs - target) ** 2).sum(1).mean())

    s = Solver(L, n, full.limits, eps)
    X = dev(x0).requires_grad_(True)
    P = dev(p0).requires_grad_(True)
    M = dev(w).requires_grad_(True)
    out = fmm_trajectory(s, tick, steps, K, [steps], X, P, M)
    loss = ((out[0] - dev(target)) ** 2).sum(-1).mean()
    loss.backward()
    torch.cuda.synchronize()
    gx, gp, gm = (t.grad.double().cpu().numpy() for t in (X, P, M))
    s.close()
    g_pos = gx + gp          # x0 and x_pre moved together: velocity held
    g_vel = -dt * gp         # x_pre = x0 - v dt
    l0 = loss64(*base)
    assert abs(float(loss.detach()) - l0) <= 1e-3 * l0
    h = 1e-3
    groups = {"position": ([], []), "velocity": ([], []), "mass": ([], [])}
    for i in rng.choice(n, 32, replace=False):
        for d in range(2):
            xp, pp, xm, pm = base[0].copy(), base[1].copy(), base[0].copy(), base[1].copy()
            xp[i, d] += h; pp[i, d] += h
            xm[i, d] -= h; pm[i, d] -= h
            groups["position"][0].append(g_pos[i, d])
            groups["position"][1].append((loss64(xp, pp, base[2]) - loss64(xm, pm, base[2])) / (2 * h))
            pp, pm = base[1].copy(), base[1].copy()
            pp[i, d] -= h * dt
            pm[i, d] += h * dt
            groups["velocity"][0].append(g_vel[i, d])
            groups["velocity"][1].append((loss64(base[0], pp, base[2]) - loss64(base[0], pm, base[2])) / (2 * h))
    for i in rng.choice(n, 128, replace=False):
        mp, mm = base[2].copy(), base[2].copy()
        mp[i] += h
        mm[i] -= h
        groups["mass"][0].append(gm[i])
        groups["mass"][1].append((loss64(base[0], base[1], mp) - loss64(base[0], base[1], mm)) / (2 * h))
    for name, (a, f) in groups.items():
        f = np.array(f)
        print(f"{name}: rel {rel(a, f):.3e} over {len(f)} parameters, |fd| max {np.abs(f).max():.3e}")
        assert np.linalg.norm(f) > 0, name
        assert rel(a, f) <= bar, (name, rel(a, f))
