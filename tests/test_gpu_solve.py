"""Whole-path GPU parity: solve, step and graph replay against the CPU oracle.

Gate (float32 forces): global relative L2 error <= 1e-5 and
max |dF_i| / RMS|F| <= 1e-4 against the oracle on the same inputs; the
permutation, head and sorted cell indices are bit-exact.  At 2^20 particles
(the two-disk collision workload) the same gates hold against the oracle,
plus momentum conservation.
"""
import numpy as np
import pytest
import torch

import npfmm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

RTOL_L2 = 1e-5       # global relative L2 of forces
RTOL_MAX = 1e-4      # max |dF_i| / RMS |F|


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = float(np.linalg.norm(b))
    d = float(np.linalg.norm(a - b))
    return d / nb if nb > 0 else d


def max_rel(a, b):
    b = np.asarray(b, np.float64).reshape(-1, 2)
    rms = float(np.sqrt((b ** 2).sum(1).mean()))
    d = float(np.abs(np.asarray(a, np.float64).reshape(-1, 2) - b).max())
    return d / rms if rms > 0 else d


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def gpu_solve(L, pos, w, limits, eps, capacity=None):
    from nbfmm_amd import Solver
    n = pos.shape[0]
    s = Solver(L, capacity or max(n, 1), limits, eps)
    out = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    s.solve(n, dev(pos), dev(w), out)
    torch.cuda.synchronize()
    r = out.cpu().numpy()
    return r, s


def test_solve_fixture(golden):
    g = golden("solve_disk_L4_N2048.npz")
    eff, s = gpu_solve(int(g["L"]), g["position"], g["weight"], g["limits"], float(g["eps"]))
    assert rel_l2(eff, g["fmm_effect"]) < RTOL_L2
    assert max_rel(eff, g["fmm_effect"]) < RTOL_MAX
    ref, _ = orc.solve(int(g["L"]), g["position"], g["weight"], g["limits"], float(g["eps"]))
    assert rel_l2(eff, ref) < RTOL_L2
    assert s.flags() == 0
    s.close()


@pytest.mark.parametrize("L,n", [(0, 1), (0, 777), (1, 64), (2, 1000), (3, 1024), (4, 16384), (5, 5000),
                                 (6, 65536), (7, 20000), (8, 3000)])
def test_solve_matches_oracle(L, n):
    rng = np.random.default_rng(1000 + 17 * L + n)
    lim = np.array([-1.5, -2.0, 2.5, 1.0], np.float32)
    pos = np.stack([rng.uniform(-1.5, 2.5, n), rng.uniform(-2.0, 1.0, n)], 1).astype(np.float32)
    # clustered half to get dense and empty cells
    k = n // 2
    pos[:k] = (rng.normal(0, 0.2, (k, 2)) + [0.5, -0.5]).clip([-1.49, -1.99], [2.49, 0.99]).astype(np.float32)
    w = rng.exponential(1.0, n).astype(np.float32)
    eff, s = gpu_solve(L, pos, w, lim, 1e-3, capacity=n + 13)
    ref, clamped = orc.solve(L, pos, w, lim, 1e-3)
    assert clamped == 0
    assert rel_l2(eff, ref) < RTOL_L2, rel_l2(eff, ref)
    assert max_rel(eff, ref) < RTOL_MAX, max_rel(eff, ref)
    # sorted indices / perm / head bit-exact
    pre = orc.predo(L, pos, w, lim)
    np.testing.assert_array_equal(s.read_buffer("perm", n).cpu().numpy(), pre["perm"])
    np.testing.assert_array_equal(s.read_buffer("head").cpu().numpy(), pre["head"])
    np.testing.assert_array_equal(s.read_buffer("index", n).cpu().numpy(), pre["index"])
    s.close()


def test_solve_reuse_smaller_n_and_zero():
    """One solver, calls of decreasing size (solve.cu:17 allows N <= max)."""
    from nbfmm_amd import Solver
    rng = np.random.default_rng(5)
    lim = np.array([0, 0, 1, 1], np.float32)
    s = Solver(5, 4096, lim, 1e-3)
    for n in (4096, 1500, 1, 0, 333):
        pos = rng.uniform(0, 1, (n, 2)).astype(np.float32)
        w = rng.uniform(0.5, 1.5, n).astype(np.float32)
        out = torch.zeros((max(n, 1), 2), dtype=torch.float32, device="cuda")
        s.solve(n, dev(pos) if n else out, dev(w) if n else out[:, 0].contiguous(), out)
        torch.cuda.synchronize()
        if n:
            ref, _ = orc.solve(5, pos, w, lim, 1e-3)
            assert rel_l2(out.cpu().numpy()[:n], ref) < RTOL_L2
        head = s.read_buffer("head").cpu().numpy()
        assert head[-1] == n and np.all(np.diff(head) >= 0)
    s.close()


def test_capacity_error():
    from nbfmm_amd import NbfmmError, Solver
    s = Solver(3, 100, (0, 0, 1, 1), 0.1)
    x = torch.zeros((200, 2), device="cuda")
    with pytest.raises(NbfmmError) as e:
        s.solve(200, x, x[:, 0].contiguous(), x)
    assert e.value.code == 4
    s.close()


def test_out_of_box_is_flagged_and_clamped():
    """A particle at x_max (H5: the reference indexes cell B) or outside the box
    is clamped into the edge cell and raises NBFMM_FLAG_OUT_OF_BOX."""
    from nbfmm_amd import FLAG_OUT_OF_BOX
    rng = np.random.default_rng(2)
    lim = np.array([-1, -1, 1, 1], np.float32)
    pos = rng.uniform(-0.9, 0.9, (500, 2)).astype(np.float32)
    pos[0] = (1.0, 0.0)     # exactly x_max
    pos[1] = (-1.5, 0.2)    # outside
    w = np.ones(500, np.float32)
    eff, s = gpu_solve(3, pos, w, lim, 1e-3)
    assert s.flags(clear=True) & FLAG_OUT_OF_BOX
    assert s.flags() == 0
    ref, clamped = orc.solve(3, pos, w, lim, 1e-3)
    assert clamped == 2
    assert np.all(np.isfinite(eff))
    assert rel_l2(eff, ref) < RTOL_L2
    s.close()


def test_step_equals_solve_then_update():
    """Stars::update (update.cu:47-55): the fused step gives the same bits as
    solve followed by updateDevice."""
    from nbfmm_amd import Solver, update
    from nbfmm_amd.workloads import disk
    wl = disk(n=4096)
    s = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    x1, p1, w = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    x2, p2 = x1.clone(), p1.clone()
    a1 = torch.zeros_like(x1)
    a2 = torch.zeros_like(x1)
    for _ in range(3):
        s.step(wl.n, wl.tick, x1, p1, w, a1)
        s.solve(wl.n, x2, w, a2)
        update(wl.n, wl.tick, x2, p2, a2)
    torch.cuda.synchronize()
    assert torch.equal(a1, a2) and torch.equal(x1, x2) and torch.equal(p1, p2)
    s.close()


def test_graph_replay_equals_eager():
    """hipGraph replay of the step is bit-identical to eager launches; the
    replays are checked one by one with eager work in between."""
    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import disk
    wl = disk(n=8192)
    s1 = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    s2 = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    w = dev(wl.weight)
    x1, p1 = dev(wl.position), dev(wl.position_pre)
    x2, p2 = x1.clone(), p1.clone()
    a1, a2 = torch.zeros_like(x1), torch.zeros_like(x1)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(4):
            s1.run(wl.n, wl.tick, 1, x1, p1, w, a1, use_graph=True)
            s2.run(wl.n, wl.tick, 1, x2, p2, w, a2, use_graph=False)
            stream.synchronize()
            assert torch.equal(x1, x2) and torch.equal(p1, p2) and torch.equal(a1, a2)
    s1.close()
    s2.close()


def test_run_matches_oracle_trajectory():
    """Ten Verlet steps of the disk workload (N=16384, 32x32 leaves) against
    the oracle's Stars::update loop; positions agree to float32 round-off."""
    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import disk
    wl = disk(n=16384)
    s = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    x, p, w = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    a = torch.zeros_like(x)
    s.run(wl.n, wl.tick, wl.steps, x, p, w, a, use_graph=False)
    torch.cuda.synchronize()
    ox, op, oa, clamped = orc.run(wl.num_level, np.array(wl.limits, np.float32), wl.eps, wl.tick, wl.steps,
                                  wl.position, wl.position_pre, wl.weight)
    assert clamped == 0
    xg = x.cpu().numpy()
    disp = np.abs(ox - wl.position).max()
    assert np.abs(xg - ox).max() < 1e-3 * disp + 1e-6
    assert rel_l2(a.cpu().numpy(), oa) < RTOL_L2
    s.close()


def test_full_size_two_disk_collision():
    """2^20 particles, 512x512 leaves: bit-exact perm/head against the
    oracle's counting sort, force parity against the oracle's solve, and
    momentum conservation."""
    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import two_disk_collision
    wl = two_disk_collision()
    lim = np.array(wl.limits, np.float32)
    s = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    x, w = dev(wl.position), dev(wl.weight)
    a = torch.zeros_like(x)
    s.solve(wl.n, x, w, a)
    torch.cuda.synchronize()
    assert s.flags() == 0
    pre = orc.predo(wl.num_level, wl.position, wl.weight, lim)
    np.testing.assert_array_equal(s.read_buffer("perm", wl.n).cpu().numpy(), pre["perm"])
    np.testing.assert_array_equal(s.read_buffer("head").cpu().numpy(), pre["head"])
    eff = a.cpu().numpy()
    ref, _ = orc.solve(wl.num_level, wl.position, wl.weight, lim, wl.eps)
    assert rel_l2(eff, ref) < RTOL_L2, rel_l2(eff, ref)
    assert max_rel(eff, ref) < RTOL_MAX, max_rel(eff, ref)
    mom = (wl.weight[:, None].astype(np.float64) * eff).sum(0)
    scale = (wl.weight[:, None].astype(np.float64) * np.abs(eff)).sum(0)
    assert np.all(np.abs(mom) < 1e-5 * scale)
    s.close()


@pytest.mark.parametrize("L,n", [(2, 3000), (6, 30000), (8, 50000)])
def test_solve_equals_stage_pipeline_bitwise(L, n):
    """solve() runs a fused upward pass (P2M + all M2M levels) and folds
    L2L/L2P/postdo into the P2P kernel; the eight stages run one by one
    (solve.cu:19-26) must give the same bits, multipoles included."""
    from nbfmm_amd import Solver
    rng = np.random.default_rng(77 + L)
    lim = (-1.0, -1.0, 1.0, 1.0)
    pos = rng.normal(0, 0.3, (n, 2)).clip(-0.999, 0.999).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    s = Solver(L, n, lim, 1e-3)
    x, m = dev(pos), dev(w)
    a1 = torch.zeros_like(x)
    a2 = torch.zeros_like(x)
    s.solve(n, x, m, a1)
    cp1 = s.read_buffer("cell_position")
    cw1 = s.read_buffer("cell_weight")
    s.predo(n, x, m)
    s.p2p(n)
    s.p2m(n)
    s.m2m()
    s.m2l()
    s.l2l()
    s.l2p(n)
    s.postdo(n, a2)
    cp2 = s.read_buffer("cell_position")
    cw2 = s.read_buffer("cell_weight")
    torch.cuda.synchronize()
    assert torch.equal(a1, a2)
    assert torch.equal(cp1, cp2) and torch.equal(cw1, cw2)
    s.close()


def _leaf_layout(L, counts, seed):
    """Particles placed inside chosen leaves: counts[(ix, iy)] particles in
    leaf (ix, iy) of the box [-1, 1]^2, uniform inside the leaf."""
    rng = np.random.default_rng(seed)
    B = 1 << (L + 1)
    w = 2.0 / B
    pts = []
    for (ix, iy), c in counts.items():
        lo = np.array([-1.0 + ix * w, -1.0 + iy * w])
        pts.append(lo + w * (0.1 + 0.8 * rng.random((c, 2))))
    pos = np.concatenate(pts).astype(np.float32)
    perm = rng.permutation(pos.shape[0])
    return pos[perm], (0.5 + rng.random(pos.shape[0])).astype(np.float32)


@pytest.mark.parametrize("case", ["one_leaf", "exact_powers", "mixed_classes"])
def test_leaf_size_classes(case):
    """Extreme leaf populations: everything in one leaf (runs of 64 inside a
    single leaf, a long P2M sum), every occupied leaf holding exactly 2^k
    particles, and a mix of leaf sizes from 1 to 2^11 scattered over the
    grid.  All match the oracle."""
    from nbfmm_amd import Solver
    L = 4
    B = 1 << (L + 1)
    if case == "one_leaf":
        counts = {(7, 9): 5000}                       # top class holds one leaf of 5000
    elif case == "exact_powers":
        k = 3                                         # every leaf exactly 2^3: class 3 holds n >> 3 leaves
        counts = {(ix, iy): 1 << k for ix in range(B) for iy in range(0, B, 2)}
    else:
        rng = np.random.default_rng(5)
        counts = {}
        for j in range(300):
            ix, iy = int(rng.integers(B)), int(rng.integers(B))
            counts[(ix, iy)] = int(2 ** rng.integers(0, 12) + rng.integers(0, 3))
    pos, m = _leaf_layout(L, counts, seed=17)
    n = pos.shape[0]
    lim = np.array([-1, -1, 1, 1], np.float32)
    eps = 1e-3
    s = Solver(L, n, lim, eps)
    acc = torch.zeros((n, 2), device="cuda")
    s.solve(n, dev(pos), dev(m), acc)
    torch.cuda.synchronize()
    ref, _ = orc.solve(L, pos, m, lim, eps)
    got = acc.cpu().numpy()
    assert rel_l2(got, ref) <= RTOL_L2, (case, rel_l2(got, ref))
    assert max_rel(got, ref) <= RTOL_MAX, (case, max_rel(got, ref))
    s.close()


def test_graph_run_on_the_null_stream():
    """run(use_graph=True) on torch's default (null) stream: the replays go to
    the solver's own stream, ordered with the null stream, and equal eager
    steps bit for bit."""
    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import disk
    wl = disk(n=6000, steps=5)
    a = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    b = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    Xa, Pa, W = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    Xb, Pb = Xa.clone(), Pa.clone()
    Aa, Ab = torch.zeros_like(Xa), torch.zeros_like(Xa)
    a.run(wl.n, wl.tick, 5, Xa, Pa, W, Aa)  # default stream, default use_graph=True
    Xa.add_(0.0)  # work queued after the call on the null stream sees its results
    for _ in range(5):
        b.step(wl.n, wl.tick, Xb, Pb, W, Ab)
    torch.cuda.synchronize()
    assert torch.equal(Xa, Xb) and torch.equal(Pa, Pb) and torch.equal(Aa, Ab)
    a.close()
    b.close()
