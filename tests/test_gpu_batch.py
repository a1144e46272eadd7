"""Batched systems: every stage is one launch over all systems, and each
system's results are bit-identical to a separate Solver on the same inputs
(same device code, same summation order).  Graph replay equals eager."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


@pytest.fixture(scope="module")
def small_batch():
    from nbfmm_amd.workloads import many_disks
    # uneven sizes and depths, from a single leaf-level system up to 512x512
    bw = many_disks(num_systems=9, n_min=200, n_max=40000, seed=7, steps=3)
    levels = [0, 1, 2, 3, 4, 5, 6, 7, 8]
    for w, L in zip(bw.systems, levels):
        w.num_level = L
    return bw


def test_batch_solve_equals_single_solvers(small_batch):
    from nbfmm_amd import Batch, Solver
    bw = small_batch
    x, _, w = bw.concat()
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems])
    X, W = dev(x), dev(w)
    A = torch.zeros_like(X)
    b.solve(X, W, A)
    torch.cuda.synchronize()
    off = b.offsets
    for k, s in enumerate(bw.systems):
        one = Solver(s.num_level, s.n, s.limits, s.eps)
        a = torch.zeros((s.n, 2), device="cuda")
        one.solve(s.n, dev(s.position), dev(s.weight), a)
        torch.cuda.synchronize()
        assert torch.equal(A[off[k]:off[k + 1]], a), k
        ref, _ = orc.solve(s.num_level, s.position, s.weight, np.array(s.limits, np.float32), s.eps)
        g = a.cpu().numpy()
        assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-5
        one.close()
    assert not b.flags().any()
    b.close()


def test_batch_run_graph_equals_single_steps(small_batch):
    from nbfmm_amd import Batch, Solver
    bw = small_batch
    x, p, w = bw.concat()
    ticks = [s.tick for s in bw.systems]
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=ticks)
    X, P, W = dev(x), dev(p), dev(w)
    A = torch.zeros_like(X)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(3):
            b.run(1, X, P, W, A, use_graph=True)
            stream.synchronize()
    off = b.offsets
    for k, s in enumerate(bw.systems):
        one = Solver(s.num_level, s.n, s.limits, s.eps)
        xs, ps, ws = dev(s.position), dev(s.position_pre), dev(s.weight)
        a = torch.zeros_like(xs)
        for _ in range(3):
            one.step(s.n, s.tick, xs, ps, ws, a)
        torch.cuda.synchronize()
        assert torch.equal(X[off[k]:off[k + 1]], xs), k
        assert torch.equal(P[off[k]:off[k + 1]], ps), k
        assert torch.equal(A[off[k]:off[k + 1]], a), k
        one.close()
    b.close()


def test_batch_runs_eager_and_graph_equal_steps(small_batch):
    """A run of several steps, eager (every step but the last makes the next
    sort's first pass, no graph) and replayed as graphs, equals as many
    batch.step calls bit for bit, in positions and in the last step's
    accelerations."""
    from nbfmm_amd import Batch
    bw = small_batch
    x, p, w = bw.concat()
    ticks = [s.tick for s in bw.systems]
    args = ([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
            [s.eps for s in bw.systems])
    stream = torch.cuda.Stream()
    out = []
    for how in ("steps", "eager", "graph"):
        b = Batch(*args, ticks=ticks)
        X, P, W = dev(x), dev(p), dev(w)
        A = torch.zeros_like(X)
        with torch.cuda.stream(stream):
            for n in (3, 1, 5):
                if how == "steps":
                    for _ in range(n):
                        b.step(X, P, W, A)
                else:
                    b.run(n, X, P, W, A, use_graph=how == "graph")
        stream.synchronize()
        assert not b.flags().any(), how
        out.append((X, P, A))
        b.close()
    for got in out[1:]:
        for t_ref, t in zip(out[0], got):
            assert torch.equal(t_ref, t)


def test_batch_balance_keeps_results(small_batch):
    """Sharing the P2P grid by pair counts moves work between blocks only:
    solve gives the same bits before and after balance(), and a graph run
    (which balances on its first call, then captures) equals single solvers."""
    from nbfmm_amd import Batch, Solver
    bw = small_batch
    x, p, w = bw.concat()
    ticks = [s.tick for s in bw.systems]
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=ticks)
    X, P, W = dev(x), dev(p), dev(w)
    A0, A1 = torch.zeros_like(X), torch.zeros_like(X)
    b.solve(X, W, A0)
    b.balance(X, W)
    b.solve(X, W, A1)
    torch.cuda.synchronize()
    assert torch.equal(A0, A1)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        b.run(4, X, P, W, A1, use_graph=True)
        stream.synchronize()
    off = b.offsets
    for k, s in enumerate(bw.systems):
        one = Solver(s.num_level, s.n, s.limits, s.eps)
        xs, ps, ws = dev(s.position), dev(s.position_pre), dev(s.weight)
        a = torch.zeros_like(xs)
        for _ in range(4):
            one.step(s.n, s.tick, xs, ps, ws, a)
        torch.cuda.synchronize()
        assert torch.equal(X[off[k]:off[k + 1]], xs), k
        assert torch.equal(A1[off[k]:off[k + 1]], a), k
        one.close()
    assert not b.flags().any()
    b.close()


def test_batch_rejects_bad_systems():
    from nbfmm_amd import Batch, NbfmmError
    with pytest.raises(NbfmmError):
        Batch([100, 0], [3, 3], [(0, 0, 1, 1)] * 2, [1e-3, 1e-3])
    with pytest.raises(NbfmmError):
        Batch([100, 10], [3, 9], [(0, 0, 1, 1)] * 2, [1e-3, 1e-3])
