"""Warm-started sort: solve/step/run start their sort from the solver's
previous sort of the same N.  Everything they produce (forces, states, the
sorted buffers) must equal, bit for bit, a solver that sorts from scratch
every call, whatever the particles did in between: small drifts (the warm
passes place the few movers), large jumps, crowded leaves (more than 64
movers into one leaf, merged 64 at a time), buffer writes and changes of N
(cold again)."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def pair(wl):
    from nbfmm_amd import Solver
    warm = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    cold = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    cold.warm_start(False)
    return warm, cold


def sorted_state(s, n):
    return [s.read_buffer(b, n).cpu() for b in ("position", "weight", "perm", "index")] + \
           [s.read_buffer("head").cpu()]


def assert_same_state(a, b, n):
    for x, y in zip(sorted_state(a, n), sorted_state(b, n)):
        assert torch.equal(x, y)
    # the occupied-leaf count each sort keeps (reported by pair_count): a
    # warm sort that switched to the cold passes once kept a stale count,
    # and its occupied-leaf list then ran off the end of its array
    for s in (a, b):
        head = s.read_buffer("head").cpu().numpy()
        assert s.pair_count(n)[1] == int((np.diff(head) > 0).sum())


def test_steps_bitwise_equal_cold():
    from nbfmm_amd.workloads import two_disk_collision
    wl = two_disk_collision(n=1 << 16, num_level=6, steps=12)
    warm, cold = pair(wl)
    Xw, Pw, W = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    Xc, Pc = Xw.clone(), Pw.clone()
    Aw, Ac = torch.zeros_like(Xw), torch.zeros_like(Xw)
    for k in range(wl.steps):
        warm.step(wl.n, wl.tick, Xw, Pw, W, Aw)
        cold.step(wl.n, wl.tick, Xc, Pc, W, Ac)
        torch.cuda.synchronize()
        assert torch.equal(Aw, Ac) and torch.equal(Xw, Xc) and torch.equal(Pw, Pc), k
    assert_same_state(warm, cold, wl.n)
    warm.close()
    cold.close()


def test_graph_run_bitwise_equal_cold():
    from nbfmm_amd.workloads import two_disk_collision
    wl = two_disk_collision(n=1 << 15, num_level=5, steps=29)
    warm, cold = pair(wl)
    Xw, Pw, W = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    Xc, Pc = Xw.clone(), Pw.clone()
    Aw, Ac = torch.zeros_like(Xw), torch.zeros_like(Xw)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        warm.run(wl.n, wl.tick, 4, Xw, Pw, W, Aw, use_graph=True)   # cold eager step, then warm replays
        warm.run(wl.n, wl.tick, 6, Xw, Pw, W, Aw, use_graph=True)   # same graph again
        warm.run(wl.n, wl.tick, 19, Xw, Pw, W, Aw, use_graph=True)  # two 8-step graphs, one single, the last
    st.synchronize()
    for _ in range(29):
        cold.step(wl.n, wl.tick, Xc, Pc, W, Ac)
    torch.cuda.synchronize()
    assert torch.equal(Aw, Ac) and torch.equal(Xw, Xc) and torch.equal(Pw, Pc)
    warm.close()
    cold.close()


@pytest.mark.parametrize("jump", ["drift", "drift_many", "shuffle", "cluster", "crowd", "crowd_fine", "crowd_coarse",
                                  "weights", "dense_few", "dense_many"])
def test_solve_sequences(jump):
    """solve at x1, then at x2: small drift (warm placement), unrelated
    positions (most particles move), 8 groups of 500 particles into 8 leaves
    (leaves with 65..1024 movers: sorted in LDS), half the particles into one
    leaf (thousands of movers), new masses only (no movers), and a leaf of
    3000 particles (its stayers walked in 47 chunks of 64, several in flight)
    that gains 30 movers (ranked in registers) or 200 (sorted in LDS).  The
    32 x 32 leaves take the one-block row passes; crowd_fine repeats the crowd
    on 256 x 256 leaves, where the row passes are a block per row, and
    crowd_coarse on 8 x 8 leaves (most of the one block's threads idle; the
    way back from the crowd merges about 150 movers into each leaf);
    drift_many drifts 70000 particles on 32 x 32 leaves, past the one-block
    row passes' particle limit."""
    import dataclasses
    from nbfmm_amd.workloads import disk
    wl = disk(n=70000 if jump == "drift_many" else 20000)
    if jump == "crowd_fine":
        wl = dataclasses.replace(wl, num_level=7)
    elif jump == "crowd_coarse":
        wl = dataclasses.replace(wl, num_level=2)
    warm, cold = pair(wl)
    rng = np.random.default_rng(9)
    x1, m1 = wl.position, wl.weight
    if jump.startswith("dense"):
        # the dense leaf at x1[0] in both states; movers come from elsewhere
        x1 = x1.copy()
        x1[1:3000] = x1[0] + 1e-6 * rng.random((2999, 2)).astype(np.float32)
        nm = 30 if jump == "dense_few" else 200
        x2 = x1.copy()
        x2[5000:5000 + nm] = x1[0] + 1e-6 * rng.random((nm, 2)).astype(np.float32)
        m2 = m1
    elif jump.startswith("drift"):
        x2 = x1 + 2e-3 * rng.standard_normal(x1.shape).astype(np.float32)
        m2 = m1
    elif jump == "shuffle":
        x2 = x1[rng.permutation(wl.n)]
        m2 = m1
    elif jump == "cluster":
        x2 = x1.copy()
        for k in range(8):
            x2[500 * k: 500 * (k + 1)] = x1[1000 + 1234 * k] + 1e-6 * rng.random((500, 2)).astype(np.float32)
        m2 = m1
    elif jump.startswith("crowd"):
        x2 = x1.copy()
        x2[: wl.n // 2] = x1[0] + 1e-6 * rng.random((wl.n // 2, 2)).astype(np.float32)
        m2 = m1
    else:
        x2 = x1
        m2 = (m1 * (0.5 + rng.random(wl.n))).astype(np.float32)
    for x, m in ((x1, m1), (x2, m2), (x1, m1)):
        aw, ac = torch.zeros((wl.n, 2), device="cuda"), torch.zeros((wl.n, 2), device="cuda")
        warm.solve(wl.n, dev(x), dev(m), aw)
        cold.solve(wl.n, dev(x), dev(m), ac)
        torch.cuda.synchronize()
        assert torch.equal(aw, ac)
        assert_same_state(warm, cold, wl.n)
    ref, _ = orc.solve(wl.num_level, x1, m1, np.array(wl.limits, np.float32), wl.eps)
    got = aw.cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    warm.close()
    cold.close()


def test_buffer_write_and_n_change_start_cold():
    from nbfmm_amd.workloads import disk
    wl = disk(n=12000)
    warm, cold = pair(wl)
    x, m = dev(wl.position), dev(wl.weight)
    seq = [wl.n, wl.n, wl.n, 7000, 7000, wl.n]
    for k, n in enumerate(seq):
        if k == 2:
            # overwrite the stored sort through the buffer API: the next sort must start cold
            warm.write_buffer("perm", torch.arange(wl.n, dtype=torch.int32, device="cuda"), wl.n)
        aw, ac = torch.zeros((n, 2), device="cuda"), torch.zeros((n, 2), device="cuda")
        warm.solve(n, x[:n], m[:n], aw)
        cold.solve(n, x[:n], m[:n], ac)
        torch.cuda.synchronize()
        assert torch.equal(aw, ac), (k, n)
        assert_same_state(warm, cold, n)
    warm.close()
    cold.close()


def test_adjoint_after_warm_sort():
    from nbfmm_amd.workloads import disk
    wl = disk(n=8192)
    warm, cold = pair(wl)
    rng = np.random.default_rng(4)
    x = dev(wl.position)
    m = dev(wl.weight)
    g = dev(rng.standard_normal((wl.n, 2)))
    a = torch.zeros_like(x)
    warm.solve(wl.n, x, m, a)                      # leaves a sort to start from
    x2 = x + 1e-3 * dev(rng.standard_normal((wl.n, 2)))
    out = []
    for s in (warm, cold):
        gp, gw = torch.zeros_like(x), torch.zeros_like(m)
        s.solve_adjoint(wl.n, x2, m, g, gp, gw)
        torch.cuda.synchronize()
        out.append((gp, gw))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    warm.close()
    cold.close()


def test_run_after_caller_edits_arrays():
    """Between two runs the caller moves every particle and changes masses:
    the next run's first step reads the caller's arrays again.  Inside a run
    the P2P of every step but the last makes the next sort's first pass, so
    runs of 1 (nothing pending), 2 (the eager step makes the pass, the last
    graph takes it) and more steps, and an eager step after a run, must all
    leave nothing pending."""
    from nbfmm_amd.workloads import two_disk_collision
    wl = two_disk_collision(n=1 << 14, num_level=5, steps=6)
    warm, cold = pair(wl)
    rng = np.random.default_rng(21)
    Xw, Pw, Ww = dev(wl.position), dev(wl.position_pre), dev(wl.weight)
    Xc, Pc, Wc = Xw.clone(), Pw.clone(), Ww.clone()
    Aw, Ac = torch.zeros_like(Xw), torch.zeros_like(Xw)
    st = torch.cuda.Stream()
    for rnd, steps in enumerate((3, 1, 2, 4)):
        with torch.cuda.stream(st):
            warm.run(wl.n, wl.tick, steps, Xw, Pw, Ww, Aw, use_graph=True)
            if rnd == 2:
                warm.step(wl.n, wl.tick, Xw, Pw, Ww, Aw)
        st.synchronize()
        for _ in range(steps + (rnd == 2)):
            cold.step(wl.n, wl.tick, Xc, Pc, Wc, Ac)
        torch.cuda.synchronize()
        assert torch.equal(Aw, Ac) and torch.equal(Xw, Xc) and torch.equal(Pw, Pc), rnd
        shift = dev(1e-2 * rng.standard_normal((wl.n, 2)))
        scale = dev(0.5 + rng.random(wl.n))
        for X, W in ((Xw, Ww), (Xc, Wc)):
            X += shift
            W *= scale
    warm.close()
    cold.close()


def _batch_pair(bw):
    from nbfmm_amd import Batch
    args = ([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
            [s.eps for s in bw.systems])
    warm = Batch(*args, ticks=[s.tick for s in bw.systems])
    warm.warm_start(True)
    cold = Batch(*args, ticks=[s.tick for s in bw.systems])
    cold.warm_start(False)
    return warm, cold


def test_batch_warm_equals_cold():
    """Every system of a batch starts its sort from the batch's previous sort:
    solves with drift and with shuffled positions (most particles move),
    then graph runs, all bitwise equal to a batch that sorts cold."""
    from nbfmm_amd.workloads import many_disks
    bw = many_disks(num_systems=6, n_min=500, n_max=30000, seed=13, steps=4)
    for w, L in zip(bw.systems, [0, 2, 3, 5, 6, 7]):
        w.num_level = L
    warm, cold = _batch_pair(bw)
    x, p, m = bw.concat()
    rng = np.random.default_rng(2)
    off = warm.offsets
    xs = [x, x + 1e-3 * rng.standard_normal(x.shape).astype(np.float32)]
    shuffled = xs[1].copy()
    k = 3  # scramble one system only
    shuffled[off[k]:off[k + 1]] = shuffled[off[k]:off[k + 1]][rng.permutation(off[k + 1] - off[k])]
    xs.append(shuffled)
    M = dev(m)
    for xi in xs:
        aw, ac = torch.zeros((x.shape[0], 2), device="cuda"), torch.zeros((x.shape[0], 2), device="cuda")
        warm.solve(dev(xi), M, aw)
        cold.solve(dev(xi), M, ac)
        torch.cuda.synchronize()
        assert torch.equal(aw, ac)
    Xw, Pw = dev(x), dev(p)
    Xc, Pc = Xw.clone(), Pw.clone()
    Aw, Ac = torch.zeros_like(Xw), torch.zeros_like(Xw)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        warm.run(3, Xw, Pw, M, Aw, use_graph=True)
        warm.run(1, Xw, Pw, M, Aw, use_graph=True)
        warm.run(2, Xw, Pw, M, Aw, use_graph=True)
        warm.run(4, Xw, Pw, M, Aw, use_graph=True)
        cold.run(10, Xc, Pc, M, Ac, use_graph=True)
    st.synchronize()
    assert torch.equal(Aw, Ac) and torch.equal(Xw, Xc) and torch.equal(Pw, Pc)
    warm.close()
    cold.close()


def _jump_state(wl, count, rng):
    """x_pre such that the first Verlet step (tiny tick, so the forces barely
    matter) carries `count` particles into one leaf and leaves the rest."""
    x = wl.position.astype(np.float64)
    target = x.copy()
    target[:count] = x[0] + 1e-6 * rng.random((count, 2))
    pre = 2 * x - target
    return wl.position, pre.astype(np.float32)


@pytest.mark.parametrize("count", [600, 5000])
def test_graph_run_switches_to_cold_inside_the_chain(count):
    """A replayed run whose second step finds more than 512 movers entering
    one leaf: that step's warm-started sort switches to the cold passes on the
    device (inside the captured graph), and every step still equals a solver
    that sorts cold, bit for bit."""
    from nbfmm_amd.workloads import disk
    wl = disk(n=20000)
    rng = np.random.default_rng(31)
    x, pre = _jump_state(wl, count, rng)
    warm, cold = pair(wl)
    tick = 1e-6
    Xw, Pw, W = dev(x), dev(pre), dev(wl.weight)
    Xc, Pc = Xw.clone(), Pw.clone()
    Aw, Ac = torch.zeros_like(Xw), torch.zeros_like(Xw)
    st = torch.cuda.Stream()
    for steps in (2, 3):  # the second step of the first run sorts the crowd
        with torch.cuda.stream(st):
            warm.run(wl.n, tick, steps, Xw, Pw, W, Aw, use_graph=True)
            for _ in range(steps):
                cold.step(wl.n, tick, Xc, Pc, W, Ac)
        st.synchronize()
        assert torch.equal(Aw, Ac) and torch.equal(Xw, Xc) and torch.equal(Pw, Pc), steps
        assert_same_state(warm, cold, wl.n)
        if steps == 2:
            # the crowded leaf holds the jumpers
            head = warm.read_buffer("head").cpu().numpy()
            assert np.diff(head).max() >= count
    warm.close()
    cold.close()


def test_batch_switches_to_cold_for_one_system():
    """In a batch, one system's leaf gains thousands of movers (its sort
    switches to the cold passes) while the others stay warm: bitwise equal to
    a batch that sorts cold."""
    from nbfmm_amd.workloads import many_disks
    bw = many_disks(num_systems=4, n_min=3000, n_max=30000, seed=19, steps=3)
    warm, cold = _batch_pair(bw)
    x, p, m = bw.concat()
    off = warm.offsets
    rng = np.random.default_rng(6)
    x2 = x.copy()
    k = 2
    n2 = (off[k + 1] - off[k]) // 2
    x2[off[k]:off[k] + n2] = x[off[k]] + 1e-6 * rng.random((n2, 2)).astype(np.float32)
    M = dev(m)
    for xi in (x, x2, x):
        aw, ac = torch.zeros((x.shape[0], 2), device="cuda"), torch.zeros((x.shape[0], 2), device="cuda")
        warm.solve(dev(xi), M, aw)
        cold.solve(dev(xi), M, ac)
        torch.cuda.synchronize()
        assert torch.equal(aw, ac)
    assert not warm.flags().any()
    warm.close()
    cold.close()
