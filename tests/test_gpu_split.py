"""P2P runs whose source union exceeds the split threshold (2048 sources;
256 in systems of at most 65536 particles, whose runs are too few to fill the
GPU) are summed in segments by several waves and combined in segment order.  The
forces must match the oracle, be the same bits for every kernel
configuration (which decides how many waves help), equal graph replay and
batches, and never raise the wait flag.  The adjoint P2P sums such runs
whole, one wave each; on the same clustered inputs it must match central
differences of the fp64 oracle and equal batches."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def clustered(n=12000, seed=5):
    """The reference's disk launcher geometry: a disk of radius 0.75 in a
    16 x 12 box on a 32 x 32 leaf grid (L = 4), so a few leaves hold
    hundreds of stars each and runs see thousands of sources."""
    rng = np.random.default_rng(seed)
    r = 0.75 * np.sqrt(rng.random(n))
    a = 2 * np.pi * rng.random(n)
    x = np.stack([r * np.cos(a), r * np.sin(a)], 1).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return x, w, 4, np.array([-8, -6, 8, 6], np.float32), 0.1


def max_union(x, lim, L):
    """Largest staged union of a run (rows y-1..y+1, leaves x_first-1..x_last+1)."""
    B = 2 ** (L + 1)
    cell = (lim[2:] - lim[:2]) / B
    ij = np.clip(np.floor((x - lim[:2]) / cell).astype(int), 0, B - 1)
    ks = np.sort(ij[:, 1] * B + ij[:, 0])
    head = np.concatenate([[0], np.cumsum(np.bincount(ks, minlength=B * B))])
    best = 0
    for y in range(B):
        for t0 in range(head[y * B], head[(y + 1) * B], 64):
            t1 = min(t0 + 64, head[(y + 1) * B])
            xf, xl = ks[t0] % B, ks[t1 - 1] % B
            u = sum(head[r * B + min(xl + 1, B - 1) + 1] - head[r * B + max(xf - 1, 0)]
                    for r in (y - 1, y, y + 1) if 0 <= r < B)
            best = max(best, u)
    return best


def test_clustered_solve_matches_oracle_in_every_configuration():
    from nbfmm_amd import Solver
    from nbfmm_amd.autotune import set_config, variants
    x, w, L, lim, eps = clustered()
    assert max_union(x, lim, L) > 4 * 2048  # runs split into several segments
    n = x.shape[0]
    s = Solver(L, n, lim, eps)
    X, W = dev(x), dev(w)
    ref = None
    for v in range(len(variants())):
        for bpc in (1, 4, 16):
            set_config(s, v, bpc)
            a = torch.zeros_like(X)
            s.solve(n, X, W, a)
            torch.cuda.synchronize()
            if ref is None:
                ref = a.clone()
            assert torch.equal(a, ref), (v, bpc)
    assert s.flags() == 0
    want, _ = orc.solve(L, x, w, lim, eps)
    got = ref.cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5
    s.close()


def test_clustered_p2p_stage_matches_oracle():
    """The P2P-only mode (buffer API) sums split runs the same way."""
    from nbfmm_amd import Solver
    x, w, L, lim, eps = clustered(n=6000, seed=8)
    n = x.shape[0]
    s = Solver(L, n, lim, eps)
    s.predo(n, dev(x), dev(w))
    s.p2p(n)
    got = s.read_buffer("effect", n).cpu().numpy()
    pos = s.read_buffer("position", n).cpu().numpy()
    wt = s.read_buffer("weight", n).cpu().numpy()
    idx = s.read_buffer("index", n).cpu().numpy()
    head = s.read_buffer("head").cpu().numpy()
    want = orc.p2p(L, pos, wt, idx, head, eps)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5
    assert s.flags() == 0
    s.close()


def test_clustered_graph_run_equals_steps():
    from nbfmm_amd import Solver
    x, w, L, lim, eps = clustered(n=8000, seed=2)
    n = x.shape[0]
    xp = x * np.float32(0.999)
    tick = 1e-3
    a = Solver(L, n, lim, eps)
    b = Solver(L, n, lim, eps)
    Xa, Pa, Wa = dev(x), dev(xp), dev(w)
    Xb, Pb, Wb = dev(x), dev(xp), dev(w)
    Aa, Ab = torch.zeros_like(Xa), torch.zeros_like(Xb)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        a.run(n, tick, 5, Xa, Pa, Wa, Aa, use_graph=True)
        st.synchronize()
    for _ in range(5):
        b.step(n, tick, Xb, Pb, Wb, Ab)
    torch.cuda.synchronize()
    assert torch.equal(Xa, Xb) and torch.equal(Pa, Pb) and torch.equal(Aa, Ab)
    assert a.flags() == 0 and b.flags() == 0
    a.close()
    b.close()


def test_clustered_system_in_a_batch():
    from nbfmm_amd import Batch, Solver
    from nbfmm_amd.workloads import disk
    xc, wc, Lc, limc, epsc = clustered(n=10000, seed=4)
    light = disk(n=5000)
    xs = [xc, light.position]
    ws = [wc, light.weight]
    b = Batch([xc.shape[0], light.n], [Lc, light.num_level], [tuple(limc), light.limits], [epsc, light.eps])
    X, W = dev(np.concatenate(xs)), dev(np.concatenate(ws))
    A = torch.zeros_like(X)
    b.solve(X, W, A)
    torch.cuda.synchronize()
    off = b.offsets
    for k, (L, lim, eps) in enumerate([(Lc, tuple(limc), epsc), (light.num_level, light.limits, light.eps)]):
        n = xs[k].shape[0]
        s = Solver(L, n, lim, eps)
        a = torch.zeros((n, 2), device="cuda")
        s.solve(n, dev(xs[k]), dev(ws[k]), a)
        torch.cuda.synchronize()
        assert torch.equal(A[off[k]:off[k + 1]], a), k
        s.close()
    assert not b.flags().any()
    b.close()


def inside(x, lim, L, margin=1e-3):
    """Moves every point at least `margin` of a leaf width inside its leaf
    (fp64), so that fp32 and fp64 bin it alike and small steps keep it."""
    B = 2 ** (L + 1)
    cell = (lim[2:].astype(np.float64) - lim[:2]) / B
    f = (x.astype(np.float64) - lim[:2]) / cell
    k = np.floor(f)
    f = k + np.clip(f - k, margin, 1 - margin)
    return (lim[:2] + f * cell).astype(np.float32)


def test_clustered_adjoint_matches_fd_directions():
    """The adjoint on runs with thousands of sources.  Pinned by central
    differences of the fp64 oracle solve (cells held fixed) along
    the computed gradient itself and along a random direction, for the
    positions and for the masses."""
    from nbfmm_amd import Solver
    x, w, L, lim, eps = clustered(n=12000, seed=6)
    x = inside(x, lim, L)
    assert max_union(x, lim, L) > 4 * 2048
    n = x.shape[0]
    rng = np.random.default_rng(9)
    g = rng.normal(0, 1, (n, 2)).astype(np.float32)
    s = Solver(L, n, lim, eps)
    gx = torch.zeros((n, 2), device="cuda")
    gm = torch.zeros(n, device="cuda")
    s.solve_adjoint(n, dev(x), dev(w), dev(g), gx, gm)
    torch.cuda.synchronize()
    assert s.flags() == 0
    s.close()
    gx = gx.double().cpu().numpy()
    gm = gm.double().cpu().numpy()
    x64, w64, g64 = x.astype(np.float64), w.astype(np.float64), g.astype(np.float64)
    lim64 = lim.astype(np.float64)

    def loss(p, m):
        return float((g64 * orc.solve_f64(L, p, m, lim64, eps)).sum())

    h = 1e-5
    for v in (gx / np.linalg.norm(gx), rng.normal(0, 1, (n, 2)) / np.sqrt(2 * n)):
        fd = (loss(x64 + h * v, w64) - loss(x64 - h * v, w64)) / (2 * h)
        assert abs(fd - float((gx * v).sum())) < 1e-4 * np.linalg.norm(gx), (fd, float((gx * v).sum()))
    for u in (gm / np.linalg.norm(gm), rng.normal(0, 1, n) / np.sqrt(n)):
        fd = (loss(x64, w64 + h * u) - loss(x64, w64 - h * u)) / (2 * h)
        assert abs(fd - float((gm * u).sum())) < 1e-4 * np.linalg.norm(gm), (fd, float((gm * u).sum()))


def test_clustered_adjoint_identities_and_batch():
    """Translation invariance and mass homogeneity of the adjoint, and a
    clustered system in a batch: the same bits as alone, for the solve and
    the Verlet-step adjoints."""
    from nbfmm_amd import Batch, Solver
    from nbfmm_amd.workloads import disk
    xc, wc, Lc, limc, epsc = clustered(n=10000, seed=3)
    n = xc.shape[0]
    rng = np.random.default_rng(2)
    g = rng.normal(0, 1, (n, 2)).astype(np.float32)
    s = Solver(Lc, n, limc, epsc)
    a = torch.zeros((n, 2), device="cuda")
    gx = torch.zeros((n, 2), device="cuda")
    gm = torch.zeros(n, device="cuda")
    s.solve(n, dev(xc), dev(wc), a)
    s.solve_adjoint(n, dev(xc), dev(wc), dev(g), gx, gm)
    lam, pi, gw = dev(g), torch.zeros((n, 2), device="cuda"), torch.zeros(n, device="cuda")
    s.step_adjoint(n, 1e-3, dev(xc), dev(wc), lam, pi, gw)
    torch.cuda.synchronize()
    assert s.flags() == 0
    gxn, gmn, an = gx.double().cpu().numpy(), gm.double().cpu().numpy(), a.double().cpu().numpy()
    assert np.all(np.abs(gxn.sum(0)) < 1e-4 * np.abs(gxn).sum(0))
    lhs = float((wc.astype(np.float64) * gmn).sum())
    rhs = float((g.astype(np.float64) * an).sum())
    assert abs(lhs - rhs) < 1e-4 * np.abs(wc * gmn).sum()

    light = disk(n=5000)
    gl = rng.normal(0, 1, (light.n, 2)).astype(np.float32)
    b = Batch([n, light.n], [Lc, light.num_level], [tuple(limc), light.limits], [epsc, light.eps],
              ticks=[1e-3, light.tick])
    X = dev(np.concatenate([xc, light.position]))
    W = dev(np.concatenate([wc, light.weight]))
    G = dev(np.concatenate([g, gl]))
    GX, GM = torch.zeros_like(X), torch.zeros_like(W)
    b.solve_adjoint(X, W, G, GX, GM)
    LAM, PI, GW = G.clone(), torch.zeros_like(X), torch.zeros_like(W)
    b.step_adjoint(X, W, LAM, PI, GW)
    torch.cuda.synchronize()
    assert torch.equal(GX[:n], gx) and torch.equal(GM[:n], gm)
    assert torch.equal(LAM[:n], lam) and torch.equal(PI[:n], pi) and torch.equal(GW[:n], gw)
    assert not b.flags().any()
    b.close()
    s.close()


def test_clustered_p2m_bit_exact():
    """Leaves of hundreds to thousands of stars: the fused solve's P2M (a
    whole wave per leaf above 256 particles, segment sums added in order) and
    the stage P2M equal the oracle's segment-order P2M bit for bit, on the
    solver's own sorted buffers."""
    from nbfmm_amd import Solver
    x, w, L, lim, eps = clustered(n=12000, seed=11)
    n = x.shape[0]
    B = 2 ** (L + 1)
    s = Solver(L, n, lim, eps)
    a = torch.zeros((n, 2), device="cuda")
    s.solve(n, dev(x), dev(w), a)
    torch.cuda.synchronize()
    pos = s.read_buffer("position", n).cpu().numpy()
    wt = s.read_buffer("weight", n).cpu().numpy()
    head = s.read_buffer("head").cpu().numpy()
    counts = np.diff(head)
    assert counts.max() > 1000 and ((counts > 16) & (counts <= 256)).any()
    rp, rw = orc.p2m(L, pos, wt, head)
    cp = s.read_buffer("cell_position").cpu().numpy()[:B * B]
    cw = s.read_buffer("cell_weight").cpu().numpy()[:B * B]
    np.testing.assert_array_equal(cw, rw[:B * B])
    np.testing.assert_array_equal(cp, rp[:B * B])
    s.p2m(n)  # the stage kernel, one thread per leaf
    torch.cuda.synchronize()
    np.testing.assert_array_equal(s.read_buffer("cell_weight").cpu().numpy()[:B * B], rw[:B * B])
    np.testing.assert_array_equal(s.read_buffer("cell_position").cpu().numpy()[:B * B], rp[:B * B])
    s.close()


FLAG_LOST = 4


@pytest.mark.parametrize("polls", [1, 3])
def test_forced_give_up_solve_matches_oracle(polls):
    """The wait bound set to 1 or 3 polls from the host: waves holding a
    segment give up on the segments before theirs, abandon the run and sum
    its whole union alone; waves with a heavy run that finds the slot taken
    sum it alone.  Every kernel still ends normally, each run's epilogue runs
    once, and a run summed alone is folded at its segment bounds, so the
    forces equal the default solver's bit for bit (and the oracle's to
    rounding) in every configuration."""
    from nbfmm_amd import FLAG_P2P_WAIT, Solver
    from nbfmm_amd.autotune import set_config, variants
    # about two heavy runs per block on the smallest grid, so that a heavy run
    # also finds its block's slot taken
    x, w, L, lim, eps = clustered(n=30000, seed=12)
    assert max_union(x, lim, L) > 8 * 2048
    n = x.shape[0]
    s = Solver(L, n, lim, eps)
    s.set_p2p_wait_polls(polls)
    X, W = dev(x), dev(w)
    want, _ = orc.solve(L, x, w, lim, eps)
    ref = Solver(L, n, lim, eps)
    a_ref = torch.zeros_like(X)
    ref.solve(n, X, W, a_ref)
    torch.cuda.synchronize()
    assert ref.flags() == 0
    bits = a_ref.cpu().numpy()
    ref.close()
    raised = 0
    for v in range(len(variants())):
        for bpc in (1, 16):
            set_config(s, v, bpc)
            a = torch.zeros_like(X)
            s.solve(n, X, W, a)
            torch.cuda.synchronize()
            got = a.cpu().numpy()
            assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5, (v, bpc)
            np.testing.assert_array_equal(got, bits, err_msg=str((v, bpc)))
            f = s.flags(clear=True)
            assert not f & FLAG_LOST, (v, bpc, f)
            raised |= f
    assert raised & FLAG_P2P_WAIT  # the give-up path was taken
    s.close()


def test_forced_give_up_steps_match_default():
    """Graph-replayed Verlet steps (the epilogue updates x and x_pre in place,
    so a run finished twice or not at all would show at once) with the wait
    bound at 1 poll: the trajectory equals the default solver's bit for bit
    (a run summed alone is folded at its segment bounds)."""
    from nbfmm_amd import FLAG_P2P_WAIT, Solver
    x, w, L, lim, eps = clustered(n=8000, seed=13)
    n = x.shape[0]
    xp = x * np.float32(0.999)
    tick = 1e-3
    forced = Solver(L, n, lim, eps)
    forced.set_p2p_wait_polls(1)
    ref = Solver(L, n, lim, eps)
    Xa, Pa, Wa = dev(x), dev(xp), dev(w)
    Xb, Pb, Wb = dev(x), dev(xp), dev(w)
    Aa, Ab = torch.zeros_like(Xa), torch.zeros_like(Xb)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        forced.run(n, tick, 6, Xa, Pa, Wa, Aa, use_graph=True)
        ref.run(n, tick, 6, Xb, Pb, Wb, Ab, use_graph=True)
    st.synchronize()
    assert torch.equal(Aa, Ab) and torch.equal(Xa, Xb) and torch.equal(Pa, Pb)
    f = forced.flags()
    assert f & FLAG_P2P_WAIT and not f & FLAG_LOST
    assert ref.flags() == 0
    forced.close()
    ref.close()


def test_forced_give_up_in_a_batch():
    from nbfmm_amd import Batch
    from nbfmm_amd.workloads import disk
    xc, wc, Lc, limc, epsc = clustered(n=10000, seed=14)
    light = disk(n=5000)
    b = Batch([xc.shape[0], light.n], [Lc, light.num_level], [tuple(limc), light.limits], [epsc, light.eps])
    b.set_p2p_wait_polls(1)
    X, W = dev(np.concatenate([xc, light.position])), dev(np.concatenate([wc, light.weight]))
    A = torch.zeros_like(X)
    b.solve(X, W, A)
    torch.cuda.synchronize()
    off = b.offsets
    for k, (x, w, L, lim, eps) in enumerate([(xc, wc, Lc, limc, epsc),
                                             (light.position, light.weight, light.num_level,
                                              np.array(light.limits, np.float32), light.eps)]):
        want, _ = orc.solve(L, x, w, lim, eps)
        got = A[off[k]:off[k + 1]].cpu().numpy()
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5, k
        # and bit for bit the system's own solver at the default wait bound
        from nbfmm_amd import Solver
        s = Solver(L, x.shape[0], lim, eps)
        a1 = torch.zeros_like(A[off[k]:off[k + 1]])
        s.solve(x.shape[0], X[off[k]:off[k + 1]], W[off[k]:off[k + 1]], a1)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got, a1.cpu().numpy(), err_msg=str(k))
        s.close()
    assert not (b.flags() & FLAG_LOST).any()
    b.close()
