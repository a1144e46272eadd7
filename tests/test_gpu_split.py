"""P2P runs whose source union exceeds 2048 (clusters of dense leaves) are
summed in segments by several waves and combined in segment order.  The
result must match the oracle, be the same bits for every kernel
configuration (which decides how many waves help and when a run is summed
by one wave alone), equal graph replay and batches, and never raise the
wait flag."""
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
