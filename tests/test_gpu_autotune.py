"""P2P autotuner: every kernel configuration gives bit-identical results,
the tuner picks one and a cached choice is reused without timing."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def test_all_configurations_bit_identical():
    from nbfmm_amd import Solver, Batch
    from nbfmm_amd.autotune import set_config, variants
    from nbfmm_amd.workloads import disk
    wl = disk(n=30000)
    s = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    x, w = dev(wl.position), dev(wl.weight)
    ref = None
    for v in range(len(variants())):
        for bpc in (1, 4, 16, 32):
            set_config(s, v, bpc)
            a = torch.zeros_like(x)
            s.solve(wl.n, x, w, a)
            torch.cuda.synchronize()
            if ref is None:
                ref = a.clone()
            assert torch.equal(a, ref), (v, bpc)
    s.close()


def test_tune_and_cache(tmp_path):
    from nbfmm_amd import Solver
    from nbfmm_amd.autotune import get_config, tune, variants
    from nbfmm_amd.workloads import disk
    wl = disk(n=16384)
    s = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    x, w = dev(wl.position), dev(wl.weight)
    path = str(tmp_path / "tune.json")
    v, b, ms = tune(s, wl.n, x, w, cache_path=path, repeats=2)
    # the candidate grids are 16, 24, 32, 48 and 64 waves per CU
    assert 0 <= v < len(variants()) and ms > 0
    assert b in {max(1, wpc // variants()[v][2]) for wpc in (16, 24, 32, 48, 64)}
    assert get_config(s) == (v, b)
    s2 = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
    v2, b2, ms2 = tune(s2, wl.n, x, w, cache_path=path)
    assert (v2, b2) == (v, b) and ms2 is None
    assert get_config(s2) == (v, b)
    s.close()
    s2.close()


def test_batch_variants_bit_identical():
    """Each variant launches its own block size (and batch launch map); all
    give the bits of the default, system by system."""
    from nbfmm_amd import Batch
    from nbfmm_amd.autotune import variants
    from nbfmm_amd.workloads import many_disks
    bw = many_disks(num_systems=5, n_min=300, n_max=30000, seed=11, steps=2)
    for w, L in zip(bw.systems, [0, 2, 4, 6, 8]):
        w.num_level = L
    x, p, w = bw.concat()
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=[s.tick for s in bw.systems])
    X, P, W = dev(x), dev(p), dev(w)
    ref = None
    configs = [(v, bpc) for v in range(len(variants())) for bpc in (1, 8, 16)]
    for v, bpc in configs:
        b.set_p2p_config(v, bpc)
        assert b.p2p_config() == (v, bpc)
        A = torch.zeros_like(X)
        b.solve(X, W, A)
        X1, P1, A1 = X.clone(), P.clone(), torch.zeros_like(X)
        b.step(X1, P1, W, A1)
        torch.cuda.synchronize()
        got = (A, X1, P1, A1)
        if ref is None:
            ref = tuple(t.clone() for t in got)
        for r, g in zip(ref, got):
            assert torch.equal(r, g), (v, bpc)
    b.close()
