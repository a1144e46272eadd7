"""Stage-isolated GPU tests: the reference's eight CppUnit tests
(check/core/solver/{predo,postdo,p2p,p2m,m2m,m2l,l2l,l2p}.cu) re-expressed
through the C ABI's buffer access.

Each test writes the stage's inputs into the solver's buffers (the
NBFMM_CHECK members), runs the one stage, reads the outputs back and compares
them with (a) the fixture's float64 values from the reference's test formula
at the reference's tolerance (abs 1e-4) and (b) the CPU oracle on the same
inputs: bit-exact for the integer and tree stages, rel 1e-5 for P2P (the GPU
kernel uses the hardware sqrt and reciprocal).
"""
import numpy as np
import pytest
import torch

import npfmm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def st(golden):
    return golden("stage_L3_N250.npz")


@pytest.fixture(scope="module")
def solver(st):
    from nbfmm_amd import Solver
    s = Solver(int(st["L"]), 256, st["limits"], float(st["eps"]))
    yield s
    s.close()


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_predo(st, solver):
    n = int(st["N"])
    solver.predo(n, dev(st["position"]), dev(st["weight"]))
    pos = host(solver.read_buffer("position", n))
    w = host(solver.read_buffer("weight", n))
    idx = host(solver.read_buffer("index", n))
    perm = host(solver.read_buffer("perm", n))
    head = host(solver.read_buffer("head"))
    np.testing.assert_array_equal(idx, st["predo_index"])
    np.testing.assert_array_equal(perm, st["predo_perm"])
    np.testing.assert_array_equal(head, st["predo_head"])
    np.testing.assert_array_equal(pos, st["predo_position"])
    np.testing.assert_array_equal(w, st["predo_weight"])
    ref = orc.predo(int(st["L"]), st["position"], st["weight"], st["limits"])
    np.testing.assert_array_equal(perm, ref["perm"])
    np.testing.assert_array_equal(head, ref["head"])


def test_postdo(st, solver):
    n = int(st["N"])
    solver.write_buffer("effect", dev(st["position"]), n)
    solver.write_buffer("perm", dev(st["perm"], torch.int32), n)
    out = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    solver.postdo(n, out)
    np.testing.assert_array_equal(host(out), st["postdo_effect_origin"])


def _load_particles(st, solver):
    n = int(st["N"])
    solver.write_buffer("position", dev(st["position"]), n)
    solver.write_buffer("weight", dev(st["weight"]), n)
    solver.write_buffer("index", dev(st["index"], torch.int32), n)
    solver.write_buffer("head", dev(st["head"], torch.int32))
    return n


def test_p2p(st, solver):
    n = _load_particles(st, solver)
    solver.p2p(n)
    eff = host(solver.read_buffer("effect", n))
    np.testing.assert_allclose(eff, st["p2p_effect"], rtol=0, atol=1e-4)
    ref = orc.p2p(int(st["L"]), st["position"], st["weight"], st["index"], st["head"], float(st["eps"]))
    assert np.linalg.norm(eff - ref) / np.linalg.norm(ref) < 1e-5


def test_p2m(st, solver):
    n = _load_particles(st, solver)
    solver.p2m(n)
    B = npfmm.base_dim(int(st["L"]))
    cp = host(solver.read_buffer("cell_position"))[:B * B]
    cw = host(solver.read_buffer("cell_weight"))[:B * B]
    np.testing.assert_allclose(cp, st["p2m_cell_position"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(cw, st["p2m_cell_weight"], rtol=0, atol=1e-4)
    rp, rw = orc.p2m(int(st["L"]), st["position"], st["weight"], st["head"])
    np.testing.assert_array_equal(cp, rp[:B * B])
    np.testing.assert_array_equal(cw, rw[:B * B])


def _levels_mask(L, B):
    m = np.zeros((L, B, B), bool)
    for l in range(L):
        m[l, ::1 << l, ::1 << l] = True
    return m


def test_m2m(st, solver):
    L = int(st["L"])
    B = npfmm.base_dim(L)
    cp = np.zeros_like(st["cube_position"])
    cw = np.zeros_like(st["cube_weight"])
    cp[:B * B] = st["cube_position"][:B * B]
    cw[:B * B] = st["cube_weight"][:B * B]
    solver.write_buffer("cell_position", dev(cp))
    solver.write_buffer("cell_weight", dev(cw))
    solver.m2m()
    gp = host(solver.read_buffer("cell_position"))
    gw = host(solver.read_buffer("cell_weight"))
    m = _levels_mask(L, B).reshape(-1)
    np.testing.assert_allclose(gp[m], st["m2m_cube_position"][m], rtol=0, atol=1e-4)
    np.testing.assert_allclose(gw[m], st["m2m_cube_weight"][m], rtol=0, atol=1e-4)
    rp, rw = orc.m2m(L, cp, cw)
    np.testing.assert_array_equal(gp[m], rp[m])
    np.testing.assert_array_equal(gw[m], rw[m])


def test_m2l(st, solver):
    L = int(st["L"])
    B = npfmm.base_dim(L)
    solver.write_buffer("cell_position", dev(st["cube_position"]))
    solver.write_buffer("cell_weight", dev(st["cube_weight"]))
    solver.m2l()
    ge = host(solver.read_buffer("cell_effect"))
    m = _levels_mask(L, B).reshape(-1)
    np.testing.assert_allclose(ge[m], st["m2l_cube_effect"][m], rtol=0, atol=1e-4)
    re_ = orc.m2l(L, float(st["eps"]), st["cube_position"], st["cube_weight"])
    np.testing.assert_array_equal(ge[m], re_[m])


def test_l2l(st, solver):
    L = int(st["L"])
    B = npfmm.base_dim(L)
    solver.write_buffer("cell_effect", dev(st["cube_position"]))
    solver.l2l()
    ge = host(solver.read_buffer("cell_effect"))[:B * B]
    np.testing.assert_allclose(ge, st["l2l_leaf_effect"], rtol=0, atol=1e-4)
    re_ = orc.l2l(L, st["cube_position"])[:B * B]
    np.testing.assert_array_equal(ge, re_)


def test_l2p(st, solver):
    n = int(st["N"])
    solver.write_buffer("effect", dev(st["position"]), n)
    solver.write_buffer("index", dev(st["index"], torch.int32), n)
    solver.write_buffer("cell_effect", dev(st["cube_position"]))
    solver.l2p(n)
    eff = host(solver.read_buffer("effect", n))
    np.testing.assert_allclose(eff, st["l2p_effect"], rtol=0, atol=1e-4)
    ref = orc.l2p(int(st["L"]), st["index"], st["cube_position"], st["position"])
    np.testing.assert_array_equal(eff, ref)


def test_update_matches_fixture(golden):
    from nbfmm_amd import update
    g = golden("verlet.npz")
    cur = dev(g["position_cur"])
    pre = dev(g["position_pre"])
    update(cur.shape[0], float(g["tick"]), cur, pre, dev(g["acceleration"]))
    np.testing.assert_array_equal(host(cur), g["expect_cur"])
    np.testing.assert_array_equal(host(pre), g["expect_pre"])
