"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

The reference ships no golden vectors: its unit tests draw inputs at run time
from std::default_random_engine (check/core/solver.cu:19-69).  These fixtures
rebuild the same constructions with a seeded numpy generator and store the
expected outputs computed by the float64 restatements of the reference's own
test formulas (tests/npfmm.py).  Nothing here calls the oracle or the HIP
library.

Fixtures
  stage_L3_N250.npz  the unit-test fixture of check/core/solver.cu:17-69:
                     L=3 (B=16), N=250, limits (0,-1,8,2); random positions,
                     exponential(1) weights, random cell data on every level,
                     a random monotone head, a random perm and the index that
                     head implies; expected outputs of predo, p2p, p2m, m2m,
                     m2l, l2l, l2p and postdo.
  solve_disk_L4_N2048.npz  a reduced cfg1: uniform disk of radius 1 in
                     [-1.1,1.1]^2, masses U[0.5,1.5], L=4 (32x32 leaves),
                     eps=1e-3; N=2048 is a multiple of 1024 (the case the
                     reference's extractHead mishandles).  Expected FMM forces
                     (float64 test formulas) and the float64 direct sum.
  verlet.npz         random states and accelerations with the float32
                     op-by-op result of update.cu:39-41.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import npfmm  # noqa: E402


def stage_fixture(seed=20160615):
    rng = np.random.default_rng(seed)
    L, N = 3, 250
    B = npfmm.base_dim(L)
    C = B * B
    limits = np.array([0, -1, 8, 2], np.float32)
    eps = np.float32(0.1)  # the reference's hard-coded epsilon (kernel_function.hpp:35)
    pos = np.stack([rng.uniform(limits[0], limits[2], N), rng.uniform(limits[1], limits[3], N)], 1).astype(np.float32)
    w = rng.exponential(1.0, N).astype(np.float32)
    # random cell data on every level at the cube slots used by that level
    cube_pos = np.zeros((L, B, B, 2), np.float32)
    cube_w = np.zeros((L, B, B), np.float32)
    for l in range(L):
        s = 1 << l
        D = B >> l
        cube_pos[l, ::s, ::s, 0] = rng.uniform(limits[0], limits[2], (D, D))
        cube_pos[l, ::s, ::s, 1] = rng.uniform(limits[1], limits[3], (D, D))
        cube_w[l, ::s, ::s] = rng.exponential(1.0, (D, D))
    # random monotone head (check/core/solver.cu:51-59)
    inc = rng.integers(0, 256, C)
    csum = np.concatenate([[0], np.cumsum(inc)])
    head = (csum * N) // csum[-1]
    perm = rng.permutation(N).astype(np.int32)
    index = np.zeros((N, 2), np.int32)
    for c in range(C):
        index[head[c]:head[c + 1]] = (c % B, c // B)

    out = dict(L=np.int32(L), N=np.int32(N), limits=limits, eps=eps, position=pos, weight=w,
               cube_position=cube_pos.reshape(-1, 2), cube_weight=cube_w.reshape(-1),
               head=head.astype(np.int32), perm=perm, index=index)

    # predo on the random positions/weights
    pre = npfmm.predo_numpy(pos, w, limits, L)
    for k in ("index", "perm", "head", "position", "weight"):
        out["predo_" + k] = pre[k]
    # p2p on the random sorted-like inputs (index consistent with head)
    out["p2p_effect"] = npfmm.p2p_allpairs(pos, w, index, float(eps))
    # p2m on the same inputs
    lp, lw = npfmm.p2m_leaf(pos, w, head, B)
    out["p2m_cell_position"] = lp
    out["p2m_cell_weight"] = lw
    # m2m from random level-0 cells
    P, W = npfmm.m2m_levels(cube_pos[0].astype(np.float64), cube_w[0].astype(np.float64), L)
    out["m2m_cube_position"] = npfmm.levels_to_cube(P, L, 2)
    out["m2m_cube_weight"] = npfmm.levels_to_cube(W, L, 1)
    # m2l on the random cube
    lv_pos = npfmm.cube_to_levels(cube_pos.reshape(-1, 2), L, 2)
    lv_w = [a[..., 0] for a in npfmm.cube_to_levels(cube_w.reshape(-1, 1), L, 1)]
    E = [npfmm.m2l_level_rule(lv_pos[l].astype(np.float64), lv_w[l].astype(np.float64), float(eps))
         for l in range(L)]
    out["m2l_cube_effect"] = npfmm.levels_to_cube(E, L, 2)
    # l2l on the random cube positions used as effects (check/core/solver/l2l.cu:22)
    E2 = npfmm.l2l_pushdown([a.astype(np.float64) for a in lv_pos])
    out["l2l_leaf_effect"] = E2[0].reshape(-1, 2)
    # l2p: initial effect = random positions, cell effect = level-0 random cell positions
    out["l2p_effect"] = pos.astype(np.float64) + cube_pos[0].astype(np.float64)[index[:, 1], index[:, 0]]
    # postdo: effect = random positions scattered by perm
    eo = np.zeros_like(pos)
    eo[perm] = pos
    out["postdo_effect_origin"] = eo
    return out


def uniform_disk(rng, n, radius=1.0):
    r = radius * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    return np.stack([r * np.cos(th), r * np.sin(th)], 1).astype(np.float32)


def solve_fixture(seed=1001):
    rng = np.random.default_rng(seed)
    L, N = 4, 2048
    limits = np.array([-1.1, -1.1, 1.1, 1.1], np.float32)
    eps = np.float32(1e-3)
    pos = uniform_disk(rng, N)
    w = rng.uniform(0.5, 1.5, N).astype(np.float32)
    fmm = npfmm.fmm_numpy(L, pos, w, limits, float(eps))
    direct = npfmm.direct64(pos, w, float(eps))
    return dict(L=np.int32(L), N=np.int32(N), limits=limits, eps=eps, position=pos, weight=w, fmm_effect=fmm,
                direct_effect=direct)


def verlet_fixture(seed=7):
    rng = np.random.default_rng(seed)
    n = 1000
    x = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    xp = (x + rng.normal(0, 1e-3, (n, 2))).astype(np.float32)
    a = rng.normal(0, 50, (n, 2)).astype(np.float32)
    tick = np.float32(1e-2)
    new, prev = npfmm.verlet_f32(tick, x, xp, a)
    return dict(tick=tick, position_cur=x, position_pre=xp, acceleration=a, expect_cur=new, expect_pre=prev)


def main():
    np.savez_compressed(os.path.join(HERE, "stage_L3_N250.npz"), **stage_fixture())
    np.savez_compressed(os.path.join(HERE, "solve_disk_L4_N2048.npz"), **solve_fixture())
    np.savez_compressed(os.path.join(HERE, "verlet.npz"), **verlet_fixture())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
