"""Independent float64 numpy formulations of the nbfmm stages.

These restate the brute-force host formulas of the reference's unit tests
(check/core/solver/*.cu), which are written differently from the reference
kernels (all-pairs masks instead of head ranges, the O(cells^2) parent-distance
rule instead of the 6x6 window, push-down instead of pull-down L2L).  They pin
the CPU oracle and write the golden fixtures; the oracle itself never uses
them.

Cube layout helpers use the reference's [levels][B][B] cube: level-l cell
(X, Y) sits at row Y<<l, column X<<l of level slice l.
"""
from __future__ import annotations

import numpy as np


def base_dim(L):
    return 1 << (L + 1)


def cell_size_f32(limits, B):
    lim = np.asarray(limits, np.float32)
    return (np.float32(lim[2] - lim[0]) / np.float32(B), np.float32(lim[3] - lim[1]) / np.float32(B))


def kernel64(t, s, m, eps):
    """include/nbfmm/core/kernel_function.hpp:29-40 in float64 (broadcasting)."""
    d = s - t
    r = np.sqrt((d * d).sum(-1))
    tmp = m / (r * r * r + eps)
    return d * tmp[..., None]


def compute_index_f32(pos, limits, L):
    """floor((p - lo)/cell) evaluated in IEEE float32 (predo.cu:37-38, :112-113)."""
    B = base_dim(L)
    cw, ch = cell_size_f32(limits, B)
    lim = np.asarray(limits, np.float32)
    pos = np.asarray(pos, np.float32)
    ix = np.floor((pos[:, 0] - lim[0]) / cw)
    iy = np.floor((pos[:, 1] - lim[1]) / ch)
    return np.stack([ix, iy], 1).astype(np.int64)


def predo_numpy(pos0, w0, limits, L):
    """Stable argsort / bincount statement of predo (check/core/solver/predo.cu:50-66)."""
    B = base_dim(L)
    idx = compute_index_f32(pos0, limits, L)
    key = idx[:, 0] + idx[:, 1] * B
    perm = np.argsort(key, kind="stable")
    head = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=B * B))])
    return dict(index=idx[perm].astype(np.int32), perm=perm.astype(np.int32), head=head.astype(np.int32),
                position=np.asarray(pos0, np.float32)[perm], weight=np.asarray(w0, np.float32)[perm])


def p2p_allpairs(pos, w, index, eps, chunk=2048):
    """check/core/solver/p2p.cu:34-41: all pairs with |dix|<=1, |diy|<=1, i != j."""
    pos = np.asarray(pos, np.float64)
    w = np.asarray(w, np.float64)
    index = np.asarray(index, np.int64)
    n = pos.shape[0]
    out = np.zeros((n, 2))
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        k = kernel64(pos[a:b, None, :], pos[None, :, :], w[None, :], eps)
        near = (np.abs(index[a:b, None, 0] - index[None, :, 0]) <= 1) & \
               (np.abs(index[a:b, None, 1] - index[None, :, 1]) <= 1)
        near[np.arange(b - a), np.arange(a, b)] = False
        out[a:b] = (k * near[..., None]).sum(1)
    return out


def p2m_leaf(pos, w, head, B):
    """check/core/solver/p2m.cu:36-46: leaf COM and mass; empty -> (0,0), 0."""
    pos = np.asarray(pos, np.float64)
    w = np.asarray(w, np.float64)
    C = B * B
    cpos = np.zeros((C, 2))
    cw = np.zeros(C)
    for c in range(C):
        a, b = int(head[c]), int(head[c + 1])
        if b > a:
            cw[c] = w[a:b].sum()
            cpos[c] = (pos[a:b] * w[a:b, None]).sum(0) / cw[c]
    return cpos, cw


def cube_to_levels(cube, L, comps):
    """Dense per-level arrays [D_l, D_l, comps] from a reference cube [levels*B*B, comps]."""
    B = base_dim(L)
    c = np.asarray(cube).reshape(-1, B, B, comps)
    return [c[l, ::1 << l, ::1 << l].copy() for l in range(max(L, 1))]


def levels_to_cube(levels, L, comps, dtype=np.float64):
    B = base_dim(L)
    cube = np.zeros((max(L, 1), B, B, comps), dtype)
    for l, a in enumerate(levels):
        cube[l, ::1 << l, ::1 << l] = np.asarray(a).reshape(cube[l, ::1 << l, ::1 << l].shape)
    return cube.reshape(-1, comps) if comps > 1 else cube.reshape(-1)


def m2m_levels(leaf_pos, leaf_w, L):
    """check/core/solver/m2m.cu:32-48: parent = mass-weighted COM of the 4 children.

    leaf_pos [D0, D0, 2] indexed [Y, X]; returns per-level (pos, w) lists.
    A zero-mass parent gets COM (0, 0) (m2m.cu:48-50).
    """
    P = [np.asarray(leaf_pos, np.float64)]
    W = [np.asarray(leaf_w, np.float64)]
    for _ in range(1, max(L, 1)):
        p, w = P[-1], W[-1]
        mw = p * w[..., None]
        sw = w[0::2, 0::2] + w[0::2, 1::2] + w[1::2, 0::2] + w[1::2, 1::2]
        sp = mw[0::2, 0::2] + mw[0::2, 1::2] + mw[1::2, 0::2] + mw[1::2, 1::2]
        with np.errstate(invalid="ignore", divide="ignore"):
            com = np.where(sw[..., None] != 0, sp / np.where(sw == 0, 1, sw)[..., None], 0.0)
        P.append(com)
        W.append(sw)
    return P, W


def m2l_level_rule(pos, w, eps):
    """check/core/solver/m2l.cu:34-50 on one level: parents within one parent cell
    of each other and the cells themselves not adjacent (O(cells^2))."""
    D = pos.shape[0]
    ys, xs = np.meshgrid(np.arange(D), np.arange(D), indexing="ij")
    X = xs.ravel()
    Y = ys.ravel()
    PX, PY = X // 2, Y // 2
    P = pos.reshape(-1, 2)
    Wf = w.ravel()
    inter = (np.abs(PX[:, None] - PX[None, :]) <= 1) & (np.abs(PY[:, None] - PY[None, :]) <= 1) & \
            ((np.abs(X[:, None] - X[None, :]) > 1) | (np.abs(Y[:, None] - Y[None, :]) > 1))
    out = np.zeros((D * D, 2))
    for t in range(D * D):
        s = np.nonzero(inter[t])[0]
        if s.size:
            out[t] = kernel64(P[t][None, :], P[s], Wf[s], eps).sum(0)
    return out.reshape(D, D, 2)


def l2l_pushdown(effs):
    """check/core/solver/l2l.cu:26-37: from the top, add each cell's effect to its 4 children."""
    E = [np.array(e, np.float64) for e in effs]
    for l in range(len(E) - 1, 0, -1):
        E[l - 1] += np.repeat(np.repeat(E[l], 2, 0), 2, 1)
    return E


def fmm_numpy(L, pos0, w0, limits, eps):
    """Whole solve (solve.cu:11-27) from the test formulas, in float64, original order."""
    pos0 = np.asarray(pos0, np.float32).reshape(-1, 2)
    pre = predo_numpy(pos0, w0, limits, L)
    B = base_dim(L)
    pos = pre["position"].astype(np.float64)
    w = pre["weight"].astype(np.float64)
    eff = p2p_allpairs(pos, w, pre["index"], eps)
    if L > 0:
        lp, lw = p2m_leaf(pos, w, pre["head"], B)
        P, W = m2m_levels(lp.reshape(B, B, 2), lw.reshape(B, B), L)
        E = [m2l_level_rule(P[l], W[l], eps) for l in range(L)]
        E = l2l_pushdown(E)
        idx = pre["index"]
        eff = eff + E[0][idx[:, 1], idx[:, 0]]
    out = np.zeros_like(eff)
    out[pre["perm"]] = eff
    return out


def direct64(pos, w, eps, targets=None):
    pos = np.asarray(pos, np.float64).reshape(-1, 2)
    w = np.asarray(w, np.float64)
    t = np.arange(pos.shape[0]) if targets is None else np.asarray(targets)
    k = kernel64(pos[t, None, :], pos[None, :, :], w[None, :], eps)
    k[np.arange(t.size), t] = 0.0
    return k.sum(1)


def verlet_f32(tick, xcur, xpre, acc):
    """update.cu:39-41 evaluated op by op in float32: ((2x - xp) + (a*tick)*tick)."""
    x = np.asarray(xcur, np.float32)
    t = np.float32(tick)
    new = (np.float32(2) * x - np.asarray(xpre, np.float32)) + (np.asarray(acc, np.float32) * t) * t
    return new.astype(np.float32), x.copy()
