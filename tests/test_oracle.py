"""Pins the CPU oracle (oracle/nbfmm_oracle.cpp) before anything is checked against it.

Each stage is compared with the float64 restatement of the reference's own
unit-test formula (tests/npfmm.py, frozen into tests/golden/*.npz), and the
whole solve with closed forms and invariants that fix it independently:
L=0 equals the direct sum, the exact regime (at most one particle per cell on
every level) equals the direct sum, momentum is conserved, mass is conserved
up the tree, and the Verlet update matches a constant-acceleration closed form.
CPU only.
"""
import numpy as np
import pytest

import npfmm
from oracle import pyoracle as orc


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def max_rel_to_rms(a, b):
    a = np.asarray(a, np.float64).reshape(-1, 2)
    b = np.asarray(b, np.float64).reshape(-1, 2)
    rms = np.sqrt((b ** 2).sum(1).mean())
    return np.abs(a - b).max() / max(rms, 1e-300)


@pytest.fixture(scope="module")
def stage(golden):
    return golden("stage_L3_N250.npz")


def test_predo_matches_fixture(stage):
    L = int(stage["L"])
    out = orc.predo(L, stage["position"], stage["weight"], stage["limits"])
    assert out["clamped"] == 0
    np.testing.assert_array_equal(out["index"], stage["predo_index"])
    np.testing.assert_array_equal(out["perm"], stage["predo_perm"])
    np.testing.assert_array_equal(out["head"], stage["predo_head"])
    np.testing.assert_array_equal(out["position"], stage["predo_position"])
    np.testing.assert_array_equal(out["weight"], stage["predo_weight"])
    # check/core/solver/predo.cu:50-66: every particle lies in its cell box and
    # head ranges agree with index
    B = npfmm.base_dim(L)
    cw, ch = npfmm.cell_size_f32(stage["limits"], B)
    lim = stage["limits"]
    p, ix = out["position"], out["index"]
    assert np.all(p[:, 0] >= ix[:, 0] * cw + lim[0]) and np.all(p[:, 0] < (ix[:, 0] + 1) * cw + lim[0])
    assert np.all(p[:, 1] >= ix[:, 1] * ch + lim[1]) and np.all(p[:, 1] < (ix[:, 1] + 1) * ch + lim[1])
    h = out["head"]
    for c in range(B * B):
        assert np.all(ix[h[c]:h[c + 1], 0] == c % B) and np.all(ix[h[c]:h[c + 1], 1] == c // B)


def test_p2p_matches_fixture(stage):
    L = int(stage["L"])
    eff = orc.p2p(L, stage["position"], stage["weight"], stage["index"], stage["head"], float(stage["eps"]))
    np.testing.assert_allclose(eff, stage["p2p_effect"], rtol=1e-5, atol=1e-4)
    assert rel_l2(eff, stage["p2p_effect"]) < 1e-6


def test_p2m_matches_fixture(stage):
    L = int(stage["L"])
    B = npfmm.base_dim(L)
    cpos, cw = orc.p2m(L, stage["position"], stage["weight"], stage["head"])
    np.testing.assert_allclose(cpos[:B * B], stage["p2m_cell_position"], rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(cw[:B * B], stage["p2m_cell_weight"], rtol=1e-5, atol=1e-4)


def test_m2m_matches_fixture(stage):
    L = int(stage["L"])
    B = npfmm.base_dim(L)
    cpos = np.zeros_like(stage["cube_position"])
    cw = np.zeros_like(stage["cube_weight"])
    cpos[:B * B] = stage["cube_position"][:B * B]
    cw[:B * B] = stage["cube_weight"][:B * B]
    cpos, cw = orc.m2m(L, cpos, cw)
    for l in range(L):
        s = 1 << l
        got_p = cpos.reshape(L, B, B, 2)[l, ::s, ::s]
        got_w = cw.reshape(L, B, B)[l, ::s, ::s]
        exp_p = stage["m2m_cube_position"].reshape(L, B, B, 2)[l, ::s, ::s]
        exp_w = stage["m2m_cube_weight"].reshape(L, B, B)[l, ::s, ::s]
        np.testing.assert_allclose(got_p, exp_p, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(got_w, exp_w, rtol=1e-5, atol=1e-4)


def test_m2l_matches_fixture(stage):
    L = int(stage["L"])
    B = npfmm.base_dim(L)
    ceff = orc.m2l(L, float(stage["eps"]), stage["cube_position"], stage["cube_weight"])
    for l in range(L):
        s = 1 << l
        got = ceff.reshape(L, B, B, 2)[l, ::s, ::s]
        exp = stage["m2l_cube_effect"].reshape(L, B, B, 2)[l, ::s, ::s]
        np.testing.assert_allclose(got, exp, rtol=1e-5, atol=1e-4)


def test_l2l_matches_fixture(stage):
    L = int(stage["L"])
    B = npfmm.base_dim(L)
    ceff = orc.l2l(L, stage["cube_position"])
    np.testing.assert_allclose(ceff[:B * B], stage["l2l_leaf_effect"], rtol=1e-5, atol=1e-4)


def test_l2p_matches_fixture(stage):
    L = int(stage["L"])
    eff = orc.l2p(L, stage["index"], stage["cube_position"], stage["position"])
    np.testing.assert_allclose(eff, stage["l2p_effect"], rtol=1e-6, atol=1e-5)


def test_postdo_matches_fixture(stage):
    out = orc.postdo(stage["perm"], stage["position"])
    np.testing.assert_array_equal(out, stage["postdo_effect_origin"])


def test_solve_matches_fixture(golden):
    g = golden("solve_disk_L4_N2048.npz")
    eff, clamped = orc.solve(int(g["L"]), g["position"], g["weight"], g["limits"], float(g["eps"]))
    assert clamped == 0
    assert rel_l2(eff, g["fmm_effect"]) < 1e-5
    assert max_rel_to_rms(eff, g["fmm_effect"]) < 1e-4
    # the monopole-only scheme's own error against the direct sum (about 0.16
    # at this shape): reported, loosely bounded
    err = rel_l2(eff, g["direct_effect"])
    assert 0.01 < err < 0.5


@pytest.mark.parametrize("L,n", [(0, 300), (1, 513), (2, 1024), (3, 700), (5, 2048)])
def test_solve_matches_numpy_fmm_random(L, n):
    rng = np.random.default_rng(100 + L)
    limits = np.array([-2.0, -1.0, 3.0, 2.5], np.float32)
    pos = np.stack([rng.uniform(-2, 3, n), rng.uniform(-1, 2.5, n)], 1).astype(np.float32)
    w = rng.exponential(1.0, n).astype(np.float32)
    eps = 1e-2
    eff, _ = orc.solve(L, pos, w, limits, eps)
    ref = npfmm.fmm_numpy(L, pos, w, limits, eps)
    assert rel_l2(eff, ref) < 1e-5
    assert max_rel_to_rms(eff, ref) < 1e-4


def test_level0_is_direct_sum():
    rng = np.random.default_rng(3)
    n = 400
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    lim = np.array([-1.1, -1.1, 1.1, 1.1], np.float32)
    eff64 = orc.solve_f64(0, pos.astype(np.float64), w.astype(np.float64), lim.astype(np.float64), 1e-3)
    d = orc.direct_f64(pos, w, 1e-3)
    assert rel_l2(eff64, d) < 1e-13
    eff, _ = orc.solve(0, pos, w, lim, 1e-3)
    assert rel_l2(eff, d) < 1e-5


@pytest.mark.parametrize("L", [1, 2, 3, 4, 6])
def test_exact_regime_equals_direct(L):
    """At most one particle per cell on every level: COMs are the particles
    themselves, so the FMM must reproduce every pair of the direct sum exactly
    once (a missing or doubled interaction shows here)."""
    rng = np.random.default_rng(50 + L)
    lim = np.array([-1.0, -1.0, 1.0, 1.0], np.float64)
    # one particle in each top-level (4x4) cell, placed away from cell edges
    top = 2.0 / 4
    cells = [(i, j) for i in range(4) for j in range(4)]
    pos = np.array([[lim[0] + (i + rng.uniform(0.2, 0.8)) * top, lim[1] + (j + rng.uniform(0.2, 0.8)) * top]
                    for i, j in cells])
    w = rng.uniform(0.5, 1.5, len(cells))
    eff = orc.solve_f64(L, pos, w, lim, 1e-3)
    d = orc.direct_f64(pos.astype(np.float32), w.astype(np.float32), 1e-3)
    # the fp64 oracle is fed the float32-rounded positions the direct sum sees
    eff32pos = orc.solve_f64(L, pos.astype(np.float32).astype(np.float64), w.astype(np.float32).astype(np.float64),
                             lim, 1e-3)
    assert rel_l2(eff32pos, d) < 1e-12
    assert rel_l2(eff, npfmm.direct64(pos, w, 1e-3)) < 1e-12


def test_one_particle_per_leaf_is_not_exact():
    """The converse of the exact regime: one particle per leaf only is not
    exact for L >= 2 (coarse cells hold several particles)."""
    rng = np.random.default_rng(9)
    L = 2
    B = npfmm.base_dim(L)
    lim = np.array([0.0, 0.0, 1.0, 1.0])
    pos = np.array([[(i + rng.uniform(0.2, 0.8)) / B, (j + rng.uniform(0.2, 0.8)) / B] for i in range(B) for j in range(B)])
    w = np.ones(B * B)
    eff = orc.solve_f64(L, pos, w, lim, 1e-3)
    assert rel_l2(eff, npfmm.direct64(pos, w, 1e-3)) > 1e-3


@pytest.mark.parametrize("L", [2, 4, 5])
def test_momentum_conserved(L):
    rng = np.random.default_rng(11 + L)
    n = 3000
    pos = rng.normal(0, 0.4, (n, 2)).clip(-0.99, 0.99)
    w = rng.uniform(0.5, 1.5, n)
    lim = np.array([-1.0, -1.0, 1.0, 1.0])
    eff = orc.solve_f64(L, pos, w, lim, 1e-3)
    p = (w[:, None] * eff).sum(0)
    scale = (w[:, None] * np.abs(eff)).sum(0)
    assert np.all(np.abs(p) < 1e-12 * scale)


def test_mass_conserved_up_the_tree():
    rng = np.random.default_rng(12)
    L, n = 5, 5000
    B = npfmm.base_dim(L)
    lim = np.array([-1, -1, 1, 1], np.float32)
    pos = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    pre = orc.predo(L, pos, w, lim)
    cpos, cw = orc.p2m(L, pre["position"], pre["weight"], pre["head"])
    cpos, cw = orc.m2m(L, cpos, cw)
    total = w.astype(np.float64).sum()
    com = (pos.astype(np.float64) * w[:, None]).sum(0) / total
    for l in range(L):
        s = 1 << l
        lw = cw.reshape(L, B, B)[l, ::s, ::s].astype(np.float64)
        lp = cpos.reshape(L, B, B, 2)[l, ::s, ::s].astype(np.float64)
        assert abs(lw.sum() - total) < 1e-4 * total
        assert np.allclose((lp * lw[..., None]).sum((0, 1)) / lw.sum(), com, atol=1e-5)


def test_compute_index_bitexact_and_clamp():
    rng = np.random.default_rng(5)
    L = 8
    lim = np.array([-4, -4, 4, 4], np.float32)
    B = npfmm.base_dim(L)
    cw, _ = npfmm.cell_size_f32(lim, B)
    # points on and next to cell edges, where rounding decides the cell
    edges = (lim[0] + np.arange(B, dtype=np.float32) * cw).astype(np.float32)
    xs = np.concatenate([edges, np.nextafter(edges, np.float32(-10)), np.nextafter(edges, np.float32(10)),
                         rng.uniform(-4, 4, 5000).astype(np.float32)])
    xs = xs[(xs >= lim[0]) & (xs < lim[2])]
    pos = np.stack([xs, xs[::-1]], 1).astype(np.float32)
    idx, clamped = orc.compute_index(L, pos, lim)
    ref = npfmm.compute_index_f32(pos, lim, L)
    inbox = (ref >= 0).all(1) & (ref < B).all(1)
    np.testing.assert_array_equal(idx[inbox], ref[inbox])
    assert clamped == int((~inbox).sum())
    # out-of-box and NaN inputs are clamped and counted
    bad = np.array([[4.0, 0.0], [-4.5, 0.0], [0.0, np.nan], [np.inf, 1.0]], np.float32)
    idx, clamped = orc.compute_index(L, bad, lim)
    assert clamped == 4
    assert idx.min() >= 0 and idx.max() <= B - 1
    assert tuple(idx[0]) == (B - 1, B // 2)


def test_two_body_closed_form():
    eps = 1e-3
    pos = np.array([[0.0, 0.0], [0.3, 0.4]], np.float32)
    w = np.array([2.0, 3.0], np.float32)
    lim = np.array([-1, -1, 1, 1], np.float32)
    eff, _ = orc.solve(3, pos, w, lim, eps)
    r = 0.5
    f = 1.0 / (r ** 3 + eps)
    np.testing.assert_allclose(eff[0], 3.0 * f * np.array([0.3, 0.4]), rtol=1e-6)
    np.testing.assert_allclose(eff[1], -2.0 * f * np.array([0.3, 0.4]), rtol=1e-6)


def test_verlet_matches_fixture(golden):
    g = golden("verlet.npz")
    cur, pre = orc.update(float(g["tick"]), g["position_cur"], g["position_pre"], g["acceleration"])
    np.testing.assert_array_equal(cur, g["expect_cur"])
    np.testing.assert_array_equal(pre, g["expect_pre"])


def test_verlet_constant_acceleration_closed_form():
    dt = 1e-2
    x0 = np.array([[0.1, -0.2]])
    v = np.array([[0.5, 0.25]])
    a = np.array([[-1.0, 2.0]])
    # x_{-1} chosen so the scheme is exact for constant acceleration
    xcur = x0.copy()
    xpre = x0 - v * dt + 0.5 * a * dt * dt
    for _ in range(50):
        xcur, xpre = orc.update_f64(dt, xcur, xpre, a)
    n = 50
    exact = x0 + n * v * dt + 0.5 * (n * dt) ** 2 * a
    np.testing.assert_allclose(xcur, exact, rtol=0, atol=1e-13)


def test_prune_is_stable_bool_partition():
    """Stars::prune (prune.cu:28-62): flags x < x_min | y < y_min | x > x_max |
    y > y_max; kept stars first, both groups in their previous order (what a
    radix sort of the bool keys gives: numpy's stable argsort)."""
    rng = np.random.default_rng(31)
    n = 5000
    lim = np.array([-1.0, -2.0, 3.0, 1.5], np.float32)
    xc = rng.uniform(-2, 4, (n, 2)).astype(np.float32)
    xp = rng.standard_normal((n, 2)).astype(np.float32)
    w = rng.random(n).astype(np.float32)
    # the box edges themselves stay; NaN compares false and stays
    xc[0] = (3.0, 1.5)
    xc[1] = (-1.0, -2.0)
    xc[2] = (np.nextafter(np.float32(3.0), np.float32(4.0)), 0.0)
    xc[3] = (0.0, np.nextafter(np.float32(-2.0), np.float32(-3.0)))
    xc[4] = (np.nan, 0.0)
    flags = (xc[:, 0] < lim[0]) | (xc[:, 1] < lim[1]) | (xc[:, 0] > lim[2]) | (xc[:, 1] > lim[3])
    order = np.argsort(flags, kind="stable")
    c2, p2, w2, kept = orc.prune(lim, xc, xp, w)
    assert kept == int((~flags).sum())
    assert flags[2] and flags[3] and not flags[0] and not flags[1] and not flags[4]
    np.testing.assert_array_equal(c2, xc[order])
    np.testing.assert_array_equal(p2, xp[order])
    np.testing.assert_array_equal(w2, w[order])


def test_prune_nothing_and_everything():
    lim = np.array([0, 0, 1, 1], np.float32)
    xc = np.array([[0.5, 0.5], [0.1, 0.9]], np.float32)
    c2, _, _, kept = orc.prune(lim, xc, xc, np.ones(2, np.float32))
    assert kept == 2 and np.array_equal(c2, xc)
    c2, _, _, kept = orc.prune(lim, xc + 5, xc, np.ones(2, np.float32))
    assert kept == 0 and np.array_equal(c2, xc + 5)


# ---- initial-condition models (source/nbfmm/model/*.cu) ----

def test_model_uniform_draws():
    u = np.array([orc.model_uniform(s, i, k) for s in (0, 5) for i in range(2000) for k in range(3)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    assert len(np.unique(u)) > 0.99 * u.size


@pytest.mark.parametrize("centered", [False, True])
def test_model_disk_first_step_stays_on_circle(centered):
    """disk.cu:45-55: radius (i+1)/n * R, previous position rotated back by
    the chord angle acos(1 - a tau^2 / 2r).  That angle makes the first Verlet
    step under the central pull a (kernel of the enclosed mass) land back on
    the circle exactly: |2x - x_prev - a tau^2 x/r| = r."""
    n, R, wt, tau, eps, cw = 3000, 2.0, 0.01, 0.02, 0.1, 5.0
    c = np.array([0.3, -0.7])
    xc, xp, w = orc.model_disk(n, c, R, wt, tau, eps, seed=3, center_weight=cw, centered=centered)
    i = np.arange(n)
    r32 = ((i + 1).astype(np.float32) / np.float32(n)) * np.float32(R)
    start = 1 if centered else 0
    d = xc.astype(np.float64) - c
    dp = xp.astype(np.float64) - c
    rr = np.sqrt((d ** 2).sum(1))
    # positions are c + r (cos, sin) in float32: a few ulps of |c| absolute
    np.testing.assert_allclose(rr[start:], r32[start:], rtol=2e-6, atol=3e-7)
    np.testing.assert_allclose(np.sqrt((dp ** 2).sum(1))[start:], r32[start:], rtol=2e-6, atol=3e-7)
    m = (wt * (i - 1) + cw) if centered else wt * i
    r = r32.astype(np.float64)
    a = r * m / (r ** 3 + eps)
    x2 = 2 * d - dp - (a * tau * tau / r)[:, None] * d
    dev = np.abs(np.sqrt((x2 ** 2).sum(1)) - r)
    assert np.all(dev[start:] < 1e-6 + 2e-5 * r[start:])
    assert np.all(w[start:] == np.float32(wt))
    if centered:
        assert np.all(xc[0] == c.astype(np.float32)) and w[0] == np.float32(cw)


def test_model_disk_static_and_circles():
    c = (1.0, 2.0)
    xc, xp, w = orc.model_disk(500, c, 1.5, 0.2, 0.01, rotate=False, seed=1)
    assert np.array_equal(xc, xp)
    n, R, tau, eps, cw = 256, 1.5, 0.01, 0.1, 10.0
    xc, xp, w = orc.model_circle(n, c, R, 0.1, cw, tau, eps, uniform=True)
    ang = np.arctan2(xc[1:, 1] - c[1], xc[1:, 0] - c[0]) % (2 * np.pi)
    np.testing.assert_allclose(ang, (2 * np.pi * np.arange(1, n) / n) % (2 * np.pi), atol=2e-6)
    # every star rotated back by the one chord angle of the centre star's pull
    a = R * cw / (R ** 3 + eps)
    dang = np.arccos(1 - a * tau * tau / R / 2)
    angp = np.arctan2(xp[1:, 1] - c[1], xp[1:, 0] - c[0]) % (2 * np.pi)
    np.testing.assert_allclose((ang - angp) % (2 * np.pi), dang, atol=3e-6)
    assert w[0] == np.float32(cw) and np.all(w[1:] == np.float32(0.1))


def test_model_rectangle():
    xc, xp, w = orc.model_rectangle(4000, (1.0, -1.0), 3.0, 2.0, 0.5, seed=9)
    assert np.array_equal(xc, xp)
    assert xc[:, 0].min() >= -0.5 and xc[:, 0].max() <= 2.5
    assert xc[:, 1].min() >= -2.0 and xc[:, 1].max() <= 0.0
    assert w.min() > 0 and w.max() <= 0.5


@pytest.mark.parametrize("ecc", [0.0, 0.5])
def test_double_disk_offsets_are_the_orbit_step(ecc):
    """double_disk.cu:61-86: each disk's previous positions shift by the step
    of a circular orbit about the common centre of mass, perpendicular to the
    other disk's pull a_k (|o| = sqrt(r_k a_k) tau), scaled by 2^-e, and
    pulled back by a_k tau^2 e."""
    c1, c2 = np.array([-1.0, 0.5]), np.array([1.5, -0.25])
    m1, m2, tau, eps = 3.0, 1.0, 0.01, 0.1
    o1, o2 = orc.double_disk_offsets(c1, c2, m1, m2, m1, m2, ecc, tau, eps)
    d = c2 - c1
    r = np.linalg.norm(d)
    for o, m_other, sign, rk in ((o1, m2, 1, r * m2 / (m1 + m2)), (o2, m1, -1, r * m1 / (m1 + m2))):
        e = sign * d * m_other / (r ** 3 + eps)
        a = np.linalg.norm(e)
        perp = np.array([-e[1], e[0]]) / a
        expect = perp * np.sqrt(rk * a) * tau / 2 ** ecc - e * tau * tau * ecc
        np.testing.assert_allclose(o, expect, rtol=1e-5, atol=1e-9)


# ---- video frames (source/nbfmm/display/stars/display.cu) ----

# The offsets (dx, dy) each size level adds, read off display.cu:54-211
# (board[dx + dy * width]; dy > 0 is down).
STAMP_LISTING = {
    0: [(0, 0)],
    1: [(-1, 0), (1, 0), (0, -1), (0, 1)],
    2: [(-1, -1), (-1, 1), (1, -1), (1, 1)],
    3: [(-2, 0), (2, 0), (0, -2), (0, 2)],
    4: [(-2, -1), (-1, -2), (-2, 1), (-1, 2), (2, -1), (1, -2), (2, 1), (1, 2)],
    5: [(-3, 0), (3, 0), (0, -3), (0, 3), (-2, -2), (-2, 2), (2, -2), (2, 2)],
    6: [(-3, -1), (-1, -3), (-3, 1), (-1, 3), (3, -1), (1, -3), (3, 1), (1, 3)],
    7: [(-4, 0), (4, 0), (0, -4), (0, 4), (-3, -2), (-2, -3), (-3, 2), (-2, 3), (3, -2), (2, -3), (3, 2), (2, 3)],
}


def expected_luma(width, height, stars):
    """stars: (x, y, level) pixel centres; the listing's offsets, clipped."""
    img = np.zeros((height, width), np.uint8)
    for x, y, level in stars:
        for s in range(min(level, 7) + 1):
            for dx, dy in STAMP_LISTING[s]:
                if 0 <= x + dx < width and 0 <= y + dy < height:
                    img[y + dy, x + dx] = 255
    return img


def pixel_centre(x, y, width, height, lim):
    uw = (lim[2] - lim[0]) / width
    uh = (lim[3] - lim[1]) / height
    return lim[0] + (x + 0.5) * uw, lim[3] - (y + 0.5) * uh


@pytest.mark.parametrize("level", range(9))
def test_display_stamp_is_the_reference_listing(level):
    W, H, lim = 20, 14, (-5.0, -3.5, 5.0, 3.5)
    px, py = pixel_centre(9, 6, W, H, lim)
    wt = float(level) ** 2 * 0.25 + 0.01   # sqrt(w) / 0.5 = level + a little
    f = orc.display(W, H, 0.5, lim, np.array([[px, py]], np.float32), np.array([wt], np.float32))
    assert f.size == W * H * 3 // 2
    assert np.array_equal(f[:W * H].reshape(H, W), expected_luma(W, H, [(9, 6, level)]))
    assert np.all(f[W * H:] == 128)
    assert (f[:W * H] == 255).sum() == [1, 5, 9, 13, 21, 29, 37, 49, 49][level]


def test_display_clips_at_edges_and_corners():
    W, H, lim = 24, 16, (0.0, 0.0, 24.0, 16.0)
    rng = np.random.default_rng(0)
    for x, y in [(0, 0), (W - 1, 0), (0, H - 1), (W - 1, H - 1), (1, 1), (W - 2, H - 2), (2, 7), (W - 3, 8),
                 (11, 0), (11, 1), (11, H - 1), (3, H - 3)]:
        level = int(rng.integers(0, 8))
        px, py = pixel_centre(x, y, W, H, lim)
        f = orc.display(W, H, 1.0, lim, np.array([[px, py]], np.float32), np.array([level ** 2 + 0.5], np.float32))
        assert np.array_equal(f[:W * H].reshape(H, W), expected_luma(W, H, [(x, y, level)])), (x, y, level)


def test_display_pixel_rule():
    """x = floor((p.x - x_min) / unit), y = floor((y_max - p.y) / unit) with
    unit 1 here: cell edges belong to the cell right of / below them, the top
    and left edges of the box are in, the right and bottom edges are out,
    and NaN positions are skipped."""
    W, H, lim = 16, 12, (0.0, 0.0, 16.0, 12.0)
    cases = [((3.0, 5.0), (3, 7)), ((0.0, 12.0), (0, 0)), ((15.999, 0.001), (15, 11)), ((7.5, 6.0), (7, 6))]
    for (px, py), (x, y) in cases:
        f = orc.display(W, H, 1.0, lim, np.array([[px, py]], np.float32), np.array([0.5], np.float32))
        lit = np.argwhere(f[:W * H].reshape(H, W) == 255)
        assert lit.tolist() == [[y, x]], (px, py)
    for px, py in [(16.0, 5.0), (5.0, 0.0), (-1e-6, 5.0), (5.0, 12.0001), (np.nan, 3.0), (3.0, np.nan), (1e30, 1.0)]:
        f = orc.display(W, H, 1.0, lim, np.array([[px, py]], np.float32), np.array([4.0], np.float32))
        assert not np.any(f[:W * H]), (px, py)
    # a NaN weight lights the point only
    f = orc.display(W, H, 1.0, lim, np.array([[3.5, 3.5]], np.float32), np.array([np.nan], np.float32))
    assert (f[:W * H] == 255).sum() == 1


def test_display_is_a_union_of_stamps():
    """Stars light pixels independently: the frame of many stars is the
    pixelwise maximum of their single-star frames, whatever their order."""
    W, H, lim = 40, 30, (-2.0, -1.5, 2.0, 1.5)
    rng = np.random.default_rng(3)
    pos = rng.uniform(-2.2, 2.2, (60, 2)).astype(np.float32)
    w = rng.uniform(0.0, 80.0, 60).astype(np.float32)
    full = orc.display(W, H, 1.0, lim, pos, w)
    acc = np.zeros(W * H, np.uint8)
    for i in range(60):
        acc = np.maximum(acc, orc.display(W, H, 1.0, lim, pos[i:i + 1], w[i:i + 1])[:W * H])
    assert np.array_equal(full[:W * H], acc)
    perm = rng.permutation(60)
    assert np.array_equal(orc.display(W, H, 1.0, lim, pos[perm], w[perm]), full)


def test_run_f64_cells_record_replay_and_smoothness():
    """Frozen-cell fp64 trajectories (for finite differences of a cells-fixed
    adjoint): recording gives run_f64's bits, replaying the recorded cells
    gives them again, and a perturbation that changes cells in the free run
    is smooth in the frozen one (central differences at h = 1e-3 and 1e-6
    agree), while the free run's differences jump."""
    rng = np.random.default_rng(12)
    n, L, steps = 200, 5, 20
    lim = np.array([-1, -1, 1, 1], np.float32)
    x = rng.uniform(-0.8, 0.8, (n, 2))
    v = rng.normal(0, 30.0, (n, 2))          # about a third of a leaf per step
    p = x - v * 1e-4
    w = rng.uniform(0.5, 1.5, n)
    tick = 1e-4 * np.sqrt(10.0)
    xa, pa = orc.run_f64(L, lim, 1e-3, tick, steps, x, p, w)
    xb, pb, cells = orc.run_f64_cells(L, lim, 1e-3, tick, steps, x, p, w)
    assert np.array_equal(xa, xb) and np.array_equal(pa, pb)
    xc, pc, _ = orc.run_f64_cells(L, lim, 1e-3, tick, steps, x, p, w, cells)
    assert np.array_equal(xa, xc) and np.array_equal(pa, pc)
    c = rng.normal(0, 1, (n, 2))

    def loss(xx, frozen):
        if frozen:
            out, _, _ = orc.run_f64_cells(L, lim, 1e-3, tick, steps, xx, xx - v * 1e-4, w, cells)
        else:
            out, _ = orc.run_f64(L, lim, 1e-3, tick, steps, xx, xx - v * 1e-4, w)
        return float((c * out).sum())

    def fd(i, h, frozen):
        xp, xm = x.copy(), x.copy()
        xp[i, 0] += h
        xm[i, 0] -= h
        return (loss(xp, frozen) - loss(xm, frozen)) / (2 * h)

    jumps = 0
    for i in range(40):
        a, b = fd(i, 1e-3, True), fd(i, 1e-6, True)
        assert abs(a - b) <= 1e-4 * max(abs(b), 1.0), (i, a, b)  # O(h^2) truncation only
        fa, fb = fd(i, 1e-3, False), fd(i, 1e-6, False)
        jumps += abs(fa - fb) > 1e-4 * max(abs(fb), 1.0)
    assert jumps > 0  # the free runs do change cells at h = 1e-3
