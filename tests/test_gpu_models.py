"""Initial-condition models on the GPU (source/nbfmm/model/*.cu) against the
oracle's restatement: the same counter-based draws, radii, weights and
orbit offsets; positions agree to the last bits of the device cosf/sinf
against the host's (1e-6 of the radius), weights exactly."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def host(t):
    return t.cpu().numpy()


def close(got, ref, scale):
    np.testing.assert_allclose(host(got), ref, rtol=0, atol=2e-6 * scale + 1e-7)


def test_disk_models():
    from nbfmm_amd import models
    n, c, R, wt, tau, eps = 20000, (0.5, -0.25), 2.0, 1e-4, 0.01, 0.1
    for kind in ("disk", "static", "center"):
        if kind == "disk":
            g = models.disk(n, c, R, wt, tau, eps, seed=4)
            r = orc.model_disk(n, c, R, wt, tau, eps, seed=4)
        elif kind == "static":
            g = models.disk_static(n, c, R, wt, seed=4)
            r = orc.model_disk(n, c, R, wt, 0.0, 1.0, seed=4, rotate=False)
        else:
            g = models.disk_center(n, c, R, wt, 3.0, tau, eps, seed=4)
            r = orc.model_disk(n, c, R, wt, tau, eps, seed=4, center_weight=3.0, centered=True)
        close(g[0], r[0], R)
        close(g[1], r[1], R)
        assert np.array_equal(host(g[2]), r[2]), kind


@pytest.mark.parametrize("uniform", [False, True])
def test_circle_models(uniform):
    from nbfmm_amd import models
    n, c, R, tau, eps = 4096, (0.0, 1.0), 1.5, 0.02, 0.1
    g = models.circle_center(n, c, R, 0.01, 20.0, tau, eps, uniform=uniform, seed=7)
    r = orc.model_circle(n, c, R, 0.01, 20.0, tau, eps, uniform=uniform, seed=7)
    close(g[0], r[0], R)
    close(g[1], r[1], R)
    assert np.array_equal(host(g[2]), r[2])


def test_rectangle_model():
    from nbfmm_amd import models
    g = models.rectangle(10000, (1.0, 2.0), 3.0, 1.0, 0.5, seed=11)
    r = orc.model_rectangle(10000, (1.0, 2.0), 3.0, 1.0, 0.5, seed=11)
    # no transcendental: the same float operations on the same draws
    for a, b in zip(g, r):
        assert np.array_equal(host(a), b)


@pytest.mark.parametrize("centered", [False, True])
def test_double_disk_models(centered):
    from nbfmm_amd import models
    n1, n2 = 6000, 3000
    c1, c2 = (-2.0, -1.0), (2.0, 1.0)
    args = dict(eps=0.1, seed=2, centered=centered)
    if centered:
        g = models.double_disk(n1, n2, c1, c2, 1.0, 0.7, 1e-4, 0.3, 0.01, center_weight1=2.0, center_weight2=1.0,
                               **args)
        r = orc.model_double_disk(n1, n2, c1, c2, 1.0, 0.7, 1e-4, 0.3, 0.01, cw1=2.0, cw2=1.0, **args)
    else:
        g = models.double_disk(n1, n2, c1, c2, 1.0, 0.7, 1e-4, 0.3, 0.01, **args)
        r = orc.model_double_disk(n1, n2, c1, c2, 1.0, 0.7, 1e-4, 0.3, 0.01, **args)
    close(g[0], r[0], 3.0)
    close(g[1], r[1], 3.0)
    assert np.array_equal(host(g[2]), r[2])


def test_simulate_from_model():
    """A centred disk from the model, stepped by the solver: no star lost,
    finite forces, and the first step matches the oracle's step."""
    from nbfmm_amd import Solver, models
    n, L = 30000, 5
    lim = np.array([-2, -2, 2, 2], np.float32)
    tau = np.float32(0.01)
    X, P, W = models.disk_center(n, (0.0, 0.0), 1.0, 1.0 / n, 1.0, tau, 0.1, seed=5)
    xc0, xp0, w0 = host(X), host(P), host(W)
    s = Solver(L, n, lim, 0.1)
    A = torch.zeros_like(X)
    s.step(n, float(tau), X, P, W, A)
    torch.cuda.synchronize()
    rc, rp, ra, _ = orc.run(L, lim, 0.1, float(tau), 1, xc0, xp0, w0)
    np.testing.assert_allclose(host(X), rc, rtol=0, atol=1e-6)
    assert np.array_equal(host(P), rp)
    assert s.prune(n, X, P, W) == n
    s.close()
