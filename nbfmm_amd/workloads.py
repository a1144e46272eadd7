"""Seeded synthetic initial conditions for the benchmark configurations.

The reference's own scenario generators (source/nbfmm/model/*.cu) draw from
cuRAND streams that cannot be reproduced here, so the configurations below
are restated with numpy's PCG64 generator and a fixed seed per
configuration.  Velocities are converted to the reference's two-position
Verlet state with x_prev = x - v*dt, and the Verlet step size is
tick = dt*sqrt(G) (source/nbfmm/display/stars.cu:31).

`levels` below is the reference's num_level: a workload quoted with
"k levels (2^k x 2^k leaf grid)" uses num_level = k - 1
(base_dim = 2^(num_level+1), source/nbfmm/core/solver.cu:19).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    position: np.ndarray      # float32 [N, 2]
    position_pre: np.ndarray  # float32 [N, 2]
    weight: np.ndarray        # float32 [N]
    num_level: int
    limits: tuple
    eps: float
    dt: float
    G: float
    steps: int
    notes: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.position.shape[0])

    @property
    def tick(self) -> float:
        return float(np.float32(self.dt) * np.float32(np.sqrt(self.G)))


def uniform_disk(rng, n, radius=1.0, center=(0.0, 0.0)):
    r = radius * np.sqrt(rng.uniform(0.0, 1.0, n))
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    return np.stack([center[0] + r * np.cos(th), center[1] + r * np.sin(th)], 1)


def gaussian_disk(rng, n, sigma, rmax, center=(0.0, 0.0)):
    """Isotropic 2D Gaussian of width sigma, radius truncated at rmax (resampled)."""
    out = np.empty((0, 2))
    while out.shape[0] < n:
        p = rng.normal(0.0, sigma, (2 * (n - out.shape[0]) + 16, 2))
        p = p[np.hypot(p[:, 0], p[:, 1]) < rmax]
        out = np.concatenate([out, p])
    return out[:n] + np.asarray(center)


def circular_velocity(pos, mass, center, G):
    """Counter-clockwise speed sqrt(G*M_enclosed(r)/r) about `center`."""
    d = pos - np.asarray(center)
    r = np.hypot(d[:, 0], d[:, 1])
    order = np.argsort(r, kind="stable")
    menc = np.empty_like(r)
    menc[order] = np.cumsum(mass[order]) - mass[order]
    v = np.sqrt(G * menc / np.maximum(r, 1e-12))
    t = np.stack([-d[:, 1], d[:, 0]], 1) / np.maximum(r, 1e-12)[:, None]
    return t * v[:, None]


def _finish(name, pos, vel, w, L, limits, eps, dt, G, steps, **notes):
    pos = pos.astype(np.float32)
    pre = (pos.astype(np.float64) - vel * dt).astype(np.float32)
    return Workload(name, pos, pre, w.astype(np.float32), L, tuple(float(x) for x in limits), float(eps),
                    float(dt), float(G), int(steps), dict(notes))


def disk(n=16384, seed=1001, steps=10):
    """Plain FMM single system: uniform disk of radius 1 at the origin, masses
    U[0.5,1.5], circular velocities, G=1, eps=1e-3, 5 levels (32x32 leaves)
    over [-1.1,1.1]^2, dt=1e-4."""
    rng = np.random.default_rng(seed)
    pos = uniform_disk(rng, n, 1.0)
    w = rng.uniform(0.5, 1.5, n)
    vel = circular_velocity(pos, w, (0.0, 0.0), 1.0)
    return _finish("disk", pos, vel, w, 4, (-1.1, -1.1, 1.1, 1.1), 1e-3, 1e-4, 1.0, steps)


def two_disk_collision(n=1 << 20, seed=1002, steps=100, box=4.0, num_level=8, G=None):
    """Two Gaussian disks (sigma 0.5, radius truncated at box/2 - 0.1) centred
    at (+-box/2, 0) with bulk velocities (0, +-0.3) plus circular rotation;
    masses Exp(1) clamped to [0.1, 5]; eps=1e-3; 9 levels (512x512 leaves)
    over [-box, box]^2; dt=1e-4.

    G is not specified for this workload.  With G=1 the 2^20 unit masses
    give circular speeds near 1e3 and the disks leave the box within a few
    dozen steps, so G defaults to 1/N (total mass about 1 in units of G)."""
    rng = np.random.default_rng(seed)
    if G is None:
        G = 1.0 / n
    half = n // 2
    parts, vels, ws = [], [], []
    for k, (cx, vy, m) in enumerate([(-box / 2, 0.3, half), (box / 2, -0.3, n - half)]):
        p = gaussian_disk(rng, m, 0.5, box / 2 - 0.1, (cx, 0.0))
        w = np.clip(rng.exponential(1.0, m), 0.1, 5.0)
        v = circular_velocity(p, w, (cx, 0.0), G) + np.array([0.0, vy])
        parts.append(p)
        vels.append(v)
        ws.append(w)
    pos = np.concatenate(parts)
    vel = np.concatenate(vels)
    w = np.concatenate(ws)
    return _finish("two_disk_collision", pos, vel, w, num_level, (-box, -box, box, box), 1e-3, 1e-4, G, steps,
                   G_per_N=G * n)


# G of the mass-inference workloads, in units of 1/N.  A trajectory depends
# on the other particles' masses through tick^2 * a, and with G = 1/N (total
# mass about 1) a 20 % mass perturbation moves the positions of a two-disk
# collision of 65536 particles by 7.8e-8 RMS after 100 steps: below one fp32
# ulp at |x| ~ 2 (2.4e-7), so the fit's loss would be rounding noise.  With
# G = 1e4 / N the same perturbation moves them by 2.0e-3 RMS (4.1e-4 after
# 50 steps), 40 times the difference between the fp32 and fp64 trajectories
# of the same masses (5.4e-5), while the particles stay inside the box.
FIT_G_PER_N = 1e4


def gaussian_disk_system(n=65536, seed=1003, steps=50, num_level=6, G=None):
    """Single Gaussian disk (sigma 0.5, truncated at 1.9), masses U[0.5,1.5],
    7 levels over [-2.09, 2.09]^2 (1.1 x the truncation radius), dt=1e-4.
    The mass-inference workload: G = FIT_G_PER_N / N unless given."""
    rng = np.random.default_rng(seed)
    if G is None:
        G = FIT_G_PER_N / n
    pos = gaussian_disk(rng, n, 0.5, 1.9)
    w = rng.uniform(0.5, 1.5, n)
    vel = circular_velocity(pos, w, (0.0, 0.0), G)
    R = 1.1 * 1.9
    return _finish("gaussian_disk", pos, vel, w, num_level, (-R, -R, R, R), 1e-3, 1e-4, G, steps,
                   G_choice=f"{FIT_G_PER_N:g}/N")


@dataclass
class BatchWorkload:
    name: str
    systems: list          # list of Workload, one per system
    steps: int

    @property
    def n(self) -> int:
        return int(sum(w.n for w in self.systems))

    def concat(self):
        """Concatenated position, position_pre, weight (system order)."""
        return (np.concatenate([w.position for w in self.systems]),
                np.concatenate([w.position_pre for w in self.systems]),
                np.concatenate([w.weight for w in self.systems]))


def many_disks(num_systems=256, n_min=1024, n_max=131072, seed=1004, steps=20):
    """Independent systems: N log-uniform in [n_min, n_max]; a uniform or a
    Gaussian disk (coin flip) of radius R log-uniform in [0.5, 4] (Gaussian:
    sigma R/2, truncated at R); levels = clamp(round(log4(N/16)), 3, 9), i.e.
    num_level = levels - 1; masses U[0.5,1.5]; eps 1e-3; dt 1e-4; box
    [-1.1R, 1.1R]^2.  G is not specified for this workload; it is 1/N per
    system, as in the two-disk collision."""
    rng = np.random.default_rng(seed)
    systems = []
    for k in range(num_systems):
        n = int(np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        R = float(np.exp(rng.uniform(np.log(0.5), np.log(4.0))))
        gaussian = bool(rng.integers(0, 2))
        levels = int(np.clip(np.round(np.log(n / 16.0) / np.log(4.0)), 3, 9))
        pos = gaussian_disk(rng, n, R / 2, R) if gaussian else uniform_disk(rng, n, R)
        w = rng.uniform(0.5, 1.5, n)
        G = 1.0 / n
        vel = circular_velocity(pos, w, (0.0, 0.0), G)
        box = 1.1 * R
        systems.append(_finish(f"sys{k}", pos, vel, w, levels - 1, (-box, -box, box, box), 1e-3, 1e-4, G, steps,
                               gaussian=gaussian, R=R))
    return BatchWorkload("many_disks", systems, steps)


def many_collisions(num_systems=64, n_min=4096, n_max=262144, seed=1005, steps=100, g_per_n=FIT_G_PER_N):
    """Independent two-disk collisions (the two-disk geometry with N_s
    particles): N log-uniform in [n_min, n_max], levels as in many_disks.
    The batched mass-inference workload: G = g_per_n / N per system
    (FIT_G_PER_N)."""
    rng = np.random.default_rng(seed)
    systems = []
    for k in range(num_systems):
        n = int(np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        levels = int(np.clip(np.round(np.log(n / 16.0) / np.log(4.0)), 3, 9))
        w = two_disk_collision(n=n, seed=int(rng.integers(1 << 30)), steps=steps, num_level=levels - 1,
                               G=g_per_n / n)
        w.name = f"sys{k}"
        systems.append(w)
    return BatchWorkload("many_collisions", systems, steps)


def by_name(name: str, **kw) -> Workload:
    return {"disk": disk, "two_disk_collision": two_disk_collision,
            "gaussian_disk": gaussian_disk_system}[name](**kw)
