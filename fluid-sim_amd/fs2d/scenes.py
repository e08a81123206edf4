"""Synthetic scenes of the benchmark workloads (host-side set-up; row-major
[j, i] arrays).  The reference sets its scene up in demo/App.cpp:133-160
(SOLID border, a FLUID region, FluidSim2D::create); the workloads here add
an all-fluid box with a smoothed random velocity and Gaussian density blobs.
"""
import numpy as np

EMPTY, FLUID, SOLID = 0, 1, 2


def walls(nx, ny):
    c = np.full((ny, nx), EMPTY, dtype=np.uint8)
    c[0, :] = c[-1, :] = SOLID
    c[:, 0] = c[:, -1] = SOLID
    return c


def all_fluid(n):
    """n x n box of FLUID cells inside one-cell SOLID walls."""
    c = walls(n, n)
    c[1:-1, 1:-1] = FLUID
    return c


def dam_break(n, wfrac=0.4, hfrac=0.6):
    """FLUID block over the left wfrac of the width and the bottom hfrac of
    the height, inside SOLID walls."""
    c = walls(n, n)
    c[1:int(hfrac * n), 1:int(wfrac * n)] = FLUID
    return c


def blobs(n, rng, count):
    """Sum of `count` Gaussian blobs on the cell centres of an n x n grid:
    centres uniform in [0.15, 0.85] of the side, radii in [n/64, n/16]."""
    J, I = np.mgrid[0:n, 0:n].astype(np.float64) + 0.5
    d = np.zeros((n, n))
    for _ in range(count):
        cx, cy = rng.uniform(0.15, 0.85, size=2) * n
        rad = rng.uniform(n / 64.0, n / 16.0)
        d += np.exp(-((I - cx) ** 2 + (J - cy) ** 2) / (2 * rad * rad))
    return d


def smoke_state(n, seed, sigma=8.0, nblobs=16):
    """u, v i.i.d. N(0, 1) from numpy's default_rng(seed) (u then v),
    Gaussian-smoothed with `sigma` cells (periodic), and `nblobs` density
    blobs from the same generator; FP32.  Not yet divergence-free: the caller
    projects it with the solver it runs."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    u = gaussian_filter(rng.standard_normal((n, n + 1)), sigma, mode="wrap")
    v = gaussian_filter(rng.standard_normal((n + 1, n)), sigma, mode="wrap")
    d = blobs(n, rng, nblobs)
    return u.astype(np.float32), v.astype(np.float32), d.astype(np.float32)


def target_density(n, seed=4, count=8):
    """The differentiable rollout's target: `count` blobs from default_rng(seed)."""
    return blobs(n, np.random.default_rng(seed), count).astype(np.float32)


def ensemble_members(count=256, seed=2, dt=0.002, tau=0.97, tol=1e-6, max_iters=500):
    """The ensemble workload: `count` PIC/FLIP dam breaks, side uniform in
    {64, 96, ..., 512} and fluid block width and height uniform in [0.2, 0.7]
    of the side (numpy default_rng(seed): sides first, then the fractions),
    4 particles per fluid cell, FLIP blend 0.95 (PIC weight 0.05), dx = 1 /
    side.  Returns (cells, keyword arguments of fs2d.Sim) per member."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(np.arange(64, 513, 32), size=count)
    fr = rng.uniform(0.2, 0.7, size=(count, 2))
    out = []
    for n, (w, h) in zip(sides, fr):
        n = int(n)
        out.append((dam_break(n, w, h), dict(k=2, dt=dt, dx=1.0 / n, mode=1, alpha=0.05, tau=tau, tol=tol,
                                              max_iters=max_iters)))
    return out
