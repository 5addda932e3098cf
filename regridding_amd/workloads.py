"""
Synthetic grid pairs and value stacks of the benchmark configurations.

Used by ``bench.py``, the tests and the autotuner.  Every generator is
seeded, so a configuration is the same arrays on every machine.  The
warp convention: with ``R`` the rotation by ``theta``,
``(xr, yr) = R (X, Y)``, then ``x = xr + ax sin(3 yr)`` and
``y = yr + ay cos(2 xr)``, both from the unwarped ``(xr, yr)``.
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "warped_grid",
    "plain_small",
    "plain_large",
    "batched_pairs",
    "inversion_problem",
]


def warped_grid(
    n: int | tuple[int, int],
    theta: float = 0.0,
    ax: float = 0.0,
    ay: float = 0.0,
    lo: float = -1.0,
    hi: float = 1.0,
) -> tuple[np.ndarray, np.ndarray]:
    """An ``(nx, ny)`` vertex grid: ``linspace(lo, hi)`` meshgrid (ij),
    rotated by `theta` and warped sinusoidally."""
    nx, ny = (n, n) if np.isscalar(n) else n
    X, Y = np.meshgrid(np.linspace(lo, hi, nx), np.linspace(lo, hi, ny), indexing="ij")
    c, s = np.cos(theta), np.sin(theta)
    xr = X * c - Y * s
    yr = X * s + Y * c
    return xr + ax * np.sin(3 * yr), yr + ay * np.cos(2 * xr)


def plain_small():
    """129x129 warped input (0.3 rad, warp 0.05) -> 101x101 rotated output
    over [-0.9, 0.9] (0.1 rad); values (4, 128, 128) Poisson(50)."""
    grid_input = warped_grid(129, 0.3, 0.05, 0.05)
    grid_output = warped_grid(101, 0.1, 0.0, 0.0, -0.9, 0.9)
    values = np.random.default_rng(1).poisson(50, (4, 128, 128)).astype(np.float64)
    return grid_input, grid_output, values


def plain_large(n_input: int = 2049, n_output: int = 1537, batch: int = 16):
    """2049x2049 warped input (0.3 rad, warp amplitudes U[0.01, 0.08]) ->
    1537x1537 rotated output over [-0.9, 0.9] (0.1 rad); values: a stack of
    `batch` Poisson images, one lognormal(3, 1) rate per image."""
    rng = np.random.default_rng(2)
    ax, ay = rng.uniform(0.01, 0.08, 2)
    grid_input = warped_grid(n_input, 0.3, ax, ay)
    grid_output = warped_grid(n_output, 0.1, 0.0, 0.0, -0.9, 0.9)
    lam = rng.lognormal(3, 1, batch)
    values = rng.poisson(lam[:, None, None], (batch, n_input - 1, n_input - 1))
    return grid_input, grid_output, values.astype(np.float64)


def batched_pairs(n_pairs: int = 256, seed: int = 3, n_min: int = 33, n_max: int = 1025):
    """Independent grid pairs: per-axis vertex counts log-uniform in
    [n_min, n_max], warp amplitudes U[0.01, 0.08], rotations U(0, pi/4),
    output/input spacing ratio U[0.5, 2] over [-0.9, 0.9], and 1-8 Poisson
    value arrays per pair.  Yields dicts."""
    rng = np.random.default_rng(seed)
    pairs = []
    for _ in range(n_pairs):
        nx, ny = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), 2))).astype(int)
        ax, ay = rng.uniform(0.01, 0.08, 2)
        th_in, th_out = rng.uniform(0, np.pi / 4, 2)
        ratio = rng.uniform(0.5, 2.0)
        mx = max(2, int(round((nx - 1) / ratio)) + 1)
        my = max(2, int(round((ny - 1) / ratio)) + 1)
        nb = int(rng.integers(1, 9))
        lam = rng.lognormal(3, 1)
        pairs.append(
            dict(
                grid_input=warped_grid((nx, ny), th_in, ax, ay),
                grid_output=warped_grid((mx, my), th_out, 0.0, 0.0, -0.9, 0.9),
                values=rng.poisson(lam, (nb, nx - 1, ny - 1)).astype(np.float64),
            )
        )
    return pairs


def inversion_problem(seed: int = 5, n_truth: int = 1025, n_obs: int = 769, n_views: int = 8):
    """The adjoint-inversion workload: a truth image on the rectilinear
    ``linspace(-1, 1, n_truth)`` vertex grid -- 50 Gaussian blobs (centres
    U(-0.8, 0.8), widths U(0.02, 0.1), amplitudes U(50, 500)) on a flat
    background of 10, sampled at the cell centres -- and `n_views`
    observation grids of ``n_obs`` vertices per axis over [-0.9, 0.9], each
    with its own rotation U(0, pi/4) and warp amplitudes U(0.01, 0.08).
    Returns ``(truth_grid, truth_image, [obs_grid, ...], rng)``; the rng
    continues the stream for the Poisson draws of the observations."""
    rng = np.random.default_rng(seed)
    grid = warped_grid(n_truth)
    xc = 0.25 * (grid[0][:-1, :-1] + grid[0][1:, :-1] + grid[0][:-1, 1:] + grid[0][1:, 1:])
    yc = 0.25 * (grid[1][:-1, :-1] + grid[1][1:, :-1] + grid[1][:-1, 1:] + grid[1][1:, 1:])
    image = np.full(xc.shape, 10.0)
    centres = rng.uniform(-0.8, 0.8, (50, 2))
    widths = rng.uniform(0.02, 0.1, 50)
    amps = rng.uniform(50, 500, 50)
    for (cx, cy), w, a in zip(centres, widths, amps):
        image += a * np.exp(-((xc - cx) ** 2 + (yc - cy) ** 2) / (2 * w * w))
    views = []
    for _ in range(n_views):
        theta = rng.uniform(0, np.pi / 4)
        ax, ay = rng.uniform(0.01, 0.08, 2)
        views.append(warped_grid(n_obs, theta, ax, ay, -0.9, 0.9))
    return grid, image, views, rng
