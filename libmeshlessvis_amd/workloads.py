"""Seeded synthetic particle sets for the five benchmark configurations.

  uniform_gaussian_4096    N=4096 uniform in [0,1]^3, w~U[0,1], h=0.05, Gaussian, z view, 64x64
  clustered_wendland       8-Gaussian mixture (sigma 0.08, centres U[0.2,0.8]^3, clipped to the cube),
                           h = distance to the 32nd neighbour, w = m/rho (SPH density with a
                           normalised Wendland C2 kernel, m = 1/N), Wendland basis, random view
  uniform_signed_gaussian  N uniform, w~U[-1,1], h=0.01, Gaussian, oblique view
  drifting_clusters        16 timesteps, N log-uniform in [1e4, 5e5], clusters drifting 0.01/step,
                           kNN(32) radii, Wendland, 32 random views per step
  lognormal_field          log-normal density (3-D Gaussian random field, P(k) ~ k^-3 on 128^3),
                           h = kNN(48), w = 1/N, Gaussian, 4 random views

Positions of the clustered/field sets are centred on the origin and viewed with step 0.5
(period 2 covers the projected unit cube).  Everything is numpy + scipy's cKDTree.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class View:
    u: Tuple[float, float, float]
    v: Tuple[float, float, float]


@dataclass
class Workload:
    name: str
    positions: np.ndarray          # (N, 3) float32
    weights: np.ndarray            # (N,) float32
    radii: Optional[np.ndarray]    # (N,) float32 or None
    basis: str
    cutoff: Tuple[int, int]
    samples: Tuple[int, int]
    step: Tuple[float, float]
    views: List[View] = field(default_factory=list)
    phi_mode: str = "analytic"

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    @property
    def terms_per_view(self) -> int:
        return self.n * 2 * self.cutoff[0] * self.cutoff[1]

    def group(self):
        return (self.positions, self.weights, self.radii, self.basis)


def random_view(rng: np.random.Generator) -> View:
    """Orthonormal (u, v) perpendicular to a uniformly random unit direction."""
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(n, a)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    return View(tuple(map(float, u)), tuple(map(float, v)))


def knn_radius(pos: np.ndarray, k: int, workers: int = -1) -> np.ndarray:
    """Distance from each point to its k-th nearest other point."""
    from scipy.spatial import cKDTree

    tree = cKDTree(pos)
    d, _ = tree.query(pos, k=[k + 1], workers=workers)
    return d[:, 0].astype(np.float32)


def wendland_density(pos: np.ndarray, h: np.ndarray, k: int, mass: float, workers: int = -1) -> np.ndarray:
    """rho_i = m * sum_j W(|x_i - x_j| / h_i) / h_i^3 over the k+1 nearest (self included),
    W(q) = 21/(2 pi) (1-q)^4 (1+4q) on q < 1 (normalised Wendland C2 in 3-D)."""
    from scipy.spatial import cKDTree

    tree = cKDTree(pos)
    d, _ = tree.query(pos, k=k + 1, workers=workers)
    q = np.clip(d / h[:, None].astype(np.float64), 0.0, 1.0)
    W = (21.0 / (2.0 * np.pi)) * (1.0 - q) ** 4 * (1.0 + 4.0 * q)
    return (mass * W.sum(axis=1) / h.astype(np.float64) ** 3)


def _clusters(rng, n, centres, sigma=0.08):
    comp = rng.integers(0, centres.shape[0], size=n)
    pos = centres[comp] + rng.normal(scale=sigma, size=(n, 3))
    return np.clip(pos, 0.0, 1.0)


def uniform_gaussian_4096(seed: int = 1) -> Workload:
    rng = np.random.default_rng(seed)
    n = 4096
    pos = rng.random((n, 3)).astype(np.float32)
    w = rng.random(n).astype(np.float32)
    return Workload("uniform_gaussian_4096", pos, w, np.full(n, 0.05, np.float32), "GAUSSIAN",
                    (32, 32), (64, 64), (1.0, 1.0), [View((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))], "reference")


def timing_test(n: int = 1_048_576, cutoff: int = 128, samples: int = 256, factor: int = 0) -> Workload:
    """The reference's timing harness (vis_timing_test/main.cpp:68-78): every particle at the
    origin with weight 1, SPH basis, radius 1e-5 (factor 0) or no radii (factor 1), step 1, the
    axis-aligned view, a square cutoff and image; Phi in the reference's formulas."""
    pos = np.zeros((n, 3), np.float32)
    w = np.ones(n, np.float32)
    radii = None if factor else np.full(n, 1e-5, np.float32)
    return Workload(f"timing_test_{n}_factor{factor}", pos, w, radii, "SPH", (cutoff, cutoff), (samples, samples),
                    (1.0, 1.0), [View((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))], "reference")


def clustered_wendland(n: int = 262_144, seed: int = 2, workers: int = -1) -> Workload:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.2, 0.8, size=(8, 3))
    pos = _clusters(rng, n, centres)
    h = knn_radius(pos, 32, workers)
    rho = wendland_density(pos, h, 32, 1.0 / n, workers)
    w = ((1.0 / n) / rho).astype(np.float32)
    view = random_view(rng)
    return Workload("clustered_wendland_262144" if n == 262_144 else f"clustered_wendland_{n}",
                    (pos - 0.5).astype(np.float32), w, h, "WENDLAND_D3_C2",
                    (128, 128), (256, 256), (0.5, 0.5), [view])


def uniform_signed_gaussian(n: int = 1_048_576, seed: int = 3) -> Workload:
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) - 0.5).astype(np.float32)
    w = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    view = random_view(rng)
    return Workload(f"uniform_signed_gaussian_{n}", pos, w, np.full(n, 0.01, np.float32), "GAUSSIAN",
                    (256, 256), (512, 512), (0.5, 0.5), [view])


def drifting_clusters(steps: int = 16, views_per_step: int = 32, seed: int = 4, n_lo: int = 10_000,
                      n_hi: int = 500_000, workers: int = -1) -> List[Workload]:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.2, 0.8, size=(8, 3))
    drift = rng.normal(size=(8, 3))
    drift /= np.linalg.norm(drift, axis=1, keepdims=True)
    out = []
    for t in range(steps):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        c = centres + 0.01 * t * drift
        pos = _clusters(rng, n, c)
        h = knn_radius(pos, 32, workers)
        rho = wendland_density(pos, h, 32, 1.0 / n, workers)
        w = ((1.0 / n) / rho).astype(np.float32)
        views = [random_view(rng) for _ in range(views_per_step)]
        out.append(Workload(f"drifting_clusters_t{t}", (pos - 0.5).astype(np.float32), w, h, "WENDLAND_D3_C2",
                            (128, 128), (256, 256), (0.5, 0.5), views))
    return out


def lognormal_positions(n: int, grid: int, rng: np.random.Generator, slope: float = -3.0) -> np.ndarray:
    """n points from the density exp(delta), delta a unit-variance Gaussian random field with
    power spectrum k^slope on a grid^3 periodic lattice.  Cells are drawn with probability
    proportional to their density and points placed uniformly inside them, which is the
    distribution rejection sampling of the piecewise-constant density produces."""
    k = np.fft.fftfreq(grid) * grid
    kz = np.fft.rfftfreq(grid) * grid
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    nz = kk > 0
    amp[nz] = kk[nz] ** (slope / 2.0)
    noise = rng.normal(size=kk.shape) + 1j * rng.normal(size=kk.shape)
    delta = np.fft.irfftn(noise * amp, s=(grid, grid, grid), axes=(0, 1, 2))
    delta = (delta - delta.mean()) / delta.std()
    rho = np.exp(delta).ravel()
    p = rho / rho.sum()
    cells = rng.choice(rho.size, size=n, p=p)
    ix, iy, iz = np.unravel_index(cells, (grid, grid, grid))
    pos = (np.stack([ix, iy, iz], axis=1) + rng.random((n, 3))) / grid
    return pos


def lognormal_field(n: int = 4_194_304, seed: int = 5, views: int = 4, grid: int = 128, workers: int = -1) -> Workload:
    rng = np.random.default_rng(seed)
    pos = lognormal_positions(n, grid, rng)
    h = knn_radius(pos, 48, workers)
    w = np.full(n, 1.0 / n, np.float32)
    vs = [random_view(rng) for _ in range(views)]
    return Workload(f"lognormal_field_{n}", (pos - 0.5).astype(np.float32), w, h, "GAUSSIAN",
                    (512, 512), (1024, 1024), (0.5, 0.5), vs)


def scaled(wl: Workload, n: int, seed: int = 0) -> Workload:
    """A uniform random subset of n particles of wl (same view and grid)."""
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(wl.n, size=min(n, wl.n), replace=False))
    return Workload(wl.name + f"_sub{n}", wl.positions[idx], wl.weights[idx],
                    None if wl.radii is None else wl.radii[idx], wl.basis, wl.cutoff, wl.samples, wl.step,
                    wl.views, wl.phi_mode)
