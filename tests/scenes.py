"""Seeded synthetic scenes shared by the tests (not collected by pytest).

A scene bundles what AbstractSolver holds after buildLaplace/buildAdvection
(abstractsolver.h:111-126): grid, resolution, AABB, inside mask, Phi, lambda,
C, gravity projection, omega, and particles.  Boxes are origin-centred,
[-0.5, 0.5]^3 with h = 1/n, because the particle cell lookup assumes
aabb.max == -aabb.min (psfsolver.cpp:80-89).
"""
from __future__ import annotations

import dataclasses
import functools

import numpy as np

from parallelspectralfluids3d_amd import advection, basis
from parallelspectralfluids3d_amd.grid import Grid, box_mask, sphere_obstacle_mask


@dataclasses.dataclass
class SceneData:
    grid: Grid
    resolution: float
    aabb_min: tuple
    aabb_max: tuple
    inside: np.ndarray
    Phi: np.ndarray        # (K, F)
    lam: np.ndarray
    C: np.ndarray          # (K, K, K), C[i, b, a] = C_i(a, b)
    g: np.ndarray          # Phi^T f
    omega: np.ndarray
    V: np.ndarray          # (K, E) curl Phi

    @property
    def K(self):
        return self.Phi.shape[0]


@functools.lru_cache(maxsize=8)
def make_scene(n=8, K=16, basis_kind="analytic", obstacle=False, seed=0, dims=None):
    dims = dims or (n, n, n)
    g = Grid(*dims)
    h = 1.0 / max(dims)
    half = [0.5 * d * h for d in dims]
    amin = tuple(-x for x in half)
    amax = tuple(half)
    inside = sphere_obstacle_mask(g, h, amin) if obstacle else box_mask(g)
    tab = None
    if basis_kind == "analytic":
        Phi, lam, tab = basis.analytic_box_modes(g, K)
    else:
        Phi, lam = basis.laplacian_eigenbasis(g, h, inside, K)
    V = (basis.curl_operator(g, h) @ Phi.T).T
    if tab is not None and not obstacle:
        C = advection.build_advection_box_analytic(g, h, Phi, lam, tab, V)
    else:
        C = advection.build_advection(g, h, inside, Phi, lam, V)
    gvec = basis.project(Phi, basis.gravity_field(g, inside))
    omega = basis.unit_energy_omega(K, seed)
    return SceneData(g, h, amin, amax, inside, Phi, lam, C, gvec, omega, V)


def particles_uniform(sd: SceneData, N, seed=1, lifetime=np.inf, fluid_only=False):
    """N particles uniform in the AABB (or in fluid voxels), w = lifetime."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(sd.aabb_min, np.float64), np.asarray(sd.aabb_max, np.float64)
    if not fluid_only:
        p = lo + (hi - lo) * rng.random((N, 3))
    else:
        x, y, z = sd.grid.voxel_coords()
        vox = np.flatnonzero(sd.inside)
        pick = vox[rng.integers(0, vox.size, N)]
        cell = np.stack([x[pick], y[pick], z[pick]], 1).astype(np.float64)
        p = lo + (cell + rng.random((N, 3))) * sd.resolution
    out = np.zeros((N, 4), np.float32)
    out[:, :3] = p
    out[:, 3] = lifetime
    return out
