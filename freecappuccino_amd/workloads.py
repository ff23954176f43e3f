"""Synthetic workloads: meshes, coefficients and right-hand sides of the five benchmark cases.

These are input synthesis only (seeded, host side).  Geometry, CSR structure,
assembly and the solves run in the GPU library.

Cases (sizes are the full ones; tests pass smaller ``n``):
  1. hex box 40^3, Laplacian (mu = -1), Dirichlet wall x=0, RHS = discrete
     divergence of a seeded U[-1,1]^3 cell velocity (inner-face central fluxes)
  2. hex box 128^3, interior vertices perturbed +-20% (seeded), RHS = 16
     Gaussian blobs, zero mean
  3. 96^3 hex -> 6 tets (Kuhn), y-grading 1.1 toward y=0 (max aspect 100),
     seeded cell permutation; upwind convection (Taylor-Green) + diffusion
     (viscosity 1e-3), BiCGStab/DILU
  4. training: 64^3 perturbed meshes, one seed per system, blob RHS
  5. hex 192^3 perturbed, x-grading 1.05 toward the x=0 wall, 32 blobs
"""
from __future__ import annotations

import numpy as np

from . import mesh as M


def case_mesh(case: int, n: int | None = None, seed: int = 0) -> M.PolyMesh:
    if case == 1:
        return M.hex_box(n or 40)
    if case == 2:
        return M.hex_box(n or 128, perturb=0.2, seed=seed + 1)
    if case == 3:
        return M.tet_box(n or 96, grade_axis=1, grade_ratio=1.1, grade_max_aspect=100.0,
                         permute_seed=seed + 3)
    if case == 4:
        return M.hex_box(n or 64, perturb=0.2, seed=seed + 1000)
    if case == 5:
        return M.hex_box(n or 192, perturb=0.2, seed=seed + 5, grade_axis=0, grade_ratio=1.05,
                         grade_max_aspect=100.0)
    raise ValueError(case)


def cell_velocity(n_cells: int, seed: int):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, size=(3, n_cells))
    return u[0].copy(), u[1].copy(), u[2].copy()


def blob_rhs(xc, yc, zc, vol, nblobs: int, seed: int, wmin: float = 0.02, wmax: float = 0.2):
    """Sum of seeded Gaussian blobs f(x_c), shifted to zero volume-weighted mean, times vol."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, size=(nblobs, 3))
    w = rng.uniform(wmin, wmax, size=nblobs)
    amp = rng.choice([-1.0, 1.0], size=nblobs)
    f = np.zeros_like(xc)
    for b in range(nblobs):
        r2 = (xc - c[b, 0]) ** 2 + (yc - c[b, 1]) ** 2 + (zc - c[b, 2]) ** 2
        f += amp[b] * np.exp(-r2 / (2.0 * w[b] ** 2))
    f -= np.dot(f, vol) / vol.sum()
    return f * vol


def taylor_green(xc, yc, zc, amplitude: float = 1.0):
    """u = A (sin X cos Y cos Z, -cos X sin Y cos Z, 0) with X = 2 pi x etc."""
    X, Y, Z = 2 * np.pi * xc, 2 * np.pi * yc, 2 * np.pi * zc
    u = amplitude * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -amplitude * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(xc)
    return u, v, w
