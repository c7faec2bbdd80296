"""Synthetic pressure-solve workloads (inputs only; all solving is on the GPU).

* poisson_2d:   128 x 128 uniform quads stored unstructured (one-cell-thick
                extrusion, EMPTY front/back), manufactured p = sin(pi x) sin(pi y),
                Dirichlet walls (apps/verification/poisson_mms.cpp:16-25).
* box_pressure: 3-D hex box with jittered interior vertices, randomly
                permuted cells, optional wall-clustered layers; log-normal
                1/a_P field with a correlation length in cells, its median
                set so the operator has a unit mean diagonal; all-Neumann
                pressure system with the reference cell eliminated; RHS seeded
                N(0,1) projected to zero mean.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import FvMesh, Mesh, generate_mesh


def correlated_gaussian(mesh: Mesh, dims, corr_cells: float, seed: int) -> np.ndarray:
    """Unit-variance Gaussian random field per cell: white noise on the cell
    lattice filtered by a Gaussian of standard deviation corr/2 cells (so the
    covariance decays as exp(-(r/corr)^2)), standardised, mapped to cells."""
    from scipy.ndimage import gaussian_filter

    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nz, ny, nx))
    s = corr_cells / 2.0
    g = gaussian_filter(w, sigma=(s if nz > 1 else 0.0, s, s), mode="wrap")
    g -= g.mean()
    g /= g.std()
    ijk = mesh.cell_ijk.reshape(-1, 3)
    return g[ijk[:, 2], ijk[:, 1], ijk[:, 0]]


@dataclass
class PressureSystem:
    name: str
    mesh: Mesh
    n: int
    rowp: np.ndarray
    cols: np.ndarray
    vals: np.ndarray
    b: np.ndarray
    rau: np.ndarray
    extra: dict


def box_pressure(name, dims, jitter, sigma, corr_cells, seed, wall_layers=0, first_aspect=0.0,
                 fv: FvMesh | None = None, device=0) -> PressureSystem:
    nx, ny, nz = dims
    mesh = generate_mesh("box", nx, ny, nz, jitter=jitter, permute=True, seed=seed, wall_layers=wall_layers,
                         first_aspect=first_aspect)
    fv = fv or FvMesh(mesh, device)
    nc = mesh.n_cells
    g = correlated_gaussian(mesh, dims, corr_cells, seed + 1)
    rau = np.repeat(np.exp(sigma * g)[None], 3, 0).ravel()
    pt = np.zeros(mesh.n_patches, np.uint8)
    pv = np.zeros(mesh.n_patches)
    phi0, p0 = np.zeros(mesh.n_faces), np.zeros(nc)
    rp, cols, vals, _ = fv.pressure_assemble(pt, pv, 1.0, rau, phi0, p0, 0)
    # Median of the log-normal field: chosen so the assembled operator has a
    # unit mean diagonal. With a median of 1 the entries are O(h^4) and the
    # reference's absolute thresholds (coarse LU pivot 128 eps max(1, |A|),
    # amg_preconditioner.h:2627-2629) reject a healthy coarsest level.
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    scale = 1.0 / vals[cols == rows].mean()
    rau = rau * scale
    rp, cols, vals, _ = fv.pressure_assemble(pt, pv, 1.0, rau, phi0, p0, 0)
    rng = np.random.default_rng(seed + 2)
    bfull = rng.standard_normal(nc)
    bfull -= bfull.mean()
    b = np.ascontiguousarray(bfull[1:])  # reference cell 0 eliminated
    return PressureSystem(name, mesh, len(rp) - 1, rp, cols, vals, b, rau, dict(fv=fv, dims=dims))


def c2_box(seed=2, device=0, scale=1):
    """128^3 jittered (+-15 %) permuted box, sigma = 1, correlation 8 cells."""
    n = 128 // scale
    return box_pressure("box128_jitter15_lognormal1", (n, n, n), 0.15, 1.0, 8.0 / scale, seed, device=device)


def c5_box(seed=5, device=0, n=256):
    """256^3 box, 30 wall-clustered layers at each y wall (first-cell aspect
    ratio 1000), jitter +-10 %, permuted, sigma = 1.5."""
    layers = max(1, round(30 * n / 256))
    return box_pressure(f"box{n}_walls{layers}_ar1000_lognormal1.5", (n, n, n), 0.10, 1.5, 8.0 * n / 256, seed,
                        wall_layers=layers, first_aspect=1000.0, device=device)


def poisson_2d(n=128, device=0):
    """Dirichlet MMS on the unit square: -lap p = 2 pi^2 sin(pi x) sin(pi y)."""
    mesh = generate_mesh("extruded2d", n, n, 1)
    fv = FvMesh(mesh, device)
    fc, fs, cc, vol = fv.geometry()
    cc = cc.reshape(-1, 3)
    fc = fc.reshape(-1, 3)
    exact = np.sin(math.pi * cc[:, 0]) * np.sin(math.pi * cc[:, 1])
    source = 2.0 * math.pi ** 2 * exact
    fval = np.full(mesh.n_faces, np.nan)
    for f in range(mesh.n_internal_faces, mesh.n_faces):
        if mesh.patch_kind[mesh.patch_of_face[f]] != 1:
            fval[f] = math.sin(math.pi * fc[f, 0]) * math.sin(math.pi * fc[f, 1])
    fneu = np.zeros(mesh.n_faces, np.uint8)
    rp, cols, vals, rhs = fv.diffusion_assemble(1.0, source, fval, fneu)
    return PressureSystem("poisson2d_128_mms", mesh, len(rp) - 1, rp, cols, vals, rhs, None,
                          dict(fv=fv, exact=exact, source=source, fval=fval, fneu=fneu, vol=vol))
