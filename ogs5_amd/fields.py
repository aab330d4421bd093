"""Seeded synthetic input fields for the workloads.

Heterogeneous conductivity enters the reference per element through
``$PERMEABILITY_DISTRIBUTION`` / ``#MEDIUM_PROPERTIES_DISTRIBUTED``
(FEM/rf_mmp_new.cpp:1974, :5725; lookup in PermeabilityTensor :4771-4781).
Here the per-element value is generated directly.

Log-conductivity is a stationary Gaussian field with exponential covariance
C(r) = sigma^2 exp(-r / lambda), sampled by the randomisation (spectral)
method of Kraichnan / Shinozuka: G(x) = sigma sqrt(2/M) sum_m cos(k_m.x+phi_m)
with wave vectors drawn from the field's spectral density.  For the
exponential covariance in d dimensions that density is proportional to
(1 + lambda^2 |k|^2)^(-(d+1)/2), i.e. k = u / lambda with u a d-variate
Student-t vector with one degree of freedom.  On a structured cell grid the
sum is evaluated separably with complex exponentials.
"""
from __future__ import annotations

import numpy as np


def _wave_vectors(rng: np.random.Generator, dim: int, n_modes: int, corr_len: float):
    z = rng.standard_normal((n_modes, dim))
    w = rng.chisquare(1.0, size=(n_modes, 1))
    u = z / np.sqrt(w)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n_modes)
    return u / corr_len, phi


def gaussian_field_grid(shape, h, corr_len: float, sigma: float, seed: int,
                        n_modes: int = 256) -> np.ndarray:
    """Zero-mean Gaussian field at the cell centres of a structured grid.

    shape = cells per axis (x fastest in the returned flat array), h = cell
    size per axis.  Returns float64 [prod(shape)].
    """
    dim = len(shape)
    rng = np.random.Generator(np.random.PCG64(seed))
    k, phi = _wave_vectors(rng, dim, n_modes, corr_len)
    centres = [(np.arange(shape[a]) + 0.5) * h[a] for a in range(dim)]
    amp = np.exp(1j * phi)  # [M]
    # E[a][m, i] = exp(i k_ma x_i): the mode sum factorises per axis
    E = [np.exp(1j * np.outer(k[:, a], centres[a])) for a in range(dim)]
    if dim == 1:
        out = (amp @ E[0]).real
    elif dim == 2:
        out = ((E[1].T * amp) @ E[0]).real  # [y, x]
    else:
        out = np.empty((shape[2], shape[1], shape[0]))
        for z in range(shape[2]):
            out[z] = ((E[1].T * (amp * E[2][:, z])) @ E[0]).real
    return (sigma * np.sqrt(2.0 / n_modes) * out).ravel()


def lognormal_conductivity(mesh, mean_ln: float, sigma_ln: float, corr_len: float,
                           seed: int, n_modes: int = 256) -> np.ndarray:
    """Per-element K = exp(mean_ln + G) for a structured mesh (cells -> elements)."""
    g = gaussian_field_grid(mesh.shape, mesh.h, corr_len, sigma_ln, seed, n_modes)
    return np.exp(mean_ln + g)[mesh.cell_of_elem]


def gauss_point_coords(mesh, n_gauss: int = 3) -> np.ndarray:
    """Physical coordinates of every Gauss point: [n_elem, ngp, 3].

    Interpolates node coordinates with the element's own shape functions at
    the reference Gauss points (ShapeFunctionPool, FEM/ShapeFunctionPool.cpp:109).
    """
    from .elements import shape_tables

    N, _ = shape_tables(mesh.etype, n_gauss)  # [ngp, npe]
    c = mesh.conn
    X = np.stack([mesh.x[c], mesh.y[c], mesh.z[c]], axis=2)  # [E, npe, 3]
    return np.einsum("ga,ead->egd", N, X)


def divergence_free_velocity(points: np.ndarray, u_mean: float, rel_amp: float,
                             length: float, seed: int, n_modes: int = 8) -> np.ndarray:
    """v = curl(0, 0, psi) with psi = u_mean*y + random low modes.

    Exactly divergence-free: v = (dpsi/dy, -dpsi/dx, 0).  The random part is
    scaled so that |v - (u_mean,0,0)| stays below rel_amp*u_mean, which sets
    the spread of the grid Peclet number.  points: [..., 3] -> [..., 3].
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    kk = rng.integers(1, 4, size=(n_modes, 2)) * (2.0 * np.pi / length)
    kk *= rng.choice([-1.0, 1.0], size=(n_modes, 2))
    ph = rng.uniform(0, 2 * np.pi, size=n_modes)
    amp = rng.standard_normal(n_modes)
    # bound on |grad psi'| is sum |amp_m| |k_m|; scale to rel_amp*u_mean
    bound = np.sum(np.abs(amp) * np.linalg.norm(kk, axis=1))
    amp *= rel_amp * u_mean / bound
    x = points[..., 0]
    y = points[..., 1]
    vx = np.full(x.shape, u_mean, dtype=np.float64)
    vy = np.zeros(x.shape, dtype=np.float64)
    for m in range(n_modes):
        arg = kk[m, 0] * x + kk[m, 1] * y + ph[m]
        c = np.cos(arg) * amp[m]
        vx += c * kk[m, 1]
        vy -= c * kk[m, 0]
    return np.stack([vx, vy, np.zeros_like(vx)], axis=-1)
