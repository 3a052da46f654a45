"""Twilight-zone (manufactured-solution) forcing.

The reference adds a forcing f to the update and to the residual when
`tz=true` (src/pseudo_transient/update.jl:23-71, residuals.jl:49-99), built by
`define_forcing_wavy` (src/twilight_zone.jl:54-194) for the exact solution
u = exp(-((x-x0)^2 + (y-y0)^2)/fwhm) * exp(-damp*t) on its wavy test grid, with
the nonlinear conductivity kappa = T^(5/2).  That function hard-codes the grid
(Lx = 12, six waves) and evaluates the metric derivatives analytically; here
the same forcing is built for any mesh, from the analytic derivatives of the
exact solution in physical coordinates.

As in the reference, f depends linearly on the discrete conductivity alpha of
the current pseudo-iteration (twilight_zone.jl:182-191):

    f = u_t - div(alpha(u_e) grad u_e) - S
      = [u_t - alpha'(u_e) |grad u_e|^2 - S]  +  alpha * [-lap u_e]
      =            A                        +  alpha *     B

so the library receives the two arrays (A, B) and adds A + B*alpha in the
fused sweep (cdiff_set_forcing).
"""
from __future__ import annotations

import numpy as np


def gaussian_exact(mesh, t, center, fwhm=1.0, damp=4.0, no_t=False):
    """u_e at the cell centroids (full array)."""
    X = _centroids(mesh)
    r2 = sum((x - c) ** 2 for x, c in zip(X, center))
    u = np.exp(-r2 / fwhm)
    if not no_t:
        u = u * np.exp(-damp * t)
    return np.asfortranarray(u)


def _centroids(mesh):
    c = mesh.centroid_coordinates
    return (c.x, c.y) if mesh.ndim == 2 else (c.x, c.y, c.z)


def gaussian_forcing(mesh, t, center, law, *, fwhm=1.0, damp=4.0, no_t=False, rho=1.0, cp=1.0, source=None,
                     axisymmetric=False):
    """(A, B) such that f = A + B*alpha makes the Gaussian exact solution
    satisfy the discrete equations up to truncation error.

    ``law``: a PowerLawConductivity (kappa0 |rho|^a |T|^b).  ``axisymmetric``:
    2D RZ geometry (x = r), whose Laplacian has the extra (1/r) du/dr."""
    X = _centroids(mesh)
    u = gaussian_exact(mesh, t, center, fwhm, damp, no_t)
    d = [-2.0 * (x - c) / fwhm for x, c in zip(X, center)]  # grad u / u
    grad2 = sum(g * g for g in d) * u * u  # |grad u|^2
    lap = u * (sum(g * g for g in d) - 2.0 * mesh.ndim / fwhm)
    if axisymmetric:
        if mesh.ndim != 2:
            raise ValueError("axisymmetric geometry is 2D (x = r, y = z)")
        lap = lap + d[0] * u / X[0]
    b = law.t_exp
    rho_term = law.kappa0 * (abs(rho) ** law.rho_exp) / (rho * cp)
    with np.errstate(divide="ignore", invalid="ignore"):
        dalpha = np.where(u != 0.0, rho_term * b * np.abs(u) ** (b - 1.0) * np.sign(u), 0.0) if b != 0 else 0.0 * u
    ut = 0.0 * u if no_t else -damp * u
    S = 0.0 if source is None else source
    A = ut - dalpha * grad2 - S
    B = -lap
    return np.asfortranarray(A), np.asfortranarray(B)
