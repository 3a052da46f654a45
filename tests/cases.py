"""Seeded synthetic problems shared by the tests (the benchmark configurations
at small sizes).  Test infrastructure."""
from __future__ import annotations

import numpy as np

from curvilineardiffusionmanuscript_amd.mesh import (
    CurvilinearGrid2D,
    CurvilinearGrid3D,
    sheared_map_2d,
    wavy_map_2d,
    wavy_map_3d,
)


def gaussian_2d(n, amplitude=0.03, nhalo=1, ni=None):
    """Config-1 physics: wavy map, T0 = 1 + 10 exp(-r^2/0.01) at centroids."""
    ni = n if ni is None else ni
    mesh = CurvilinearGrid2D.from_mapping(wavy_map_2d(amplitude), (ni, n), nhalo)
    xc, yc = mesh.centroid_coordinates.x, mesh.centroid_coordinates.y
    T = np.asfortranarray(1.0 + 10.0 * np.exp(-((xc - 0.5) ** 2 + (yc - 0.5) ** 2) / 0.01))
    return mesh, T


def hot_disk_2d(n, seed=0, amplitude=0.05, nhalo=1, radius=0.05, background=0.01):
    """Config-2 physics: wavy map, background 0.01, seeded hot disk T=1."""
    mesh = CurvilinearGrid2D.from_mapping(wavy_map_2d(amplitude), (n, n), nhalo)
    rng = np.random.default_rng(seed)
    cx, cy = rng.uniform(0.3, 0.7, size=2)
    xc, yc = mesh.centroid_coordinates.x, mesh.centroid_coordinates.y
    T = np.where((xc - cx) ** 2 + (yc - cy) ** 2 <= radius**2, 1.0, background)
    return mesh, np.asfortranarray(T)


def gaussian_spots_3d(n, seed=0, amplitude=0.04, nhalo=1, nspots=8, background=0.01):
    """Config-4 physics: 3D wavy map, 8 seeded Gaussian spots on background 0.01."""
    mesh = CurvilinearGrid3D.from_mapping(wavy_map_3d(amplitude), (n, n, n), nhalo)
    rng = np.random.default_rng(seed)
    x, y, z = (mesh.centroid_coordinates.x, mesh.centroid_coordinates.y, mesh.centroid_coordinates.z)
    T = np.full(mesh.dims_full, background)
    for _ in range(nspots):
        a = rng.uniform(0.5, 1.0)
        w = rng.uniform(0.02, 0.05)
        c = rng.uniform(0.15, 0.85, size=3)
        T = T + a * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / w**2)
    return mesh, np.asfortranarray(T)


def sheared_2d(n, shear=0.3, nhalo=1):
    mesh = CurvilinearGrid2D.from_mapping(sheared_map_2d(shear), (n, n), nhalo)
    return mesh
