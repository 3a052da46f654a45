"""Curvilinear meshes and their FP64 metrics.

The reference takes every metric array on its hot path from CurvilinearGrids.jl
(not part of the reference repository).  This module defines the mesh objects
the solver reads, with the field names the hot path uses:

* ``cell_center_metrics``: inverse metrics d(xi_a)/d(x_c) at cell centres and
  the forward metrics d(x_c)/d(xi_a) and Jacobian J
  (read at src/pseudo_transient/flux_divergence.jl:11-14,
  src/pseudo_transient/mesh_metric_cache.jl:57-61,
  src/pseudo_transient/PseudoTransientSolver.jl:127);
* ``edge_metrics``: conservative metrics xi_hat = J * grad(xi) and J on the
  i+1/2, j+1/2 (k+1/2) faces (mesh_metric_cache.jl:51-85);
* ``centroid_coordinates``, ``coords()``, ``nnodes``, ``nhalo``, index limits.

Index space is the logical cell index (xi = i), so metrics are per cell width.
Discretization (FP64, second order):

* node coordinates are known on the domain plus ``nhalo`` ghost layers (from an
  analytic map, or extrapolated linearly from the domain nodes);
* cell-centre forward metrics average the node differences of the cell's
  edges along each logical direction; J = det; inverse metrics = inverse matrix;
* a face's tangential derivatives come from the face's own nodes, its normal
  derivative is the mean of the two adjacent cells'; xi_hat = J * inverse.

Arrays are Fortran-ordered numpy arrays of shape (NI, NJ[, NK]) including
halos, so ``a[i, j]`` and the memory layout match the reference's Julia arrays.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from types import SimpleNamespace
from typing import Callable, Sequence

import numpy as np


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def _extend_nodes(x: np.ndarray, nh: int) -> np.ndarray:
    """Linearly extrapolate ``nh`` ghost node layers on every side."""
    for ax in range(x.ndim):
        for _ in range(nh):
            lo = 2.0 * np.take(x, [0], axis=ax) - np.take(x, [1], axis=ax)
            hi = 2.0 * np.take(x, [-1], axis=ax) - np.take(x, [-2], axis=ax)
            x = np.concatenate([lo, x, hi], axis=ax)
    return x


@dataclass
class _Limits:
    ilo: int
    ihi: int
    jlo: int
    jhi: int
    klo: int = 0
    khi: int = 0


class CurvilinearGrid2D:
    """2D curvilinear mesh (stand-in for CurvilinearGrids.CurvilinearGrid2D).

    ``x``, ``y``: node coordinates of the domain, shape (ni+1, nj+1) for ni x nj
    cells.  ``xg``, ``yg`` (optional) hold the nodes including ``nhalo`` ghost
    layers; without them the ghost nodes are extrapolated linearly.
    """

    ndim = 2

    def __init__(self, x, y, nhalo: int = 1, xg=None, yg=None):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        if x.shape != y.shape or x.ndim != 2:
            raise ValueError("x and y must be 2D node arrays of equal shape")
        self.nhalo = h = int(nhalo)
        if h < 1:
            raise ValueError("nhalo must be >= 1")
        self.nnodes = x.shape
        self.ncells = (x.shape[0] - 1, x.shape[1] - 1)
        self._x, self._y = x, y
        X = _extend_nodes(x, h) if xg is None else np.asarray(xg, dtype=np.float64)
        Y = _extend_nodes(y, h) if yg is None else np.asarray(yg, dtype=np.float64)
        ni, nj = self.ncells
        NI, NJ = ni + 2 * h, nj + 2 * h
        assert X.shape == (NI + 1, NJ + 1)
        self.dims_full = (NI, NJ)
        self.domain_limits = SimpleNamespace(
            cell=_Limits(h, h + ni - 1, h, h + nj - 1), node=_Limits(h, h + ni, h, h + nj)
        )
        self.iterators = SimpleNamespace(
            cell=SimpleNamespace(
                full=(slice(0, NI), slice(0, NJ)), domain=(slice(h, h + ni), slice(h, h + nj))
            )
        )
        self._XG, self._YG = X, Y
        xc = 0.25 * (X[:-1, :-1] + X[1:, :-1] + X[:-1, 1:] + X[1:, 1:])
        yc = 0.25 * (Y[:-1, :-1] + Y[1:, :-1] + Y[:-1, 1:] + Y[1:, 1:])
        self.centroid_coordinates = SimpleNamespace(x=_f(xc), y=_f(yc))
        self._metrics = None

    def nodes_full(self):
        """Node coordinates including the ghost nodes, shape dims_full + 1."""
        return self._XG, self._YG

    @property
    def cell_center_metrics(self):
        if self._metrics is None:
            self._metrics = self._build(self._XG, self._YG)
        return self._metrics[0]

    @property
    def edge_metrics(self):
        if self._metrics is None:
            self._metrics = self._build(self._XG, self._YG)
        return self._metrics[1]

    # -- metrics (computed on first use; the solver can build them on the device
    #    from nodes_full() with the same arithmetic) -------------------------
    @staticmethod
    def _forward(X, Y):
        def dxi(A):
            return 0.5 * ((A[1:, :-1] - A[:-1, :-1]) + (A[1:, 1:] - A[:-1, 1:]))

        def deta(A):
            return 0.5 * ((A[:-1, 1:] - A[:-1, :-1]) + (A[1:, 1:] - A[1:, :-1]))

        xxi, xeta, yxi, yeta = dxi(X), deta(X), dxi(Y), deta(Y)
        return xxi, xeta, yxi, yeta, xxi * yeta - xeta * yxi

    def _build(self, X, Y):
        # cell-centre forward metrics
        xxi, xeta, yxi, yeta, J = self._forward(X, Y)
        cell_center_metrics = SimpleNamespace(
            J=_f(J),
            xi=SimpleNamespace(x1=_f(yeta / J), x2=_f(-xeta / J)),
            eta=SimpleNamespace(x1=_f(-yxi / J), x2=_f(xxi / J)),
            x1=SimpleNamespace(xi=_f(xxi), eta=_f(xeta)),
            x2=SimpleNamespace(xi=_f(yxi), eta=_f(yeta)),
        )
        NI, NJ = self.dims_full

        def face(axis):
            # face a+1/2 of cell (i, j) along `axis`; last layer has no outer
            # neighbour and keeps the cell's own values (never read)
            xq = np.empty((NI, NJ))
            yq = np.empty((NI, NJ))
            xt = np.empty((NI, NJ))
            yt = np.empty((NI, NJ))
            if axis == 0:
                # normal (xi) derivative: mean of cells i, i+1
                xq[:-1] = 0.5 * (xxi[:-1] + xxi[1:])
                yq[:-1] = 0.5 * (yxi[:-1] + yxi[1:])
                xq[-1], yq[-1] = xxi[-1], yxi[-1]
                # tangential (eta) derivative along the face nodes (i+1, j..j+1)
                xt[:] = X[1:, 1:] - X[1:, :-1]
                yt[:] = Y[1:, 1:] - Y[1:, :-1]
                x_xi, x_eta, y_xi, y_eta = xq, xt, yq, yt
            else:
                xq[:, :-1] = 0.5 * (xeta[:, :-1] + xeta[:, 1:])
                yq[:, :-1] = 0.5 * (yeta[:, :-1] + yeta[:, 1:])
                xq[:, -1], yq[:, -1] = xeta[:, -1], yeta[:, -1]
                xt[:] = X[1:, 1:] - X[:-1, 1:]
                yt[:] = Y[1:, 1:] - Y[:-1, 1:]
                x_xi, x_eta, y_xi, y_eta = xt, xq, yt, yq
            Jf = x_xi * y_eta - x_eta * y_xi
            return SimpleNamespace(
                J=_f(Jf),
                xi_hat=SimpleNamespace(x1=_f(y_eta), x2=_f(-x_eta)),
                eta_hat=SimpleNamespace(x1=_f(-y_xi), x2=_f(x_xi)),
            )

        return cell_center_metrics, SimpleNamespace(i_half=face(0), j_half=face(1))

    # -- accessors used by the solver --------------------------------------
    def coords(self):
        """Domain node coordinates (CurvilinearGrids.coords)."""
        return self._x, self._y

    def cell_metric_list(self):
        """[xi_x, xi_y, eta_x, eta_y]: cell[a*2 + c] = d(xi_a)/d(x_c)."""
        m = self.cell_center_metrics
        return [m.xi.x1, m.xi.x2, m.eta.x1, m.eta.x2]

    def edge_metric_list(self):
        """Per face axis b: [xi_hat.x1, xi_hat.x2, eta_hat.x1, eta_hat.x2, J]."""
        out = []
        for f in (self.edge_metrics.i_half, self.edge_metrics.j_half):
            out += [f.xi_hat.x1, f.xi_hat.x2, f.eta_hat.x1, f.eta_hat.x2, f.J]
        return out

    def phys_dims(self):
        """(L, spacing) as src/pseudo_transient/PseudoTransientSolver.jl:122-139."""
        dom = self.iterators.cell.domain
        J = self._forward(self._XG, self._YG)[4] if self._metrics is None else self.cell_center_metrics.J
        ds = float(np.sqrt(np.min(J[dom])))
        x, y = self.coords()
        L = min(abs(x.max() - x.min()), abs(y.max() - y.min()))
        return float(L), (ds, ds)

    def cellsize_withhalo(self):
        return self.dims_full

    @classmethod
    def from_mapping(cls, fn: Callable, ncells: Sequence[int], nhalo: int = 1, extent=((0.0, 1.0), (0.0, 1.0))):
        """Mesh of ``fn(xi, eta) -> (x, y)`` over logical ``extent`` with ``ncells`` cells,
        the ghost nodes taken from the same map."""
        ni, nj = ncells
        h = nhalo
        (a0, a1), (b0, b1) = extent
        p = np.arange(-h, ni + h + 1, dtype=np.float64)
        q = np.arange(-h, nj + h + 1, dtype=np.float64)
        XI, ETA = np.meshgrid(a0 + (a1 - a0) * p / ni, b0 + (b1 - b0) * q / nj, indexing="ij")
        XG, YG = fn(XI, ETA)
        XG = np.asarray(XG, dtype=np.float64)
        YG = np.asarray(YG, dtype=np.float64)
        return cls(XG[h : h + ni + 1, h : h + nj + 1], YG[h : h + ni + 1, h : h + nj + 1], nhalo, XG, YG)


class CurvilinearGrid3D:
    """3D curvilinear mesh (stand-in for CurvilinearGrids.CurvilinearGrid3D)."""

    ndim = 3

    def __init__(self, x, y, z, nhalo: int = 1, xg=None, yg=None, zg=None):
        x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
        if not (x.shape == y.shape == z.shape) or x.ndim != 3:
            raise ValueError("x, y, z must be 3D node arrays of equal shape")
        self.nhalo = h = int(nhalo)
        if h < 1:
            raise ValueError("nhalo must be >= 1")
        self.nnodes = x.shape
        self.ncells = tuple(s - 1 for s in x.shape)
        self._xyz = (x, y, z)
        G = [(_extend_nodes(a, h) if g is None else np.asarray(g, dtype=np.float64)) for a, g in zip((x, y, z), (xg, yg, zg))]
        ni, nj, nk = self.ncells
        self.dims_full = (ni + 2 * h, nj + 2 * h, nk + 2 * h)
        for g in G:
            assert g.shape == tuple(d + 1 for d in self.dims_full)
        self.domain_limits = SimpleNamespace(
            cell=_Limits(h, h + ni - 1, h, h + nj - 1, h, h + nk - 1),
            node=_Limits(h, h + ni, h, h + nj, h, h + nk),
        )
        NI, NJ, NK = self.dims_full
        self.iterators = SimpleNamespace(
            cell=SimpleNamespace(
                full=(slice(0, NI), slice(0, NJ), slice(0, NK)),
                domain=(slice(h, h + ni), slice(h, h + nj), slice(h, h + nk)),
            )
        )
        self._G = G
        cen = []
        for c in range(3):
            A = G[c]
            acc = 0.0
            for u in (0, 1):
                for v in (0, 1):
                    for w in (0, 1):
                        acc = acc + A[slice(u, A.shape[0] - 1 + u), slice(v, A.shape[1] - 1 + v), slice(w, A.shape[2] - 1 + w)]
            cen.append(_f(acc / 8.0))
        self.centroid_coordinates = SimpleNamespace(x=cen[0], y=cen[1], z=cen[2])
        self._metrics = None

    def nodes_full(self):
        """Node coordinates including the ghost nodes, shape dims_full + 1."""
        return tuple(self._G)

    @property
    def cell_center_metrics(self):
        if self._metrics is None:
            self._metrics = self._build(self._G)
        return self._metrics[0]

    @property
    def edge_metrics(self):
        if self._metrics is None:
            self._metrics = self._build(self._G)
        return self._metrics[1]

    @staticmethod
    def _det_inv(M):
        """det and inverse of the 3x3 matrices M[..., c, a] by cofactors, in a
        fixed order of operations (the device kernel evaluates the same)."""
        m = [[M[..., r, q] for q in range(3)] for r in range(3)]
        c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1]
        c01 = m[1][0] * m[2][2] - m[1][2] * m[2][0]
        c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0]
        det = m[0][0] * c00 - m[0][1] * c01 + m[0][2] * c02
        adj = [
            [c00, m[0][2] * m[2][1] - m[0][1] * m[2][2], m[0][1] * m[1][2] - m[0][2] * m[1][1]],
            [m[1][2] * m[2][0] - m[1][0] * m[2][2], m[0][0] * m[2][2] - m[0][2] * m[2][0],
             m[0][2] * m[1][0] - m[0][0] * m[1][2]],
            [c02, m[0][1] * m[2][0] - m[0][0] * m[2][1], m[0][0] * m[1][1] - m[0][1] * m[1][0]],
        ]
        inv = np.empty(M.shape)
        for a in range(3):
            for c in range(3):
                inv[..., a, c] = adj[a][c] / det
        return det, inv

    @staticmethod
    def _dnode(A, axis):
        """Cell-centre derivative along `axis`: mean of the 4 parallel cell edges."""
        sl = [slice(None)] * 3

        def s(ax_vals):
            t = [slice(None)] * 3
            for ax, v in ax_vals.items():
                t[ax] = slice(1, None) if v else slice(None, -1)
            return tuple(t)

        others = [a for a in range(3) if a != axis]
        acc = 0.0
        for u in (0, 1):
            for v in (0, 1):
                hi = s({axis: 1, others[0]: u, others[1]: v})
                lo = s({axis: 0, others[0]: u, others[1]: v})
                acc = acc + (A[hi] - A[lo])
        del sl
        return 0.25 * acc

    def _build(self, G):
        X = G
        # D[c][a] = d(x_c)/d(xi_a) at cell centres
        D = [[self._dnode(X[c], a) for a in range(3)] for c in range(3)]
        M = np.stack([np.stack(D[c], axis=-1) for c in range(3)], axis=-2)  # [..., c, a]
        J, inv = self._det_inv(M)  # inv[..., a, c] = d(xi_a)/d(x_c)
        names = ("xi", "eta", "zeta")
        comp = ("x1", "x2", "x3")
        ccm = SimpleNamespace(J=_f(J))
        for a in range(3):
            setattr(ccm, names[a], SimpleNamespace(**{comp[c]: _f(inv[..., a, c]) for c in range(3)}))
        for c in range(3):
            setattr(ccm, comp[c], SimpleNamespace(**{names[a]: _f(M[..., c, a]) for a in range(3)}))
        shape = self.dims_full
        faces = []
        for b in range(3):
            Mf = np.empty(shape + (3, 3))
            # normal derivative: mean of the two adjacent cells
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            sl_lo[b] = slice(None, -1)
            sl_hi[b] = slice(1, None)
            last = [slice(None)] * 3
            last[b] = slice(-1, None)
            for c in range(3):
                col = np.empty(shape)
                col[tuple(sl_lo)] = 0.5 * (M[..., c, b][tuple(sl_lo)] + M[..., c, b][tuple(sl_hi)])
                col[tuple(last)] = M[..., c, b][tuple(last)]
                Mf[..., c, b] = col
            # tangential derivatives from the face's own 4 nodes (plane node index +1 along b)
            for a in range(3):
                if a == b:
                    continue
                o = [x for x in range(3) if x not in (a, b)][0]
                for c in range(3):
                    A = X[c]
                    acc = 0.0
                    for v in (0, 1):
                        hi = [None] * 3
                        lo = [None] * 3
                        hi[b] = lo[b] = slice(1, None)
                        hi[a] = slice(1, None)
                        lo[a] = slice(None, -1)
                        hi[o] = lo[o] = slice(1, None) if v else slice(None, -1)
                        acc = acc + (A[tuple(hi)] - A[tuple(lo)])
                    Mf[..., c, a] = 0.5 * acc
            Jf, invf = self._det_inv(Mf)
            hat = invf * Jf[..., None, None]
            ns = SimpleNamespace(J=_f(Jf))
            for a in range(3):
                setattr(ns, names[a] + "_hat", SimpleNamespace(**{comp[c]: _f(hat[..., a, c]) for c in range(3)}))
            faces.append(ns)
        return ccm, SimpleNamespace(i_half=faces[0], j_half=faces[1], k_half=faces[2])

    def coords(self):
        return self._xyz

    def cell_metric_list(self):
        m = self.cell_center_metrics
        return [getattr(getattr(m, a), c) for a in ("xi", "eta", "zeta") for c in ("x1", "x2", "x3")]

    def edge_metric_list(self):
        out = []
        for f in (self.edge_metrics.i_half, self.edge_metrics.j_half, self.edge_metrics.k_half):
            for a in ("xi_hat", "eta_hat", "zeta_hat"):
                for c in ("x1", "x2", "x3"):
                    out.append(getattr(getattr(f, a), c))
            out.append(f.J)
        return out

    def phys_dims(self):
        """(L, spacing) as src/pseudo_transient/PseudoTransientSolver.jl:141-153."""
        x, y, z = self.coords()
        spacing = (float(np.min(np.diff(x, axis=0))), float(np.min(np.diff(y, axis=1))), float(np.min(np.diff(z, axis=2))))
        L = min(abs(x.max() - x.min()), abs(y.max() - y.min()), abs(z.max() - z.min()))
        return float(L), spacing

    def cellsize_withhalo(self):
        return self.dims_full

    @classmethod
    def from_mapping(cls, fn: Callable, ncells: Sequence[int], nhalo: int = 1, extent=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0))):
        ni, nj, nk = ncells
        h = nhalo
        axes = []
        for (a0, a1), n in zip(extent, ncells):
            p = np.arange(-h, n + h + 1, dtype=np.float64)
            axes.append(a0 + (a1 - a0) * p / n)
        XI, ETA, ZETA = np.meshgrid(*axes, indexing="ij")
        XG, YG, ZG = (np.asarray(a, dtype=np.float64) for a in fn(XI, ETA, ZETA))
        d = (slice(h, h + ni + 1), slice(h, h + nj + 1), slice(h, h + nk + 1))
        return cls(XG[d], YG[d], ZG[d], nhalo, XG, YG, ZG)


def wavy_map_2d(amplitude: float):
    """x = xi + A sin(2 pi xi) sin(2 pi eta), y = eta + A sin(2 pi xi) sin(2 pi eta)."""

    def fn(xi, eta):
        s = amplitude * np.sin(2 * np.pi * xi) * np.sin(2 * np.pi * eta)
        return xi + s, eta + s

    return fn


def wavy_map_3d(amplitude: float):
    """x = xi + A sin(2 pi xi) sin(2 pi eta) sin(2 pi zeta), same perturbation for y, z."""

    def fn(xi, eta, zeta):
        s = amplitude * np.sin(2 * np.pi * xi) * np.sin(2 * np.pi * eta) * np.sin(2 * np.pi * zeta)
        return xi + s, eta + s, zeta + s

    return fn


def sheared_map_2d(shear: float):
    """Affine map x = xi + shear*eta, y = eta (constant metrics, g12 != 0)."""

    def fn(xi, eta):
        return xi + shear * eta, eta + 0.0 * xi

    return fn


def RectlinearGrid(lo, hi, ncells, nhalo: int = 1):
    """Uniform Cartesian mesh of ``ncells`` cells between corners ``lo`` and ``hi``
    (CurvilinearGrids.RectlinearGrid, as used in test/unit/test_bc.jl:49)."""
    lo = tuple(float(v) for v in lo)
    hi = tuple(float(v) for v in hi)
    if len(ncells) == 2:
        def fn(xi, eta):
            return lo[0] + (hi[0] - lo[0]) * xi, lo[1] + (hi[1] - lo[1]) * eta

        return CurvilinearGrid2D.from_mapping(fn, ncells, nhalo)

    def fn3(xi, eta, zeta):
        return (lo[0] + (hi[0] - lo[0]) * xi, lo[1] + (hi[1] - lo[1]) * eta, lo[2] + (hi[2] - lo[2]) * zeta)

    return CurvilinearGrid3D.from_mapping(fn3, ncells, nhalo)
