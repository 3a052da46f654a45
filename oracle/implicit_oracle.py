"""FP64 implicit-direct oracle -- TEST INFRASTRUCTURE.

A restatement of the reference's implicit scheme (src/implicit/), used by the
tests and the benchmarks as the independent FP64 solution the APT results are
compared against.  Nothing in the product path imports this module.

One backward-Euler step with the conductivity lagged at T^n:
  * BCs applied to T (boundary_operators.jl, via the APT oracle's applybcs),
    alpha = |kappa(rho, T) / (rho cp)| over the diffusion region
    (conductivity.jl:76-84);
  * the linear system on the diffusion region = domain + one halo layer
    (ImplictSolver.jl:70-121):
      - inner rows: the conservative 9-point (2D) stencil of stencil_2d
        (stencils.jl:14-37) built from conservative_edge_terms
        (edge_terms.jl:32-80) with face diffusivities mean(alpha_c, alpha_nb)
        (inner_operators.jl:246-256); rhs = (S dt + u) J
        (inner_operators.jl:83-84);
      - boundary rows (i = ilo/ihi, j = jlo/jhi, corners by the first matching
        face in ilo, ihi, jlo, jhi order, inner_operators.jl:188-211): 1 on
        the diagonal, -1 on the inner neighbour (Neumann) or on the cell one
        inside the opposite face (periodic) (init_matrix.jl:67-165); rhs 0, or
        the Dirichlet value (boundary_operators.jl:271-273);
  * a direct sparse solve (SciPy SuperLU; the reference uses Pardiso, KLU or
    UMFPACK, all exact to rounding), the cutoff, next_dt on the diffusion
    region, and the copy back into T (ImplictSolver.jl:198-248).

Pins (tests/test_implicit_oracle.py): the reference's known answers for the
edge terms (test/unit/test_edge_terms.jl:78-94), the unit-metric stencils
(test/unit/test_inner_operator.jl:44,76-78) and the sparsity patterns
(test/unit/test_matrix_construction.jl:21-37,70-117); on uniform grids the
system equals the 5-point backward-Euler system.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

OFF2 = [(-1, -1), (0, -1), (1, -1), (-1, 0), (0, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]


def edge_terms_2d(a_ip, a_im, a_jp, a_jm, m):
    """conservative_edge_terms (edge_terms.jl:32-80).  `m` maps the reference's
    names (J_ip, Jxi_x_ip, ...) to values; returns the eight a_J.. terms."""
    out = {}
    out["xi2_ip"] = a_ip * (m["Jxi_x_ip"] ** 2 + m["Jxi_y_ip"] ** 2) / m["J_ip"]
    out["xi2_im"] = a_im * (m["Jxi_x_im"] ** 2 + m["Jxi_y_im"] ** 2) / m["J_im"]
    out["eta2_jp"] = a_jp * (m["Jeta_x_jp"] ** 2 + m["Jeta_y_jp"] ** 2) / m["J_jp"]
    out["eta2_jm"] = a_jm * (m["Jeta_x_jm"] ** 2 + m["Jeta_y_jm"] ** 2) / m["J_jm"]
    out["xieta_ip"] = a_ip * (m["Jxi_x_ip"] * m["Jeta_x_ip"] + m["Jxi_y_ip"] * m["Jeta_y_ip"]) / (4 * m["J_ip"])
    out["xieta_im"] = a_im * (m["Jxi_x_im"] * m["Jeta_x_im"] + m["Jxi_y_im"] * m["Jeta_y_im"]) / (4 * m["J_im"])
    out["etaxi_jp"] = a_jp * (m["Jxi_x_jp"] * m["Jeta_x_jp"] + m["Jxi_y_jp"] * m["Jeta_y_jp"]) / (4 * m["J_jp"])
    out["etaxi_jm"] = a_jm * (m["Jxi_x_jm"] * m["Jeta_x_jm"] + m["Jxi_y_jm"] * m["Jeta_y_jm"]) / (4 * m["J_jm"])
    return out


def stencil_2d(e, J, dt):
    """stencil_2d (stencils.jl:14-37), in OFF2 order."""
    c = J + (e["eta2_jp"] + e["eta2_jm"] + e["xi2_ip"] + e["xi2_im"]) * dt
    sw = (-e["etaxi_jm"] - e["xieta_im"]) * dt
    se = (e["etaxi_jm"] + e["xieta_ip"]) * dt
    nw = (e["etaxi_jp"] + e["xieta_im"]) * dt
    ne = (-e["etaxi_jp"] - e["xieta_ip"]) * dt
    ee = (-e["etaxi_jp"] + e["etaxi_jm"] - e["xi2_ip"]) * dt
    ww = (e["etaxi_jp"] - e["etaxi_jm"] - e["xi2_im"]) * dt
    nn = (-e["eta2_jp"] - e["xieta_ip"] + e["xieta_im"]) * dt
    ss = (-e["eta2_jm"] + e["xieta_ip"] - e["xieta_im"]) * dt
    return [sw, ss, se, ww, c, ee, nw, nn, ne]


def stencil_3d_unit(a, dt, J=1.0):
    """The 27-point stencil of stencil_3d (stencils.jl:39-137) for orthogonal
    unit metrics (no cross terms) and face terms a (six values ip, im, jp, jm,
    kp, km), in (k, j, i) lexicographic order."""
    ip, im, jp, jm, kp, km = a
    s = np.zeros((3, 3, 3))
    s[1, 1, 1] = J + (kp + km + jp + jm + ip + im) * dt
    s[1, 1, 2], s[1, 1, 0] = -ip * dt, -im * dt
    s[1, 2, 1], s[1, 0, 1] = -jp * dt, -jm * dt
    s[2, 1, 1], s[0, 1, 1] = -kp * dt, -km * dt
    return s.ravel()


def _bc_kind(b):
    return b[0] if isinstance(b, tuple) else b


def init_pattern_2d(dims, bcs):
    """Sparsity pattern of the 2D coefficient matrix (init_matrix.jl:1-172):
    returns a CSR matrix whose stored entries are the reference's."""
    ni, nj = dims
    kinds = [_bc_kind(b) for b in bcs]
    pairs = [(kinds[0], kinds[1]), (kinds[2], kinds[3])]
    for lo, hi in pairs:
        if (lo == "periodic") != (hi == "periodic"):
            raise ValueError("Inconsistent periodic boundary conditions")
    LI = lambda i, j: i + ni * j  # noqa: E731  (column-major, 0-based)
    rows, cols, vals = [], [], []
    for j in range(nj):
        for i in range(ni):
            rows.append(LI(i, j)), cols.append(LI(i, j)), vals.append(1.0)
    for j in range(1, nj - 1):
        for i in range(1, ni - 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    if di or dj:
                        rows.append(LI(i, j)), cols.append(LI(i + di, j + dj)), vals.append(0.0)
    ilo, ihi, jlo, jhi = 0, ni - 1, 0, nj - 1
    for j in range(1, nj - 1):
        if kinds[0] in ("neumann", "periodic"):
            c = LI(1, j) if kinds[0] == "neumann" else LI(ihi - 1, j)
            rows.append(LI(ilo, j)), cols.append(c), vals.append(-1.0)
    for j in range(1, nj - 1):
        if kinds[1] in ("neumann", "periodic"):
            c = LI(ihi - 1, j) if kinds[1] == "neumann" else LI(ilo + 1, j)
            rows.append(LI(ihi, j)), cols.append(c), vals.append(-1.0)
    for i in range(1, ni - 1):
        if kinds[2] in ("neumann", "periodic"):
            c = LI(i, 1) if kinds[2] == "neumann" else LI(i, jhi - 1)
            rows.append(LI(i, jlo)), cols.append(c), vals.append(-1.0)
    for i in range(1, ni - 1):
        if kinds[3] in ("neumann", "periodic"):
            c = LI(i, jhi - 1) if kinds[3] == "neumann" else LI(i, jlo + 1)
            rows.append(LI(i, jhi)), cols.append(c), vals.append(-1.0)
    A = sp.coo_matrix((vals, (rows, cols)), shape=(ni * nj, ni * nj)).tocsr()
    A.sort_indices()
    return A


def init_pattern_3d(dims, bcs):
    """Sparsity pattern of the 3D coefficient matrix (init_matrix.jl:174-416)."""
    ni, nj, nk = dims
    kinds = [_bc_kind(b) for b in bcs]
    for a in range(3):
        if (kinds[2 * a] == "periodic") != (kinds[2 * a + 1] == "periodic"):
            raise ValueError("Inconsistent periodic boundary conditions")
    n = (ni, nj, nk)
    LI = lambda i, j, k: i + ni * (j + nj * k)  # noqa: E731
    rows, cols, vals = [], [], []
    for k in range(nk):
        for j in range(nj):
            for i in range(ni):
                rows.append(LI(i, j, k)), cols.append(LI(i, j, k)), vals.append(1.0)
    for k in range(1, nk - 1):
        for j in range(1, nj - 1):
            for i in range(1, ni - 1):
                for dk in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        for di in (-1, 0, 1):
                            if di or dj or dk:
                                rows.append(LI(i, j, k)), cols.append(LI(i + di, j + dj, k + dk)), vals.append(0.0)
    # boundary faces (interior of each face only)
    for a in range(3):
        for side in (0, 1):
            kind = kinds[2 * a + side]
            if kind not in ("neumann", "periodic"):
                continue
            o = [ax for ax in range(3) if ax != a]
            for q in range(1, n[o[1]] - 1):
                for p in range(1, n[o[0]] - 1):
                    idx = [0, 0, 0]
                    idx[o[0]], idx[o[1]] = p, q
                    idx[a] = 0 if side == 0 else n[a] - 1
                    src = list(idx)
                    if kind == "neumann":
                        src[a] = 1 if side == 0 else n[a] - 2
                    else:
                        src[a] = n[a] - 2 if side == 0 else 1
                    rows.append(LI(*idx)), cols.append(LI(*src)), vals.append(-1.0)
    A = sp.coo_matrix((vals, (rows, cols)), shape=(ni * nj * nk,) * 2).tocsr()
    A.sort_indices()
    return A


def _mean(kind, a, b):
    if kind == "harmonic":
        with np.errstate(divide="ignore", invalid="ignore"):
            return (2.0 * a * b) / (a + b)
    return (a + b) / 2.0


class ImplicitOracle2D:
    """One implicit step on a 2D mesh with nhalo = 1 (the diffusion region is
    then the full array)."""

    def __init__(self, mesh, bcs, *, mean="harmonic", kappa=(1.0, 0.0, 0.0), cp=1.0, rho=None, source=None):
        if mesh.nhalo != 1:
            raise ValueError("the oracle covers meshes with one halo layer")
        self.mesh = mesh
        self.bcs = bcs
        self.mean = mean
        self.kappa = kappa
        self.cp = cp
        self.rho = rho
        self.source = source
        self.shape = mesh.dims_full
        self.pattern = init_pattern_2d(self.shape, bcs)
        em = mesh.edge_metric_list()
        self.fi = dict(Jxi_x=em[0], Jxi_y=em[1], Jeta_x=em[2], Jeta_y=em[3], J=em[4])
        self.fj = dict(Jxi_x=em[5], Jxi_y=em[6], Jeta_x=em[7], Jeta_y=em[8], J=em[9])
        self.J = mesh.cell_center_metrics.J

    def alpha(self, T):
        k0, te, re = self.kappa
        rho = np.ones(self.shape) if self.rho is None else self.rho
        with np.errstate(invalid="ignore", over="ignore"):
            kap = np.where(np.isfinite(T), k0 * np.abs(rho) ** re * np.abs(T) ** te, 0.0)
        return np.abs(kap / (rho * self.cp))

    def assemble(self, T, dt):
        """(A, b) of the step from the BC-applied temperature T."""
        NI, NJ = self.shape
        a = self.alpha(T)
        i = np.arange(1, NI - 1)[:, None]
        j = np.arange(1, NJ - 1)[None, :]
        mf = self.mean
        a_ip = _mean(mf, a[i, j], a[i + 1, j])
        a_im = _mean(mf, a[i, j], a[i - 1, j])
        a_jp = _mean(mf, a[i, j], a[i, j + 1])
        a_jm = _mean(mf, a[i, j], a[i, j - 1])
        m = {}
        for name in ("Jxi_x", "Jxi_y", "Jeta_x", "Jeta_y", "J"):
            m[f"{name}_ip"] = self.fi[name][i, j]
            m[f"{name}_im"] = self.fi[name][i - 1, j]
            m[f"{name}_jp"] = self.fj[name][i, j]
            m[f"{name}_jm"] = self.fj[name][i, j - 1]
        e = edge_terms_2d(a_ip, a_im, a_jp, a_jm, m)
        st = stencil_2d(e, self.J[i, j], dt)
        rows, cols, vals = [], [], []
        LI = lambda ii, jj: ii + NI * jj  # noqa: E731
        r = np.broadcast_to(LI(i, j), (NI - 2, NJ - 2)).ravel()
        for (di, dj), v in zip(OFF2, st):
            rows.append(r)
            cols.append(np.broadcast_to(LI(i + di, j + dj), (NI - 2, NJ - 2)).ravel())
            vals.append(np.broadcast_to(v, (NI - 2, NJ - 2)).ravel())
        b = np.zeros(NI * NJ)
        S = 0.0 if self.source is None else self.source[i, j]
        b[r] = ((S * dt + T[i, j]) * self.J[i, j]).ravel()
        # boundary rows: the pattern's identity / -1 entries
        P = self.pattern.tocoo()
        inner = np.zeros(NI * NJ, dtype=bool)
        inner[r] = True
        keep = ~inner[P.row]
        rows.append(P.row[keep])
        cols.append(P.col[keep])
        vals.append(P.data[keep])
        kinds = [_bc_kind(bb) for bb in self.bcs]
        vals_bc = [bb[1] if isinstance(bb, tuple) else 0.0 for bb in self.bcs]
        for jj in range(NJ):
            for ii in range(NI):
                if 0 < ii < NI - 1 and 0 < jj < NJ - 1:
                    continue
                f = 0 if ii == 0 else 1 if ii == NI - 1 else 2 if jj == 0 else 3
                b[LI(ii, jj)] = vals_bc[f] if kinds[f] == "dirichlet" else 0.0
        A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(NI * NJ, NI * NJ)).tocsc()
        return A, b

    def step(self, T, dt, *, cutoff=True, tol=1e-3, maximum_u_change=0.05, timestep_growth_factor=1.1):
        """Advance T (full array, modified in place) by dt; returns next_dt."""
        import apt_oracle as O

        helper = O.OracleSolver(self.mesh, self.bcs, mean=self.mean, kappa=self.kappa, cp=self.cp)
        helper.applybcs(T, 1)
        A, b = self.assemble(T, dt)
        x = spla.spsolve(A, b)
        if cutoff:
            x = 0.5 * (np.abs(x) + x) * np.isfinite(x)
        new = x.reshape(self.shape, order="F")
        nd = helper.next_dt(new, T, dt, tol=tol, maximum_u_change=maximum_u_change,
                            timestep_growth_factor=timestep_growth_factor)
        T[...] = new
        return nd
