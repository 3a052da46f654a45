"""Pins of the implicit-direct oracle (oracle/implicit_oracle.py) against the
reference's own known answers, and its relation to the APT.

* edge terms: test/unit/test_edge_terms.jl:78-94 (values for metrics 1..20);
* unit-metric stencils: test/unit/test_inner_operator.jl:44,76-78;
* sparsity patterns: test/unit/test_matrix_construction.jl:21-37,70-117;
* uniform grid: the implicit system is the 5-point backward-Euler system, and
  the converged APT (conductivity frozen at T^n) reproduces the implicit
  solve to the iteration tolerance;
* wavy grid: the conservative (implicit) and non-conservative (APT)
  discretizations differ by O(h^2).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import apt_oracle as O
import implicit_oracle as I
from cases import gaussian_2d
from curvilineardiffusionmanuscript_amd.mesh import CurvilinearGrid2D, RectlinearGrid, wavy_map_2d


def _metrics(vals):
    names = ["J_ip", "J_im", "J_jp", "J_jm"]
    m = dict(zip(names, vals[:4]))
    order = [("Jxi_x", "ip"), ("Jxi_y", "ip"), ("Jeta_x", "ip"), ("Jeta_y", "ip"),
             ("Jxi_x", "im"), ("Jxi_y", "im"), ("Jeta_x", "im"), ("Jeta_y", "im"),
             ("Jxi_x", "jp"), ("Jxi_y", "jp"), ("Jeta_x", "jp"), ("Jeta_y", "jp"),
             ("Jxi_x", "jm"), ("Jxi_y", "jm"), ("Jeta_x", "jm"), ("Jeta_y", "jm")]
    for (n, f), v in zip(order, vals[4:]):
        m[f"{n}_{f}"] = v
    return m


def test_edge_terms_known_answers():
    # test_edge_terms.jl:26-94: alpha = 1, 2, 3, 4; J = 1, 2, 3, 4; metrics 5..20
    e = I.edge_terms_2d(1.0, 2.0, 3.0, 4.0, _metrics([1.0, 2.0, 3.0, 4.0] + [float(v) for v in range(5, 21)]))
    assert (e["xi2_ip"], e["xi2_im"], e["eta2_jp"], e["eta2_jm"]) == (61.0, 181.0, 481.0, 761.0)
    assert (e["xieta_ip"], e["xieta_im"], e["etaxi_jp"], e["etaxi_jm"]) == (20.75, 54.75, 104.75, 170.75)
    orth = [1.0, 2.0, 3.0, 4.0, 5.0, 0.0, 0.0, 8.0, 9.0, 0.0, 0.0, 12.0, 13.0, 0.0, 0.0, 16.0, 17.0, 0.0, 0.0, 20.0]
    e = I.edge_terms_2d(1.0, 2.0, 3.0, 4.0, _metrics(orth))
    assert (e["xi2_ip"], e["xi2_im"], e["eta2_jp"], e["eta2_jm"]) == (25.0, 81.0, 256.0, 400.0)
    assert (e["xieta_ip"], e["xieta_im"], e["etaxi_jp"], e["etaxi_jm"]) == (0.0, 0.0, 0.0, 0.0)


def test_unit_metric_stencils():
    one = _metrics([1.0] * 4 + [1.0, 0.0, 0.0, 1.0] * 4)
    e = I.edge_terms_2d(1.0, 1.0, 1.0, 1.0, one)
    assert I.stencil_2d(e, 1.0, 1.0) == [0, -1, 0, -1, 5, -1, 0, -1, 0]
    s3 = I.stencil_3d_unit([1.0] * 6, 1.0)
    ref = np.zeros(27)
    ref[[4, 10, 12, 14, 16, 22]] = -1
    ref[13] = 7
    np.testing.assert_array_equal(s3, ref)


def test_matrix_pattern_2d():
    m, n = 5, 6
    bcs = [("periodic", 0), ("periodic", 0), ("neumann", 0), ("neumann", 0)]
    with pytest.raises(ValueError):
        I.init_pattern_2d((m, n), [("neumann", 0), ("periodic", 0), ("neumann", 0), ("neumann", 0)])
    A = I.init_pattern_2d((m, n), bcs)
    LI = lambda i, j: (i - 1) + m * (j - 1)  # noqa: E731  (1-based like the reference)
    nz = lambda r: list(A.indices[A.indptr[r] : A.indptr[r + 1]])  # noqa: E731
    assert nz(LI(1, 4)) == sorted([LI(1, 4), LI(m - 1, 4)])
    assert nz(LI(m, 4)) == sorted([LI(2, 4), LI(m, 4)])
    assert [c + 1 for c in nz(LI(3, 3))] == [7, 8, 9, 12, 13, 14, 17, 18, 19]
    assert A.nnz == 140


def test_matrix_pattern_3d():
    dims = (5, 6, 7)
    bcs = [("periodic", 0), ("periodic", 0), ("neumann", 0), ("neumann", 0), ("periodic", 0), ("periodic", 0)]
    A = I.init_pattern_3d(dims, bcs)
    m = dims[0]
    LI = lambda i, j, k: (i - 1) + m * ((j - 1) + dims[1] * (k - 1))  # noqa: E731
    nz = lambda r: list(A.indices[A.indptr[r] : A.indptr[r + 1]])  # noqa: E731
    assert nz(LI(1, 4, 4)) == sorted([LI(1, 4, 4), LI(m - 1, 4, 4)])
    assert nz(LI(m, 4, 4)) == sorted([LI(2, 4, 4), LI(m, 4, 4)])
    assert [c + 1 for c in nz(LI(3, 3, 3))] == [37, 38, 39, 42, 43, 44, 47, 48, 49, 67, 68, 69, 72, 73, 74, 77, 78,
                                                79, 97, 98, 99, 102, 103, 104, 107, 108, 109]
    assert A.nnz == 1864


def _uniform(n):
    return RectlinearGrid((0.0, 0.0), (1.0, 1.0), (n, n))


def test_uniform_grid_is_the_five_point_system():
    n, dt, k = 24, 2e-3, 0.7
    mesh = _uniform(n)
    h = 1.0 / n
    rng = np.random.default_rng(0)
    T = np.asfortranarray(1.0 + rng.random(mesh.dims_full))
    bcs = [("dirichlet", 2.0), ("neumann", 0.0), ("neumann", 0.0), ("dirichlet", 0.5)]
    imp = I.ImplicitOracle2D(mesh, bcs, kappa=(k, 0.0, 0.0))
    Ti = T.copy(order="F")
    imp.step(Ti, dt, cutoff=False)
    # independent 5-point backward Euler with ghost-cell BCs (ghost = value or mirror)
    N = n
    idx = lambda i, j: i + N * j  # noqa: E731
    rows, cols, vals, b = [], [], [], np.zeros(N * N)
    r = k * dt / h**2
    for j in range(N):
        for i in range(N):
            p = idx(i, j)
            rows.append(p), cols.append(p), vals.append(1.0 + 4 * r)
            b[p] = T[i + 1, j + 1]
            for di, dj, face in ((-1, 0, 0), (1, 0, 1), (0, -1, 2), (0, 1, 3)):
                ii, jj = i + di, j + dj
                if 0 <= ii < N and 0 <= jj < N:
                    rows.append(p), cols.append(idx(ii, jj)), vals.append(-r)
                elif bcs[face][0] == "neumann":
                    rows.append(p), cols.append(p), vals.append(-r)  # mirror ghost: couples to the cell itself
                else:
                    b[p] += r * bcs[face][1]
    A = sp.coo_matrix((vals, (rows, cols)), shape=(N * N, N * N)).tocsc()
    x = spla.spsolve(A, b).reshape((N, N), order="F")
    np.testing.assert_allclose(Ti[1:-1, 1:-1], x, rtol=1e-12, atol=1e-12)


def test_converged_apt_matches_implicit_on_uniform_grid():
    n, dt = 32, 1e-3
    mesh = _uniform(n)
    rng = np.random.default_rng(1)
    T0 = np.asfortranarray(1.0 + rng.random(mesh.dims_full))
    bcs = [("neumann", 0.0), ("dirichlet", 1.5), ("neumann", 0.0), ("neumann", 0.0)]
    kappa = (1.0, 2.0, 0.0)
    Ti = T0.copy(order="F")
    I.ImplicitOracle2D(mesh, bcs, kappa=kappa).step(Ti, dt, cutoff=False)
    ora = O.OracleSolver(mesh, bcs, kappa=kappa)
    Ta = T0.copy(order="F")
    ora.update_conductivity(Ta)
    # conductivity frozen at T^n, as the implicit scheme lags it
    ora.step(Ta, dt, rtol=0.0, atol=1e-11, max_iter=10**6, subcycle_conductivity=False, apply_cutoff=False,
             calculate_next_dt=False)
    np.testing.assert_allclose(Ta[1:-1, 1:-1], Ti[1:-1, 1:-1], rtol=0, atol=1e-9)


def _apt_vs_implicit(mesh, T0, dt=2e-4):
    bcs = [("neumann", 0.0)] * 4
    Ti = T0.copy(order="F")
    I.ImplicitOracle2D(mesh, bcs).step(Ti, dt, cutoff=False)
    ora = O.OracleSolver(mesh, bcs)
    Ta = T0.copy(order="F")
    ora.update_conductivity(Ta)
    ora.step(Ta, dt, rtol=0.0, atol=1e-10, max_iter=10**6, calculate_next_dt=False, apply_cutoff=False)
    n = mesh.dims_full[0] - 2
    # central half of the domain: near the boundary the two schemes differ by
    # construction (the implicit corner ghosts are identity rows with rhs 0 that
    # its cross terms read; the APT zeroes the fluxes outside expand_lower)
    q = slice(1 + n // 4, 1 + 3 * n // 4)
    return float(np.max(np.abs(Ta - Ti)[q, q])), float(np.max(np.abs(Ta - T0)[q, q]))


def test_apt_equals_implicit_on_affine_grid():
    # constant metrics: the conservative and the non-conservative cross terms
    # are the same 9-point stencil, so the schemes agree to the APT tolerance
    from curvilineardiffusionmanuscript_amd.mesh import sheared_map_2d

    n = 64
    mesh = CurvilinearGrid2D.from_mapping(sheared_map_2d(0.3), (n, n), 1)
    x, y = mesh.centroid_coordinates.x, mesh.centroid_coordinates.y
    T0 = np.asfortranarray(1 + 10 * np.exp(-((x - 0.65) ** 2 + (y - 0.5) ** 2) / 0.005))
    d, inc = _apt_vs_implicit(mesh, T0)
    assert inc > 0.5 and d <= 1e-9 * inc, (d, inc)


def test_apt_and_implicit_differ_at_second_order_on_wavy_grid():
    errs = []
    for n in (32, 64, 128):
        mesh, T0 = gaussian_2d(n, amplitude=0.03)
        d, inc = _apt_vs_implicit(mesh, T0)
        errs.append(d / inc)
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert errs[-1] < 1e-4 and min(orders) > 1.7, (errs, orders)
