"""Twilight-zone (manufactured solution) verification, the reference's
`tz=true` path (update.jl:23-71, residuals.jl:49-99, twilight_zone.jl).

The Gaussian exact solution of the reference's MMS driver
(test/integrated/2d/wavy_manufactured_solution/run.jl:14-51), steady (no_t)
with kappa = T^(5/2), on the wavy map: the discrete solution must converge to
it at second order in the mesh spacing (the order the reference's study
reports, wavy_manufactured_solution/run.jl:466-469).
"""
import math

import numpy as np
import pytest

import apt_oracle as O
from curvilineardiffusionmanuscript_amd import twilight_zone as tz
from curvilineardiffusionmanuscript_amd.mesh import CurvilinearGrid2D, wavy_map_2d
from curvilineardiffusionmanuscript_amd.solver import PowerLawConductivity

LAW = PowerLawConductivity(1.0, 2.5, 0.0)
CENTER, FWHM = (0.5, 0.5), 0.02


def _setup(n):
    mesh = CurvilinearGrid2D.from_mapping(wavy_map_2d(0.03), (n, n), 1)
    T = tz.gaussian_exact(mesh, 0.0, CENTER, FWHM, no_t=True)
    A, B = tz.gaussian_forcing(mesh, 0.0, CENTER, LAW, fwhm=FWHM, no_t=True)
    return mesh, T, A, B


def _oracle_error(n, steps=20, dt=2e-3):
    mesh, T, A, B = _setup(n)
    s = O.OracleSolver(mesh, [("neumann", 0.0)] * 4, kappa=(1.0, 2.5, 0.0))
    s.set_forcing(A, B)
    s.update_conductivity(T)
    for _ in range(steps):
        s.step(T, dt, rtol=0.0, atol=1e-9, max_iter=10**6, calculate_next_dt=False)
    ue = tz.gaussian_exact(mesh, 0.0, CENTER, FWHM, no_t=True)
    return float(np.max(np.abs(T - ue)[1:-1, 1:-1]))


def test_forcing_vanishes_for_exact_discrete_identity():
    # with alpha = alpha(u_e), A + B*alpha = -div(alpha grad u_e) - S exactly:
    # check against a direct finite-difference evaluation of div(alpha grad u_e)
    mesh, _, A, B = _setup(64)
    x, y = mesh.centroid_coordinates.x, mesh.centroid_coordinates.y
    h = 1e-4

    def flux_div(xx, yy):
        def u(a, b):
            return np.exp(-((a - CENTER[0]) ** 2 + (b - CENTER[1]) ** 2) / FWHM)

        def alpha(a, b):
            return u(a, b) ** 2.5

        fx = (alpha(xx + h / 2, yy) * (u(xx + h, yy) - u(xx, yy)) - alpha(xx - h / 2, yy) * (u(xx, yy) - u(xx - h, yy))) / h**2
        fy = (alpha(xx, yy + h / 2) * (u(xx, yy + h) - u(xx, yy)) - alpha(xx, yy - h / 2) * (u(xx, yy) - u(xx, yy - h))) / h**2
        return fx + fy

    ue = tz.gaussian_exact(mesh, 0.0, CENTER, FWHM, no_t=True)
    f = A + B * ue**2.5
    np.testing.assert_allclose(f, -flux_div(x, y), rtol=1e-5, atol=1e-4 * np.max(np.abs(f)))


def test_oracle_mms_second_order():
    e = [_oracle_error(n) for n in (32, 64, 128)]
    orders = [math.log2(e[i] / e[i + 1]) for i in range(2)]
    assert orders[-1] > 1.8, (e, orders)


@pytest.mark.gpu
def test_device_mms_matches_oracle_and_order():
    from curvilineardiffusionmanuscript_amd import solver as S

    errs = []
    for n in (32, 64, 128):
        mesh, T, A, B = _setup(n)
        dev = S.PseudoTransientSolver(mesh, [S.NeumannBC()] * 4)
        dev.set_forcing(A, B)
        ora = O.OracleSolver(mesh, [("neumann", 0.0)] * 4, kappa=(1.0, 2.5, 0.0))
        ora.set_forcing(A, B)
        Td, To = T.copy(order="F"), T.copy(order="F")
        S.update_conductivity(dev, mesh, Td, 1.0, 1.0, LAW)
        ora.update_conductivity(To)
        for _ in range(20):
            sd, _ = S.nonlinear_thermal_conduction_step(dev, mesh, Td, 1.0, 1.0, LAW, 2e-3, rtol=0.0, atol=1e-9,
                                                        max_iter=10**6, calculate_next_dt=False,
                                                        show_convergence=False)
            so, _ = ora.step(To, 2e-3, rtol=0.0, atol=1e-9, max_iter=10**6, calculate_next_dt=False)
            assert sd.niter == so["niter"]
        assert np.max(np.abs(Td - To)) <= 1e-12 * np.max(np.abs(To))
        ue = tz.gaussian_exact(mesh, 0.0, CENTER, FWHM, no_t=True)
        errs.append(float(np.max(np.abs(Td - ue)[1:-1, 1:-1])))
        dev.close()
    assert math.log2(errs[1] / errs[2]) > 1.8, errs


@pytest.mark.gpu
def test_3d_mms_mixed_precision_matches_fp64_and_oracle():
    # the forcing A + B alpha on the 3D mixed-precision path (the fused 3D
    # outer pass adds it to the FP64 residual): on a 3D wavy grid the steady
    # Gaussian of the MMS driver, three steps; each mixed-precision step ends
    # at the FP64 tolerance, and its temperature agrees with the FP64 APT's
    # (device and oracle, bit-identical to each other) to dt sqrt(N) atol
    from curvilineardiffusionmanuscript_amd import solver as S
    from curvilineardiffusionmanuscript_amd.mesh import CurvilinearGrid3D, wavy_map_3d

    mesh = CurvilinearGrid3D.from_mapping(wavy_map_3d(0.03), (24, 20, 18), 1)
    center, fwhm = (0.5, 0.5, 0.5), 0.05
    T = tz.gaussian_exact(mesh, 0.0, center, fwhm, no_t=True)
    A, B = tz.gaussian_forcing(mesh, 0.0, center, LAW, fwhm=fwhm, no_t=True)
    bcs = [S.NeumannBC()] * 6
    dt, atol = 2e-3, 1e-9
    out = {}
    for prec in ("fp32_dc", "fp64"):
        dev = S.PseudoTransientSolver(mesh, bcs)
        dev.set_forcing(A, B)
        Td = T.copy(order="F")
        S.update_conductivity(dev, mesh, Td, 1.0, 1.0, LAW)
        for _ in range(3):
            st, _ = S.nonlinear_thermal_conduction_step(dev, mesh, Td, 1.0, 1.0, LAW, dt, rtol=0.0, atol=atol,
                                                        max_iter=10**6, calculate_next_dt=False,
                                                        show_convergence=False, precision=prec)
            assert st.abs_err <= atol
            if prec == "fp32_dc":
                assert st.fp64_fallback == 0
        out[prec] = Td
        dev.close()
    ora = O.OracleSolver(mesh, [("neumann", 0.0)] * 6, kappa=(1.0, 2.5, 0.0))
    ora.set_forcing(A, B)
    To = T.copy(order="F")
    ora.update_conductivity(To)
    for _ in range(3):
        ora.step(To, dt, rtol=0.0, atol=atol, max_iter=10**6, calculate_next_dt=False)
    assert np.max(np.abs(out["fp64"] - To)) <= 1e-12 * np.max(np.abs(To))
    bound = 3 * dt * math.sqrt(T.size) * atol
    assert np.max(np.abs(out["fp32_dc"] - To)) <= bound
