"""The reference's own integrated cases through the HIP path, with the numbers
its drivers hold (tests/golden/reference_integrated.json):

* sin(3 pi x) decay on uniform 50^2, 150^2 and 450^2 grids
  (test/integrated/2d/analytic/run.jl): FixedNegSymmetry faces in x, periodic
  in y, kappa = 1, harmonic mean, dt0 = 1e-7 adaptive with a 1e-4 cap, t =
  0.04.  The driver's L2 error (its own measure, sqrt(mean |u_e^2 - u^2|)) at
  each resolution must come within 5 % of the reference's logged value (its
  keyword arguments rel_tol / abs_tol / cutoff are not step! keywords and are
  ignored, so the step runs at the default tolerances, as here).  At 50^2
  the whole run is compared with the oracle as well (bit-identical).
* the planar nonlinear heat wave (test/integrated/2d/planar_nonlinear_wave/
  run.jl:26-150): 50^2 cells of the wavy grid x = xi + 0.1 sin(2 pi xi)
  sin(2 pi eta) (same for y), kappa = T^3, arithmetic mean, T = 1 at x = 0,
  T = 0 at x = 1, periodic in y, T0 = 1e-10, dt0 = 5e-8 adaptive, frozen
  conductivity per step, enforce_positivity, to t = 1.  Its first 200 steps
  are compared with the oracle (bit-identical); at t = 1 the profile along
  the first cell row must follow the analytic curve the driver plots
  (run.jl:180-237) to 0.025, and the front (where T falls through 1e-3,
  interpolated in log T) must lie within 1.5 cells (0.03) of the analytic
  front x = 0.870571 (run.jl:253).  Both bands are the discretisation error of
  a 50-cell grid with backward Euler steps: the oracle's own run lands at 0.016
  and 0.024.
"""
import json
import math
import os

import numpy as np
import pytest

import apt_oracle as O
from curvilineardiffusionmanuscript_amd import solver as S
from curvilineardiffusionmanuscript_amd.mesh import CurvilinearGrid2D, RectlinearGrid, wavy_map_2d

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_integrated.json")))


def _analytic_run(n, with_oracle=False):
    mesh = RectlinearGrid((0, 0), (1, 1), (n, n), 1)
    xc = mesh.centroid_coordinates.x
    dom = mesh.iterators.cell.domain
    T = np.zeros(mesh.dims_full, order="F")
    T[dom] = np.sin(3 * np.pi * xc[dom])
    bd = [S.FixedNegSymmetryBC(), S.FixedNegSymmetryBC(), S.PeriodicBC(), S.PeriodicBC()]
    dev = S.PseudoTransientSolver(mesh, bd, mean="harmonic")
    dev.alpha = np.ones(mesh.dims_full, order="F")  # fill!(solver.α, 1.0)
    law = S.PowerLawConductivity(1.0)
    ora = None
    if with_oracle:
        ora = O.OracleSolver(mesh, [("fixed_neg_symmetry", 0), ("fixed_neg_symmetry", 0), ("periodic", 0),
                                    ("periodic", 0)], mean="harmonic")
        ora.alpha[...] = 1.0
        To = T.copy(order="F")
    tfinal = GOLDEN["analytic_sine"]["tfinal"]
    dt, t, it = 1e-7, 0.0, 0
    while True:
        sd, nd = S.nonlinear_thermal_conduction_step(dev, mesh, T, 1.0, 1.0, law, dt, enforce_positivity=False,
                                                     subcycle_conductivity=False, show_convergence=False)
        if ora is not None:
            so, no = ora.step(To, dt, enforce_positivity=False, subcycle_conductivity=False)
            assert sd.niter == so["niter"], it
            assert nd == no, it
            np.testing.assert_array_equal(T, To)
        if t >= tfinal:
            break
        it += 1
        t += dt
        if math.isfinite(nd):
            dt = min(nd, 1e-4)
    sol = np.sin(3 * np.pi * xc[dom]) * np.exp(-9 * np.pi**2 * tfinal)
    sim = T[dom]
    dev.close()
    return math.sqrt(np.sum(np.abs(sol**2 - sim**2)) / sim.size), it + 1


@pytest.mark.parametrize("k", [0, 1, 2])
def test_analytic_sine_decay_matches_the_reference_errors(k):
    n = GOLDEN["analytic_sine"]["cells"][k]
    ref = GOLDEN["analytic_sine"]["L2"][k]
    L2, steps = _analytic_run(n, with_oracle=(n == 50))
    print(f"{n}^2: L2 = {L2:.6e} after {steps} steps; reference {ref:.6e} ({L2 / ref - 1:+.2%})")
    assert abs(L2 / ref - 1.0) <= 0.05


def _front(x, T, thr=1e-3):
    k = int(np.where(T < thr)[0][0])
    a, b = math.log(T[k - 1]), math.log(T[k])
    return x[k - 1] + (x[k] - x[k - 1]) * (a - math.log(thr)) / (a - b)


def test_planar_nonlinear_heat_wave():
    mesh = CurvilinearGrid2D.from_mapping(wavy_map_2d(0.1), (50, 50), 1)
    T = np.full(mesh.dims_full, 1e-10, order="F")
    bd = [S.DirichletBC(1.0), S.DirichletBC(0.0), S.PeriodicBC(), S.PeriodicBC()]
    bo = [("dirichlet", 1.0), ("dirichlet", 0.0), ("periodic", 0), ("periodic", 0)]
    law = S.PowerLawConductivity(1.0, 3.0, 0.0)
    dev = S.PseudoTransientSolver(mesh, bd, mean="arithmetic")
    ora = O.OracleSolver(mesh, bo, mean="arithmetic", kappa=(1.0, 3.0, 0.0))
    To = T.copy(order="F")
    S.update_conductivity(dev, mesh, T, 1.0, 1.0, law)
    ora.update_conductivity(To)
    kw = dict(apply_cutoff=False, calculate_next_dt=True, subcycle_conductivity=False, enforce_positivity=True)
    dt, t, it, iters = 5e-8, 0.0, 0, 0
    while True:
        sd, nd = S.nonlinear_thermal_conduction_step(dev, mesh, T, 1.0, 1.0, law, dt, show_convergence=False, **kw)
        iters += sd.niter
        if it < 200:
            so, no = ora.step(To, dt, **kw)
            assert sd.niter == so["niter"] and nd == no, it
            np.testing.assert_array_equal(T, To)
        if t >= 1.0:
            break
        it += 1
        t += dt
        if it % 2000 == 0:
            print(f"  step {it}: t = {t:.4f}, dt = {dt:.3e}", flush=True)
        if math.isfinite(nd):
            dt = nd
    dev.close()
    dom = mesh.iterators.cell.domain
    x = mesh.centroid_coordinates.x[dom][:, 0]
    T1 = T[dom][:, 0]
    g = GOLDEN["planar_wave"]
    xa, Ta = np.array(g["x"]), np.array(g["T"])
    sel = xa <= 0.846
    dev_profile = float(np.max(np.abs(np.interp(xa[sel], x, T1) - Ta[sel])))
    xf = _front(x, T1)
    print(f"{it + 1} steps to t = {t:.6f} ({iters} pseudo-iterations); max |T - analytic| for x <= 0.846: "
          f"{dev_profile:.4f}; front {xf:.4f} (analytic {g['front']})")
    assert dev_profile <= 0.025
    assert abs(xf - g["front"]) <= 0.03
