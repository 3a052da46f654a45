"""Two pseudo-iterations per launch (temporal blocking) against one per launch.

The 2D sweep may run iterations t and t+1 in one kernel launch; its per-cell
arithmetic is the single-iteration kernel's, so every field after a step must
be bit-identical with the pairing switched off (tuning pairs = 0).  Covered: the
boundary-condition kinds that allow pairing, ragged tiles (sizes not multiples
of the 64 x 8 tile, domains narrower than a tile), frozen conductivity with
density and source, the manufactured-solution forcing, and the FP32 sweeps of
the defect-correction path (with and without the Newton term).
"""

import numpy as np
import pytest

from cases import gaussian_2d, hot_disk_2d
from curvilineardiffusionmanuscript_amd import solver as S
from curvilineardiffusionmanuscript_amd import twilight_zone as tz

pytestmark = pytest.mark.gpu


def _run(pairs, mesh, T0, bcs, law, nsteps, dt, *, rho=1.0, source=None, forcing=None, mean="harmonic", tuning=None,
         **opts):
    dev = S.PseudoTransientSolver(mesh, bcs, mean=mean, tuning=dict(tuning or {}, pairs=int(pairs)))
    if source is not None:
        dev.source_term = source
    if forcing is not None:
        dev.set_forcing(*forcing)
    T = T0.copy(order="F")
    S.update_conductivity(dev, mesh, T, rho, 1.0, law)
    stats = []
    for _ in range(nsteps):
        st, nd = S.nonlinear_thermal_conduction_step(dev, mesh, T, rho, 1.0, law, dt, show_convergence=False, **opts)
        stats.append((st.niter, st.abs_err, st.rel_err, nd))
    fields = {
        "T": T.copy(),
        "alpha": dev.alpha,
        "res": dev.res,
        "q0": dev.q[0],
        "q1": dev.q[1],
        "qp0": dev.q_prime[0],
        "qp1": dev.q_prime[1],
    }
    dev.close()
    return stats, fields


def _same(*args, fp32=False, **kw):
    s1, f1 = _run(False, *args, **kw)
    s2, f2 = _run(True, *args, **kw)
    assert s1 == s2
    keys = ["T"] if fp32 else list(f1)
    for k in keys:
        np.testing.assert_array_equal(f2[k], f1[k], err_msg=k)
    return s1


@pytest.mark.parametrize(
    "n,ni,bcs,k0",
    [
        (45, 70, [S.NeumannBC()] * 4, 1e-3),
        (13, 30, [S.DirichletBC(2.0), S.NeumannBC(), S.FixedNegSymmetryBC(), S.DirichletBC(0.5)], 1e-3),
        (64, 128, [S.FixedNegSymmetryBC(), S.DirichletBC(1.5), S.NeumannBC(), S.NeumannBC()], 1e-3),
        # 9 rows: one partial tile row; kappa0 = 1e-3 diverges here (in the oracle too)
        (9, 200, [S.NeumannBC(), S.NeumannBC(), S.DirichletBC(3.0), S.DirichletBC(1.0)], 1e-4),
    ],
)
def test_pairs_bit_identical_fp64(n, ni, bcs, k0):
    mesh, T0 = gaussian_2d(n, ni=ni)
    st = _same(mesh, T0, bcs, S.PowerLawConductivity(k0, 3.0, 0.0), 3, 1e-4, rtol=0.0, atol=1e-7,
               max_iter=100000)
    assert all(s[0] > 4 for s in st)


def test_pairs_bit_identical_nonlinear_adaptive():
    mesh, T0 = hot_disk_2d(80, seed=3)
    bcs = [S.DirichletBC(1.0), S.NeumannBC(), S.NeumannBC(), S.NeumannBC()]
    _same(mesh, T0, bcs, S.PowerLawConductivity(1.0, 2.5, 0.0), 3, 1e-5, rtol=0.0, atol=1e-8, max_iter=200000)


def test_pairs_bit_identical_frozen_density_source():
    mesh, T0 = gaussian_2d(36, ni=64)
    rng = np.random.default_rng(3)
    rho = np.asfortranarray(1.0 + 0.5 * rng.random(mesh.dims_full))
    src = np.zeros(mesh.dims_full, order="F")
    src[1:-1, 1:-1] = 50.0 * rng.random((64, 36))
    _same(mesh, T0, [S.NeumannBC()] * 4, S.PowerLawConductivity(1.5, 1.0, 1.0), 3, 2e-4, rho=rho, source=src,
          mean="arithmetic", rtol=1e-9, atol=1e-9, max_iter=100000, subcycle_conductivity=False,
          enforce_positivity=True, apply_cutoff=True)


def test_pairs_bit_identical_forcing():
    from curvilineardiffusionmanuscript_amd.mesh import CurvilinearGrid2D, wavy_map_2d

    mesh = CurvilinearGrid2D.from_mapping(wavy_map_2d(0.03), (48, 48), 1)
    law = S.PowerLawConductivity(1.0, 2.5, 0.0)
    T0 = tz.gaussian_exact(mesh, 0.0, (0.5, 0.5), 0.02, no_t=True)
    A, B = tz.gaussian_forcing(mesh, 0.0, (0.5, 0.5), law, fwhm=0.02, no_t=True)
    _same(mesh, T0, [S.NeumannBC()] * 4, law, 3, 2e-3, forcing=(A, B), rtol=0.0, atol=1e-9, max_iter=10**6,
          calculate_next_dt=False)


@pytest.mark.parametrize("newton", ["0", "1"])
def test_pairs_bit_identical_fp32_dc(newton):
    # the conductivity-form FP32 sweep (dc_coefficient_form = 0)
    mesh, T0 = hot_disk_2d(96, seed=0)
    bcs = [S.DirichletBC(1.0), S.NeumannBC(), S.NeumannBC(), S.NeumannBC()]
    _same(mesh, T0, bcs, S.PowerLawConductivity(1.0, 2.5, 0.0), 3, 1e-5, fp32=True, rtol=0.0, atol=1e-8,
          max_iter=10**6, precision="fp32_dc", calculate_next_dt=False,
          tuning={"dc_newton": int(newton), "dc_coefficient_form": 0})
