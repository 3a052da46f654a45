"""VTK dumps of the device solver state (PseudoTransientSolver.jl:299-311,
380-412) through the C ABI's field reads."""
import glob
import os

import numpy as np
import pytest

from cases import gaussian_2d
from curvilineardiffusionmanuscript_amd import solver as S
from curvilineardiffusionmanuscript_amd import vtk as V

pytestmark = pytest.mark.gpu


def _setup(n=32):
    mesh, T0 = gaussian_2d(n)
    law = S.PowerLawConductivity(1.0, 2.0, 0.0)
    dev = S.PseudoTransientSolver(mesh, [S.NeumannBC()] * 4)
    T = T0.copy(order="F")
    S.update_conductivity(dev, mesh, T, 1.0, 1.0, law)
    return mesh, law, dev, T


def test_state_dump_after_a_step(tmp_path):
    mesh, law, dev, T = _setup()
    S.nonlinear_thermal_conduction_step(dev, mesh, T, 1.0, 1.0, law, 1e-4, rtol=0.0, atol=1e-8, max_iter=10**6,
                                        show_convergence=False)
    path = V.to_vtk(dev, mesh, dev.u, 1.0, 1, 1e-4, directory=str(tmp_path))
    _, cd, fd = V.read_vts(path)
    d = (slice(1, -1), slice(1, -1))
    np.testing.assert_array_equal(cd["u"], T[d].astype(np.float32))
    np.testing.assert_array_equal(cd["diffusivity"], dev.alpha[d].astype(np.float32))
    np.testing.assert_array_equal(cd["q1"], dev.q[0][d].astype(np.float32))
    np.testing.assert_array_equal(cd["q22"], dev.q_prime[1][d].astype(np.float32))
    np.testing.assert_array_equal(cd["residual"], dev.res[d].astype(np.float32))
    assert fd["TimeValue"][0] == 1e-4
    dev.close()


def test_failed_step_writes_the_diagnostic_dump(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    mesh, law, dev, T = _setup()
    with pytest.raises(S.MaxIterError, match="Maximum iteration limit"):
        S.nonlinear_thermal_conduction_step(dev, mesh, T, 1.0, 1.0, law, 1e-4, rtol=0.0, atol=1e-14, max_iter=20,
                                            show_convergence=False)
    files = glob.glob(os.path.join(str(tmp_path), "diffusion_*.vts"))
    assert len(files) == 1
    _, cd, _ = V.read_vts(files[0])
    assert np.all(np.isfinite(cd["u"]))
    dev.close()


def test_write_diagnostic_vtk_dumps_the_final_state(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    mesh, law, dev, T = _setup()
    st, _ = S.nonlinear_thermal_conduction_step(dev, mesh, T, 1.0, 1.0, law, 1e-4, rtol=0.0, atol=1e-8,
                                                max_iter=10**6, show_convergence=False, write_diagnostic_vtk=True)
    _, cd, _ = V.read_vts(os.path.join(str(tmp_path), V.get_filename(st.niter) + ".vts"))
    np.testing.assert_array_equal(cd["u"], T[1:-1, 1:-1].astype(np.float32))
    dev.close()
