"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
tiny grids with randomized wavy metrics and every BC kind, two APT steps.
The CPU test re-runs the oracle from the stored inputs (regression of the
oracle); the GPU test runs libcdiff from the same stored arrays through the
C ABI, with no oracle involved."""
import glob
import os
from types import SimpleNamespace

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))
KIND_NAMES = {0: "neumann", 1: "dirichlet", 2: "periodic", 3: "fixed_neg_symmetry"}


class FixtureMesh:
    """The metric arrays of a fixture, in the shape the solvers read."""

    def __init__(self, z):
        self.dims_full = tuple(int(d) for d in z["dims"])
        self.ndim = len(self.dims_full)
        self.nhalo = 1
        nm = self.ndim * self.ndim
        self._cell = [np.asfortranarray(z[f"cell{i}"]) for i in range(nm)]
        self._edge = [np.asfortranarray(z[f"edge{i}"]) for i in range(self.ndim * (nm + 1))]
        self._L = float(z["L"])
        self._spacing = tuple(float(s) for s in z["spacing"])
        h = 1
        self.iterators = SimpleNamespace(
            cell=SimpleNamespace(
                domain=tuple(slice(h, d - h) for d in self.dims_full), full=tuple(slice(0, d) for d in self.dims_full)
            )
        )

    def cell_metric_list(self):
        return self._cell

    def edge_metric_list(self):
        return self._edge

    def phys_dims(self):
        return self._L, self._spacing


def _load(path):
    z = np.load(path, allow_pickle=False)
    mesh = FixtureMesh(z)
    bcs = [(KIND_NAMES[int(k)], float(v)) for k, v in z["bcs"]]
    return z, mesh, bcs


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    import apt_oracle as O

    z, mesh, bcs = _load(path)
    s = O.OracleSolver(mesh, bcs, kappa=tuple(z["kappa"]))
    T = np.asfortranarray(z["T0"].copy())
    s.update_conductivity(T)
    for k in (1, 2):
        st, nd = s.step(T, float(z["dt"]), rtol=0.0, atol=1e-10, max_iter=100000)
        assert st["niter"] == int(z[f"niter{k}"])
        np.testing.assert_array_equal(T, z[f"T{k}"])
        assert nd == float(z[f"next_dt{k}"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_device_reproduces_golden(path):
    from curvilineardiffusionmanuscript_amd import solver as S

    z, mesh, bcs = _load(path)
    dev_bcs = []
    for kind, v in bcs:
        dev_bcs.append({"neumann": S.NeumannBC(), "periodic": S.PeriodicBC(),
                        "fixed_neg_symmetry": S.FixedNegSymmetryBC()}.get(kind) or S.DirichletBC(v))
    law = S.PowerLawConductivity(*[float(x) for x in z["kappa"]])
    dev = S.PseudoTransientSolver(mesh, dev_bcs)
    T = np.asfortranarray(z["T0"].copy())
    S.update_conductivity(dev, mesh, T, 1.0, 1.0, law)
    exact = float(z["kappa"][1]).is_integer()
    for k in (1, 2):
        st, nd = S.nonlinear_thermal_conduction_step(dev, mesh, T, 1.0, 1.0, law, float(z["dt"]), rtol=0.0,
                                                     atol=1e-10, max_iter=100000, show_convergence=False)
        assert st.niter == int(z[f"niter{k}"])
        ref = z[f"T{k}"]
        if exact:
            np.testing.assert_array_equal(T, ref)
        else:
            assert np.max(np.abs(T - ref)) <= 1e-12 * np.max(np.abs(ref))
    for a in range(mesh.ndim):
        q = dev.q[a]
        if exact:
            np.testing.assert_array_equal(q, z["q_final"][a])
        else:
            np.testing.assert_allclose(q, z["q_final"][a], rtol=1e-10, atol=1e-14)
