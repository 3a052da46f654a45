"""ctypes front end of the CPU oracle (apt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product path never imports it.

``OracleSolver`` holds the state arrays of PseudoTransientSolver
(src/pseudo_transient/PseudoTransientSolver.jl:28-46) as Fortran-ordered
numpy arrays and calls the C restatement for each pass.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libapt_oracle.so")

OK, NONFINITE, MAX_ITER, INVALID_ALPHA, INVALID_SOURCE, BAD_ARG, INVALID_SCALAR = range(7)
BC_KIND = {"neumann": 0, "dirichlet": 1, "periodic": 2, "fixed_neg_symmetry": 3}

_dp = C.POINTER(C.c_double)


class _BC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("value", C.c_double)]


class _Solver(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("h", C.c_int32),
        ("n", C.c_int64 * 3),
        ("N", C.c_int64),
        ("lo", C.c_int64 * 3),
        ("hi", C.c_int64 * 3),
        ("mean", C.c_int32),
        ("last_err_check", C.c_int32),
        ("bcs", _BC * 6),
        ("L", C.c_double),
        ("spacing", C.c_double * 3),
        ("kappa0", C.c_double),
        ("t_exp", C.c_double),
        ("rho_exp", C.c_double),
        ("cp", C.c_double),
        ("t_exp_is_int", C.c_int32),
        ("rho_exp_is_int", C.c_int32),
        ("u", _dp),
        ("u_prev", _dp),
        ("S", _dp),
        ("res", _dp),
        ("alpha", _dp),
        ("theta", _dp),
        ("dtau", _dp),
        ("rho", _dp),
        ("q", _dp * 3),
        ("qp", _dp * 3),
        ("cell", _dp * 9),
        ("edge", _dp * 30),
        ("cache", _dp * 3),
        ("tzA", _dp),
        ("tzB", _dp),
        ("tmp_u", _dp),
        ("err", C.c_char * 512),
        ("rz_r", _dp),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32),
        ("norm_kind", C.c_int32),
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("cfl", C.c_double),
        ("apply_cutoff", C.c_int32),
        ("calculate_next_dt", C.c_int32),
        ("subcycle_conductivity", C.c_int32),
        ("enforce_positivity", C.c_int32),
        ("tol", C.c_double),
        ("maximum_u_change", C.c_double),
        ("timestep_growth_factor", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [("rel_err", C.c_double), ("abs_err", C.c_double), ("niter", C.c_int32), ("nchecks", C.c_int32), ("time", C.c_double)]


def default_opts(**kw) -> Opts:
    """step! keyword defaults (PseudoTransientSolver.jl:188-202, timestep_control.jl:33-42)."""
    o = dict(
        max_iter=1500,
        norm_kind=0,
        rtol=1e-5,
        atol=float(np.sqrt(np.finfo(np.float64).eps)),
        cfl=1.0 / np.sqrt(3.0),
        apply_cutoff=0,
        calculate_next_dt=1,
        subcycle_conductivity=1,
        enforce_positivity=0,
        tol=1e-3,
        maximum_u_change=0.05,
        timestep_growth_factor=1.1,
    )
    for k, v in kw.items():
        if k not in o:
            raise KeyError(k)
        o[k] = v
    return Opts(**{k: (int(v) if isinstance(v, bool) else v) for k, v in o.items()})


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(_Solver)
        L.ora_init.argtypes = [P]
        L.ora_applybcs.argtypes = [P, _dp, C.c_int]
        L.ora_update_conductivity.argtypes = [P, _dp]
        L.ora_update_iteration_params.argtypes = [P, C.c_double, C.c_double]
        L.ora_compute_flux.argtypes = [P]
        L.ora_compute_update.argtypes = [P, C.c_double]
        L.ora_update_residual.argtypes = [P, C.c_double]
        L.ora_l2_norm.argtypes = [P, _dp, C.c_int]
        L.ora_l2_norm.restype = C.c_double
        L.ora_cutoff.argtypes = [P, _dp]
        L.ora_next_dt.argtypes = [P, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.ora_next_dt.restype = C.c_double
        L.ora_validate_scalar.argtypes = [P, _dp, C.c_char_p, C.c_int]
        L.ora_update_metric_cache.argtypes = [P]
        L.ora_schedule.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ora_is_check.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ora_step.argtypes = [P, _dp, C.c_double, C.POINTER(Opts), C.POINTER(Stats), _dp]
        L.ora_num_threads.restype = C.c_int
        L.ora_set_num_threads.argtypes = [C.c_int]
        L.ora_kappa.argtypes = [P, C.c_double, C.c_double]
        L.ora_kappa.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(_dp)


def schedule(last_iter_count: int):
    m, k = C.c_int(), C.c_int()
    lib().ora_schedule(int(last_iter_count), C.byref(m), C.byref(k))
    return m.value, k.value


def is_check(it, m, k) -> bool:
    return bool(lib().ora_is_check(it, m, k))


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class OracleSolver:
    """CPU restatement of PseudoTransientSolver for one mesh (2D or 3D)."""

    def __init__(self, mesh, bcs, *, mean="harmonic", kappa=(1.0, 0.0, 0.0), cp=1.0, rho=None, source=None,
                 axisymmetric=False):
        """``bcs``: sequence of (kind, value) in (ilo, ihi, jlo, jhi[, klo, khi]) order.
        ``kappa``: (kappa0, t_exp, rho_exp) for kappa = kappa0 |rho|^rho_exp |T|^t_exp.
        ``axisymmetric``: 2D RZ geometry with x = r (not in the reference APT):
        the operator gains (1/r) dT/dr, r = the cell centroid's x."""
        L_ = lib()
        nd = mesh.ndim
        dims = mesh.dims_full
        self.mesh = mesh
        self.ndim = nd
        self.shape = tuple(dims)
        z = lambda: np.zeros(self.shape, dtype=np.float64, order="F")  # noqa: E731
        self.u, self.u_prev, self.S, self.res = z(), z(), z(), z()
        self.alpha, self.theta, self.dtau = z(), z(), z()
        self.rho = z() + 1.0 if rho is None else np.asfortranarray(rho, dtype=np.float64)
        if source is not None:
            self.S[...] = source
        self.q = [z() for _ in range(nd)]
        self.qp = [z() for _ in range(nd)]
        self.cache = [z() for _ in range(nd)]
        self.cell = [np.asfortranarray(a) for a in mesh.cell_metric_list()]
        self.edge = [np.asfortranarray(a) for a in mesh.edge_metric_list()]
        self.tzA = None
        self.tzB = None
        s = _Solver()
        s.ndim = nd
        s.h = mesh.nhalo
        for a in range(nd):
            s.n[a] = dims[a]
        s.mean = 0 if mean == "harmonic" else 1
        s.last_err_check = 0
        if len(bcs) != 2 * nd:
            raise ValueError("need one BC per face")
        for i, (kind, val) in enumerate(bcs):
            s.bcs[i].kind = BC_KIND[kind]
            s.bcs[i].value = float(val)
        L, spacing = mesh.phys_dims()
        s.L = L
        for a in range(nd):
            s.spacing[a] = spacing[a]
        k0, te, re = kappa
        s.kappa0, s.t_exp, s.rho_exp, s.cp = float(k0), float(te), float(re), float(cp)
        s.t_exp_is_int = int(float(te).is_integer() and 0 <= te <= 3)
        s.rho_exp_is_int = int(float(re).is_integer() and 0 <= re <= 3)
        for name in ("u", "u_prev", "S", "res", "alpha", "theta", "dtau", "rho"):
            setattr(s, name, _ptr(getattr(self, name)))
        for a in range(nd):
            s.q[a] = _ptr(self.q[a])
            s.qp[a] = _ptr(self.qp[a])
            s.cache[a] = _ptr(self.cache[a])
        for i, a in enumerate(self.cell):
            s.cell[i] = _ptr(a)
        for i, a in enumerate(self.edge):
            s.edge[i] = _ptr(a)
        self.rz_r = None
        if axisymmetric:
            if nd != 2:
                raise ValueError("axisymmetric geometry is 2D (x = r, y = z)")
            self.rz_r = np.asfortranarray(mesh.centroid_coordinates.x, dtype=np.float64)
            if np.min(self.rz_r[mesh.iterators.cell.domain]) <= 0.0:
                raise ValueError("axisymmetric mesh needs r > 0 at every domain cell centre")
            s.rz_r = _ptr(self.rz_r)
        self._s = s
        rc = L_.ora_init(C.byref(s))
        if rc:
            raise OracleError(rc, "bad arguments")

    # state
    @property
    def last_err_check(self):
        return self._s.last_err_check

    @last_err_check.setter
    def last_err_check(self, v):
        self._s.last_err_check = int(v)

    @property
    def L(self):
        return self._s.L

    @property
    def spacing(self):
        return tuple(self._s.spacing[a] for a in range(self.ndim))

    def set_forcing(self, A, B):
        self.tzA = np.asfortranarray(A, dtype=np.float64)
        self.tzB = np.asfortranarray(B, dtype=np.float64)
        self._s.tzA = _ptr(self.tzA)
        self._s.tzB = _ptr(self.tzB)

    def clear_forcing(self):
        self.tzA = self.tzB = None
        self._s.tzA = None
        self._s.tzB = None

    # passes
    def applybcs(self, u, nhalo=1):
        rc = lib().ora_applybcs(C.byref(self._s), _ptr(u), nhalo)
        if rc:
            raise OracleError(rc, "bad BC")

    def update_conductivity(self, T):
        lib().ora_update_conductivity(C.byref(self._s), _ptr(np.asfortranarray(T)))

    def update_iteration_params(self, Vpdtau, dt):
        lib().ora_update_iteration_params(C.byref(self._s), Vpdtau, dt)

    def compute_flux(self):
        lib().ora_compute_flux(C.byref(self._s))

    def compute_update(self, dt):
        lib().ora_compute_update(C.byref(self._s), dt)

    def update_residual(self, dt):
        lib().ora_update_residual(C.byref(self._s), dt)

    def l2_norm(self, a, kind=0):
        return lib().ora_l2_norm(C.byref(self._s), _ptr(a), kind)

    def cutoff(self, a):
        lib().ora_cutoff(C.byref(self._s), _ptr(a))

    def next_dt(self, unew, uold, dt, tol=1e-3, maximum_u_change=0.05, timestep_growth_factor=1.1):
        return lib().ora_next_dt(C.byref(self._s), _ptr(unew), _ptr(uold), dt, tol, maximum_u_change, timestep_growth_factor)

    def validate_scalar(self, a, name, positivity=False):
        rc = lib().ora_validate_scalar(C.byref(self._s), _ptr(a), name.encode(), int(positivity))
        if rc:
            raise OracleError(rc, self._s.err.decode())

    def update_metric_cache(self):
        lib().ora_update_metric_cache(C.byref(self._s))

    def kappa(self, rho, T):
        return lib().ora_kappa(C.byref(self._s), rho, T)

    def step(self, T: np.ndarray, dt: float, **opts):
        """One time step; T (full array, Fortran order) is overwritten like copy!(T, u)."""
        o = default_opts(**opts)
        st = Stats()
        nd = C.c_double(0.0)
        assert T.flags.f_contiguous and T.dtype == np.float64
        rc = lib().ora_step(C.byref(self._s), _ptr(T), float(dt), C.byref(o), C.byref(st), C.byref(nd))
        if rc:
            raise OracleError(rc, self._s.err.decode())
        return dict(rel_err=st.rel_err, abs_err=st.abs_err, niter=st.niter, nchecks=st.nchecks), nd.value


def set_num_threads(n: int):
    lib().ora_set_num_threads(int(n))


def num_threads() -> int:
    return lib().ora_num_threads()
