"""Host-side mirror of CurvilinearDiffusion.jl's pseudo-transient API over libcdiff.

Same names, argument meaning and error behaviour as the reference:

* ``PseudoTransientSolver(mesh, bcs; backend, mean=:harmonic, T=Float64)``
  (src/pseudo_transient/PseudoTransientSolver.jl:48-103)
* ``nonlinear_thermal_conduction_step!(solver, mesh, T, rho, cp, kappa, dt; kwargs...)``
  -> ``(stats, next_dt)`` (src/nonlinear_thermal_conduction.jl:51-72)
* ``update_conductivity!(solver, mesh, T, rho, cp, kappa)`` (src/conductivity.jl:15-30)
* ``applybcs!(bcs, mesh, u, nhalo)`` (src/boundary_conditions/boundary_operators.jl:26-30)
* ``next_dt(u_new, u_prev, dt; tol, maximum_u_change, timestep_growth_factor)``
  (src/timestep_control.jl:33-54)
* BC types ``NeumannBC``, ``DirichletBC``, ``PeriodicBC``, ``FixedNegSymmetryBC``
  (boundary_operators.jl:11-17)

Julia's ``!`` cannot appear in a Python name: the mutating functions drop it
(``nonlinear_thermal_conduction_step``) and keep their in-place behaviour: the
temperature array passed in is overwritten with the new time level, exactly
as ``copy!(T, solver.u)`` (PseudoTransientSolver.jl:346).

Every compute call goes to the HIP library; nothing here evaluates the solver
on the CPU.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from types import SimpleNamespace
from typing import Any, Sequence

import numpy as np

from . import _lib as L


# ----------------------------------------------------------------------------
# errors: Julia error() / @assert of the reference, one class per status


class CurvilinearDiffusionError(RuntimeError):
    """error(...) raised by the reference solver."""

    status = None


class NonFiniteError(CurvilinearDiffusionError):
    status = L.NONFINITE


class MaxIterError(CurvilinearDiffusionError):
    status = L.MAX_ITER


class InvalidScalarError(CurvilinearDiffusionError):
    status = L.INVALID_ALPHA


class SolverAssertionError(AssertionError):
    status = L.BAD_ARG


class DeviceError(RuntimeError):
    status = L.HIP_ERROR


_STATUS_EXC = {
    L.NONFINITE: NonFiniteError,
    L.MAX_ITER: MaxIterError,
    L.INVALID_ALPHA: InvalidScalarError,
    L.INVALID_SOURCE: InvalidScalarError,
    L.BAD_ARG: SolverAssertionError,
    L.HIP_ERROR: DeviceError,
    L.NOT_READY: CurvilinearDiffusionError,
}


def _check(rc: int, handle=None):
    if rc == L.OK:
        return
    lib = L.load()
    msg = lib.cdiff_last_error(handle).decode() if handle else lib.cdiff_status_string(rc).decode()
    exc = _STATUS_EXC.get(rc, CurvilinearDiffusionError)(msg)
    exc.status = rc
    raise exc


# ----------------------------------------------------------------------------
# boundary conditions (boundary_operators.jl:11-17)


class NeumannBC:
    kind = L.BC_NEUMANN
    value = 0.0

    def __repr__(self):
        return "NeumannBC()"


class PeriodicBC:
    kind = L.BC_PERIODIC
    value = 0.0

    def __repr__(self):
        return "PeriodicBC()"


class FixedNegSymmetryBC:
    kind = L.BC_FIXED_NEG_SYMMETRY
    value = 0.0

    def __repr__(self):
        return "FixedNegSymmetryBC()"


class DirichletBC:
    """``DirichletBC(val)``.  The reference reads ``bc.val.u`` in 2D and ``bc.val``
    in 3D (boundary_operators.jl:144,185); both spellings are accepted here."""

    kind = L.BC_DIRICHLET

    def __init__(self, val):
        if hasattr(val, "u"):
            v = val.u
        elif isinstance(val, dict):
            v = val["u"]
        else:
            v = val
        self.val = val
        self.value = float(v)

    def __repr__(self):
        return f"DirichletBC({self.val!r})"


_LOCS = ("ilo", "ihi", "jlo", "jhi", "klo", "khi")


def _bc_list(bcs, ndim) -> list:
    """BCs in NamedTuple position order (boundary_operators.jl:19-28)."""
    if isinstance(bcs, dict):
        vals = list(bcs.values())
    else:
        vals = list(bcs)
    if len(vals) != 2 * ndim:
        raise ValueError(f"{ndim}D mesh needs {2 * ndim} boundary conditions")
    return vals


# ----------------------------------------------------------------------------
# conductivity law


@dataclass(frozen=True)
class PowerLawConductivity:
    """kappa(rho, T) = kappa0 * |rho|^rho_exp * |T|^t_exp, 0 for non-finite T.

    The drivers' closures (test/integrated/2d/gaussian/run.jl:85-91,129-135:
    2.5 and 2.5*abs(T)^3; wavy_manufactured_solution/run.jl: T^(5/2);
    planar_nonlinear_wave/run.jl:68: T^3) are all of this form."""

    kappa0: float = 1.0
    t_exp: float = 0.0
    rho_exp: float = 0.0

    def __call__(self, rho, T):
        T = np.asarray(T, dtype=np.float64)
        rho = np.asarray(rho, dtype=np.float64)
        with np.errstate(all="ignore"):
            v = self.kappa0 * np.abs(rho) ** self.rho_exp * np.abs(T) ** self.t_exp
        return np.where(np.isfinite(T), v, 0.0)

    def _c(self):
        return L.Kappa(float(self.kappa0), float(self.t_exp), float(self.rho_exp))


def _as_law(kappa) -> PowerLawConductivity:
    if isinstance(kappa, PowerLawConductivity):
        return kappa
    if isinstance(kappa, (int, float)):
        return PowerLawConductivity(float(kappa), 0.0, 0.0)
    if isinstance(kappa, (tuple, list)) and 1 <= len(kappa) <= 3:
        return PowerLawConductivity(*[float(x) for x in kappa])
    raise TypeError(
        "kappa must be a PowerLawConductivity (kappa0*|rho|^a*|T|^b); arbitrary callables cannot cross the C ABI"
    )


# ----------------------------------------------------------------------------


def _fortran64(a) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.float64)


_STEP_KW = {
    "max_iter": 1500,
    "rtol": 1e-5,
    "atol": math.sqrt(np.finfo(np.float64).eps),
    "error_check_interval": 2,  # accepted and unused, as in the reference (:195)
    "apply_cutoff": False,
    "calculate_next_dt": True,
    "subcycle_conductivity": True,
    "write_diagnostic_vtk": False,
    "enforce_positivity": False,
    "CFL": 1.0 / math.sqrt(3.0),
    "tol": 1e-3,
    "maximum_u_change": 0.05,
    "timestep_growth_factor": 1.1,
    "norm": "cpu",  # L2_norm flavour: "cpu" RMS (residuals.jl:8-22) or "gpu" Inf/sqrt(N) (:1-6)
    "precision": "fp64",  # "fp64" or "fp32_dc"
    "inner_iters": 0,
}


class PseudoTransientSolver:
    """Device-resident APT solver state (PseudoTransientSolver.jl:28-46)."""

    def __init__(self, mesh, bcs, *, backend=None, mean="harmonic", T=np.float64, metrics=None, axisymmetric=False,
                 tuning=None, **kwargs):
        """``metrics``: "device" builds the metric arrays on the GPU from the
        mesh's ghost-inclusive nodes (``mesh.nodes_full()``, the default when the
        mesh has them), "host" uploads the mesh's own metric arrays.
        ``axisymmetric``: 2D RZ geometry with x = r, y = z (the operator gains
        (1/r) dT/dr; r = the cell centroid's x); not in the reference APT.
        ``tuning``: overrides of the cdiff_tuning fields (kernel variants and
        defect-correction parameters, see ``set_tuning``)."""
        if T not in (np.float64, float, "Float64"):
            raise TypeError("the drop-in path keeps the reference's Float64 state (see precision= for FP32 sweeps)")
        lib = L.load()
        self.mesh = mesh
        self.ndim = mesh.ndim
        self.bcs = bcs
        self._bcl = _bc_list(bcs, self.ndim)
        self.mean = mean
        self.shape = tuple(mesh.dims_full)
        self.nhalo = mesh.nhalo
        arr = (L.BC * 6)()
        for i, b in enumerate(self._bcl):
            arr[i].kind = int(b.kind)
            arr[i].value = float(getattr(b, "value", 0.0))
        dims = (C.c_int64 * 3)(*(list(self.shape) + [1] * (3 - self.ndim)))
        h = C.c_void_p()
        self._law = PowerLawConductivity()
        self._cp = 1.0
        kap = self._law._c()
        rc = lib.cdiff_create(
            self.ndim, dims, self.nhalo, L.MEAN_HARMONIC if mean in ("harmonic", ":harmonic") else L.MEAN_ARITHMETIC,
            arr, C.byref(kap), 1.0, C.byref(h)
        )
        self._h = h
        if rc != L.OK:
            try:
                _check(rc, h if h.value else None)
            finally:
                self.close()
        self.L, self.spacing = mesh.phys_dims()
        sp = (C.c_double * 3)(*(list(self.spacing) + [0.0] * (3 - self.ndim)))
        self.axisymmetric = bool(axisymmetric)
        if self.axisymmetric:
            if self.ndim != 2:
                self.close()
                raise ValueError("axisymmetric geometry is 2D (x = r, y = z)")
            r = _fortran64(mesh.centroid_coordinates.x)
            _check(lib.cdiff_set_axisymmetric(h, L.ptr(r)), h)
        if metrics is None:
            metrics = "device" if hasattr(mesh, "nodes_full") else "host"
        if metrics == "device":
            nodes = [_fortran64(a) for a in mesh.nodes_full()]
            want = tuple(d + 1 for d in self.shape)
            if any(a.shape != want for a in nodes):
                raise ValueError(f"node arrays must have shape {want}")
            npp = (L._dp * len(nodes))(*[L.ptr(a) for a in nodes])
            _check(lib.cdiff_set_nodes(h, npp, float(self.L), sp), h)
        elif metrics == "host":
            cells = [_fortran64(a) for a in mesh.cell_metric_list()]
            edges = [_fortran64(a) for a in mesh.edge_metric_list()]
            cp = (L._dp * len(cells))(*[L.ptr(a) for a in cells])
            ep = (L._dp * len(edges))(*[L.ptr(a) for a in edges])
            _check(lib.cdiff_set_metrics(h, cp, ep, float(self.L), sp), h)
        else:
            raise ValueError('metrics must be "device" or "host"')
        self.metrics = metrics
        if tuning:
            self.set_tuning(**tuning)
        self._rho_id = None
        self.iterators = SimpleNamespace(
            domain=SimpleNamespace(cartesian=mesh.iterators.cell.domain),
            full=SimpleNamespace(cartesian=mesh.iterators.cell.full),
        )

    # -- tuning ----------------------------------------------------------------
    def get_tuning(self) -> dict:
        """The solver's cdiff_tuning as a dict."""
        t = L.Tuning()
        _check(L.load().cdiff_get_tuning(self._h, C.byref(t)), self._h)
        return {n: getattr(t, n) for n, _ in L.Tuning._fields_}

    def set_tuning(self, **overrides):
        """Change fields of the solver's cdiff_tuning (no counterpart in the
        reference): kernel variants that give the same bits (sweep2d,
        tile_rows, pairs, dc_layout2d, dc_depth2d, dc_launch_iters, ...) and the
        defect-correction parameters (dc_eta, dc_cfl_scale, ...).  Unknown
        names raise; values out of range raise SolverAssertionError."""
        t = L.Tuning()
        _check(L.load().cdiff_get_tuning(self._h, C.byref(t)), self._h)
        names = {n for n, _ in L.Tuning._fields_}
        for k, v in overrides.items():
            if k not in names:
                raise TypeError(f"unknown tuning field {k!r}")
            setattr(t, k, v)
        _check(L.load().cdiff_set_tuning(self._h, C.byref(t)), self._h)

    # -- lifetime --------------------------------------------------------------
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.load().cdiff_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state ------------------------------------------------------------------
    def _field(self, f):
        out = np.zeros(self.shape, dtype=np.float64, order="F")
        _check(L.load().cdiff_get_field(self._h, f, L.ptr(out)), self._h)
        return out

    @property
    def u(self):
        return self._field(L.FIELD_U)

    @property
    def u_prev(self):
        return self._field(L.FIELD_U_PREV)

    @property
    def res(self):
        return self._field(L.FIELD_RES)

    @property
    def alpha(self):
        return self._field(L.FIELD_ALPHA)

    @alpha.setter
    def alpha(self, a):
        _check(L.load().cdiff_set_field(self._h, L.FIELD_ALPHA, L.ptr(_fortran64(a))), self._h)

    @property
    def dtau_rho(self):
        return self._field(L.FIELD_DTAU_RHO)

    @property
    def theta_r_dtau(self):
        return self._field(L.FIELD_THETA_R_DTAU)

    @property
    def q(self):
        return tuple(self._field(L.FIELD_Q + a) for a in range(self.ndim))

    def set_flux(self, axis, values):
        _check(L.load().cdiff_set_field(self._h, L.FIELD_Q + axis, L.ptr(_fortran64(values))), self._h)

    @property
    def q_prime(self):
        return tuple(self._field(L.FIELD_QPRIME + a) for a in range(self.ndim))

    @property
    def cache(self):
        return tuple(self._field(L.FIELD_CACHE + a) for a in range(self.ndim))

    @property
    def source_term(self):
        return self._field(L.FIELD_SOURCE)

    @source_term.setter
    def source_term(self, S):
        """``copy!(solver.source_term, S)`` (test/integrated/2d/gaussian/run.jl:166)."""
        _check(L.load().cdiff_set_source(self._h, None if S is None else L.ptr(_fortran64(S))), self._h)

    def set_forcing(self, A, B):
        """Twilight-zone forcing f = A + B*alpha (update.jl:23-71)."""
        lib = L.load()
        if A is None:
            _check(lib.cdiff_set_forcing(self._h, None, None), self._h)
        else:
            _check(lib.cdiff_set_forcing(self._h, L.ptr(_fortran64(A)), L.ptr(_fortran64(B))), self._h)

    @property
    def last_err_check(self):
        v = C.c_int32()
        _check(L.load().cdiff_get_last_iter(self._h, C.byref(v)), self._h)
        return [v.value]

    @last_err_check.setter
    def last_err_check(self, v):
        v = v[0] if isinstance(v, (list, tuple)) else v
        _check(L.load().cdiff_set_last_iter(self._h, int(v)), self._h)

    def reset_state(self):
        _check(L.load().cdiff_reset_state(self._h), self._h)

    # -- device-resident stepping (temperature kept in HBM) -----------------------
    def set_temperature(self, T):
        _check(L.load().cdiff_set_temperature(self._h, L.ptr(_fortran64(T))), self._h)

    def get_temperature(self):
        out = np.zeros(self.shape, dtype=np.float64, order="F")
        _check(L.load().cdiff_get_temperature(self._h, L.ptr(out)), self._h)
        return out

    def _prepare(self, rho, cp, kappa):
        lib = L.load()
        law = _as_law(kappa)
        if law != self._law or float(cp) != self._cp:
            k = law._c()
            _check(lib.cdiff_set_kappa(self._h, C.byref(k), float(cp)), self._h)
            self._law, self._cp = law, float(cp)
        if rho is not None:
            # the density on the device is replaced whenever the values passed
            # differ from the last upload (compared by content: the caller may
            # edit its array in place between steps)
            if np.isscalar(rho):
                key = float(rho)
                if key != self._rho_id:
                    if key == 1.0:
                        _check(lib.cdiff_set_density(self._h, None), self._h)
                    else:
                        r = np.full(self.shape, key, order="F")
                        _check(lib.cdiff_set_density(self._h, L.ptr(r)), self._h)
                    self._rho_id = key
            else:
                r = _fortran64(rho)
                last = self._rho_id
                if not (isinstance(last, np.ndarray) and last.shape == r.shape and np.array_equal(last, r)):
                    _check(lib.cdiff_set_density(self._h, L.ptr(r)), self._h)
                    self._rho_id = r.copy(order="F")

    def step_resident(self, dt, rho=None, cp=1.0, kappa=1.0, **kwargs):
        """One step on the device-resident temperature; returns (stats, next_dt)."""
        self._prepare(rho, cp, kappa)
        o = _opts(**kwargs)
        st = L.Stats()
        nd = C.c_double(0.0)
        rc = L.load().cdiff_step_resident(self._h, float(dt), C.byref(o), C.byref(st), C.byref(nd))
        _check(rc, self._h)
        return _stats(st), nd.value


def _opts(**kwargs) -> L.Opts:
    kw = dict(_STEP_KW)
    for k, v in kwargs.items():
        if k in kw:
            kw[k] = v
        # unknown keywords are accepted and ignored, as the reference forwards
        # them to next_dt, which drops them (timestep_control.jl:41)
    o = L.Opts()
    L.load().cdiff_default_opts(C.byref(o))
    o.max_iter = int(min(kw["max_iter"], 2**31 - 2)) if kw["max_iter"] != math.inf else 2**31 - 2
    o.rtol = float(kw["rtol"])
    o.atol = float(kw["atol"])
    o.cfl = float(kw["CFL"])
    o.apply_cutoff = int(bool(kw["apply_cutoff"]))
    o.calculate_next_dt = int(bool(kw["calculate_next_dt"]))
    o.subcycle_conductivity = int(bool(kw["subcycle_conductivity"]))
    o.enforce_positivity = int(bool(kw["enforce_positivity"]))
    o.tol = float(kw["tol"])
    o.maximum_u_change = float(kw["maximum_u_change"])
    o.timestep_growth_factor = float(kw["timestep_growth_factor"])
    o.norm_kind = 0 if kw["norm"] in ("cpu", "rms", 0) else 1
    o.precision = L.PREC_FP64 if kw["precision"] in ("fp64", 0) else L.PREC_FP32_DC
    o.inner_iters = int(kw["inner_iters"])
    return o


def _stats(st: L.Stats):
    return SimpleNamespace(
        rel_err=st.rel_err, abs_err=st.abs_err, niter=st.niter, time=st.time, nchecks=st.nchecks, outer=st.outer,
        warm_start=bool(st.warm_start), fp64_fallback=bool(st.fp64_fallback), warm_consulted=bool(st.warm_consulted),
    )


def update_conductivity(solver: PseudoTransientSolver, mesh, T, rho, cp, kappa):
    """update_conductivity!(solver, mesh, T, rho, cp, kappa) (src/conductivity.jl:15-30)."""
    solver._prepare(rho, cp, kappa)
    _check(L.load().cdiff_update_conductivity(solver._h, L.ptr(_fortran64(T))), solver._h)


def _twilight_zone(solver: PseudoTransientSolver, mesh, rho, cp, t, tz, no_t):
    """``tz=true`` of step! (PseudoTransientSolver.jl:189-191,276,281): the
    update and the residual gain the forcing of the reference's manufactured
    solution u = exp(-|x|^2/fwhm) exp(-damp t), fwhm = 1, damp = 4, centred at
    the origin (twilight_zone.jl:101-106), at time ``t`` (``no_t``: without the
    time factor).  The reference hard-codes its wavy test grid and kappa =
    T^(5/2) (twilight_zone.jl:86-118,180-189); here the forcing is built for the
    mesh and the conductivity law of the step.  ``tz=false`` removes a forcing
    that an earlier ``tz=true`` step set (one set with ``set_forcing`` stays)."""
    if tz:
        from .twilight_zone import gaussian_forcing

        r = 1.0 if rho is None else (float(rho) if np.isscalar(rho) else _fortran64(rho))
        A, B = gaussian_forcing(
            mesh, float(t), (0.0,) * solver.ndim, solver._law, fwhm=1.0, damp=4.0, no_t=bool(no_t), rho=r,
            cp=float(cp), source=solver.source_term, axisymmetric=solver.axisymmetric,
        )
        solver.set_forcing(A, B)
        solver._tz_forcing = True
    elif getattr(solver, "_tz_forcing", False):
        solver.set_forcing(None, None)
        solver._tz_forcing = False


def nonlinear_thermal_conduction_step(
    solver: PseudoTransientSolver, mesh, T, rho, cp, kappa, dt, *, t=0, tz=False, tz_dim=1, no_t=True,
    show_convergence=True, **kwargs
):
    """nonlinear_thermal_conduction_step!(::PseudoTransientSolver, ...) -> (stats, next_dt)
    (src/nonlinear_thermal_conduction.jl:51-72).  ``T`` is overwritten in place.

    ``t``, ``tz``, ``no_t``: the twilight-zone forcing of the reference's
    manufactured solution (see ``_twilight_zone``).  ``tz_dim`` is accepted and
    unused, as in the reference's wavy-grid forcing (twilight_zone.jl:54-194)."""
    if not isinstance(T, np.ndarray) or T.dtype != np.float64:
        raise TypeError("T must be a float64 numpy array of the full (halo-inclusive) mesh size")
    if T.shape != solver.shape:
        raise ValueError(f"T has shape {T.shape}, mesh needs {solver.shape}")
    solver._prepare(rho, cp, kappa)
    _twilight_zone(solver, mesh, rho, cp, t, tz, no_t)
    o = _opts(**kwargs)
    st = L.Stats()
    nd = C.c_double(0.0)
    buf = T if T.flags.f_contiguous else np.asfortranarray(T)
    rc = L.load().cdiff_step(solver._h, L.ptr(buf), float(dt), C.byref(o), C.byref(st), C.byref(nd))
    # the reference dumps the state before raising on a non-finite error or the
    # iteration limit (PseudoTransientSolver.jl:302-314); with
    # write_diagnostic_vtk it dumps at every error check (:298-300), which the
    # device loop does not stop for: here the step's final state is dumped once
    if rc in (L.NONFINITE, L.MAX_ITER) or (rc == L.OK and kwargs.get("write_diagnostic_vtk")):
        from .vtk import to_vtk

        to_vtk(solver, mesh, solver.u, 1.0 if rho is None else rho, st.niter, st.niter)
    _check(rc, solver._h)
    if buf is not T:
        T[...] = buf
    stats = _stats(st)
    if show_convergence:
        print(f"\trel error: {stats.rel_err:.3e}, abs err: {stats.abs_err:.3e}, {stats.niter}")
    return stats, nd.value


def applybcs(bcs, mesh, u, nhalo=1, solver: PseudoTransientSolver | None = None):
    """applybcs!(bcs, mesh, u, nhalo) on the device; ``u`` is overwritten."""
    own = solver is None
    if own:
        solver = PseudoTransientSolver(mesh, bcs)
    try:
        buf = u if u.flags.f_contiguous else np.asfortranarray(u)
        _check(L.load().cdiff_apply_bcs(solver._h, L.ptr(buf)), solver._h)
        if buf is not u:
            u[...] = buf
    finally:
        if own:
            solver.close()


def next_dt(u_new, u_prev, dt, *, solver: PseudoTransientSolver, tol=1e-3, maximum_u_change=0.05,
            timestep_growth_factor=1.1, **kwargs):
    """next_dt(u_new, u_prev, dt; ...) (src/timestep_control.jl:33-54) on the device."""
    out = C.c_double()
    rc = L.load().cdiff_next_dt(
        solver._h, L.ptr(_fortran64(u_new)), L.ptr(_fortran64(u_prev)), float(dt), float(tol),
        float(maximum_u_change), float(timestep_growth_factor), C.byref(out)
    )
    _check(rc, solver._h)
    return out.value
