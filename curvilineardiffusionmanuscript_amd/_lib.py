"""ctypes binding of libcdiff.so (include/cdiff.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
curvilineardiffusionmanuscript_amd/csrc``).  Loading fails loudly when it is
missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcdiff.so")

# cdiff_status
OK, NONFINITE, MAX_ITER, INVALID_ALPHA, INVALID_SOURCE, BAD_ARG, HIP_ERROR, NOT_READY = range(8)
# cdiff_bc_kind
BC_NEUMANN, BC_DIRICHLET, BC_PERIODIC, BC_FIXED_NEG_SYMMETRY = range(4)
MEAN_HARMONIC, MEAN_ARITHMETIC = 0, 1
PREC_FP64, PREC_FP32_DC = 0, 1
# cdiff_field
FIELD_U, FIELD_U_PREV, FIELD_RES, FIELD_ALPHA, FIELD_DTAU_RHO, FIELD_THETA_R_DTAU, FIELD_SOURCE = range(7)
FIELD_Q, FIELD_QPRIME, FIELD_CACHE, FIELD_DENSITY = 8, 12, 16, 20

EXPORTS = (
    "cdiff_default_opts",
    "cdiff_default_tuning",
    "cdiff_set_tuning",
    "cdiff_get_tuning",
    "cdiff_create",
    "cdiff_set_metrics",
    "cdiff_set_nodes",
    "cdiff_set_axisymmetric",
    "cdiff_set_kappa",
    "cdiff_set_density",
    "cdiff_set_source",
    "cdiff_set_forcing",
    "cdiff_update_conductivity",
    "cdiff_step",
    "cdiff_set_temperature",
    "cdiff_get_temperature",
    "cdiff_step_resident",
    "cdiff_apply_bcs",
    "cdiff_next_dt",
    "cdiff_next_dt_device",
    "cdiff_get_field",
    "cdiff_set_field",
    "cdiff_get_last_iter",
    "cdiff_set_last_iter",
    "cdiff_reset_state",
    "cdiff_bench_sweeps",
    "cdiff_last_error",
    "cdiff_status_string",
    "cdiff_destroy",
    "cdiff_net_num_params",
    "cdiff_net_create",
    "cdiff_net_destroy",
    "cdiff_net_set_params",
    "cdiff_net_get_params",
    "cdiff_net_dataset",
    "cdiff_net_capture",
    "cdiff_net_train",
    "cdiff_net_eval",
    "cdiff_net_loss_grad",
    "cdiff_net_predict",
    "cdiff_net_last_error",
    "cdiff_set_warm_start",
    "cdiff_net_set_gate",
    "cdiff_net_set_loss",
)


class BC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("value", C.c_double)]


class Kappa(C.Structure):
    _fields_ = [("kappa0", C.c_double), ("t_exp", C.c_double), ("rho_exp", C.c_double)]


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
        ("precision", C.c_int32),
        ("inner_iters", C.c_int32),
    ]


class Tuning(C.Structure):
    """cdiff_tuning: kernel variants and defect-correction parameters."""

    _fields_ = [(n, C.c_int32) for n in (
        "sweep2d", "tile_rows", "march_rows64", "pairs", "sweep3d", "planes3d", "seg_min", "seg_max",
        "dc_coefficient_form", "dc_newton", "dc_fused3d", "dc_depth2d", "dc_layout2d", "dc_march_rows",
        "dc_launch_iters", "dc_depth3d", "dc_planes3d", "dc_planes3d_pair", "dc_outer_planes3d", "dc_check",
        "dc_max_outer", "verbose", "exact_pow", "graphs", "warm_policy")] + [(n, C.c_double) for n in (
        "dc_eta", "dc_floor", "dc_cfl_scale", "dc_dtau_scale", "dc_theta_scale")]


class Stats(C.Structure):
    _fields_ = [
        ("rel_err", C.c_double),
        ("abs_err", C.c_double),
        ("niter", C.c_int32),
        ("nchecks", C.c_int32),
        ("time", C.c_double),
        ("outer", C.c_int32),
        ("warm_start", C.c_int32),
        ("fp64_fallback", C.c_int32),
        ("warm_consulted", C.c_int32),
    ]


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libcdiff.so and declare the prototypes of include/cdiff.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(
            f"{path} is not built: run __graft_entry__.build() (hipcc, gfx950); there is no CPU fallback"
        )
    L = C.CDLL(path)
    L.cdiff_default_opts.argtypes = [C.POINTER(Opts)]
    L.cdiff_default_opts.restype = None
    L.cdiff_default_tuning.argtypes = [C.POINTER(Tuning)]
    L.cdiff_default_tuning.restype = None
    L.cdiff_set_tuning.argtypes = [_vp, C.POINTER(Tuning)]
    L.cdiff_get_tuning.argtypes = [_vp, C.POINTER(Tuning)]
    L.cdiff_create.argtypes = [C.c_int, C.POINTER(C.c_int64), C.c_int, C.c_int, C.POINTER(BC), C.POINTER(Kappa), C.c_double, C.POINTER(_vp)]
    L.cdiff_set_metrics.argtypes = [_vp, C.POINTER(_dp), C.POINTER(_dp), C.c_double, _dp]
    L.cdiff_set_nodes.argtypes = [_vp, C.POINTER(_dp), C.c_double, _dp]
    L.cdiff_set_axisymmetric.argtypes = [_vp, _dp]
    L.cdiff_set_kappa.argtypes = [_vp, C.POINTER(Kappa), C.c_double]
    L.cdiff_set_density.argtypes = [_vp, _dp]
    L.cdiff_set_source.argtypes = [_vp, _dp]
    L.cdiff_set_forcing.argtypes = [_vp, _dp, _dp]
    L.cdiff_update_conductivity.argtypes = [_vp, _dp]
    L.cdiff_step.argtypes = [_vp, _dp, C.c_double, C.POINTER(Opts), C.POINTER(Stats), _dp]
    L.cdiff_set_temperature.argtypes = [_vp, _dp]
    L.cdiff_get_temperature.argtypes = [_vp, _dp]
    L.cdiff_step_resident.argtypes = [_vp, C.c_double, C.POINTER(Opts), C.POINTER(Stats), _dp]
    L.cdiff_apply_bcs.argtypes = [_vp, _dp]
    L.cdiff_next_dt.argtypes = [_vp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    L.cdiff_next_dt_device.argtypes = [_vp, _vp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, _vp,
                                       _vp, C.c_char_p, C.c_size_t]
    L.cdiff_get_field.argtypes = [_vp, C.c_int, _dp]
    L.cdiff_set_field.argtypes = [_vp, C.c_int, _dp]
    L.cdiff_get_last_iter.argtypes = [_vp, C.POINTER(C.c_int32)]
    L.cdiff_set_last_iter.argtypes = [_vp, C.c_int32]
    L.cdiff_reset_state.argtypes = [_vp]
    L.cdiff_bench_sweeps.argtypes = [_vp, C.c_double, C.c_int, C.c_int, C.c_int, _dp]
    L.cdiff_last_error.argtypes = [_vp]
    L.cdiff_last_error.restype = C.c_char_p
    L.cdiff_status_string.argtypes = [C.c_int]
    L.cdiff_status_string.restype = C.c_char_p
    L.cdiff_destroy.argtypes = [_vp]
    _fp = C.POINTER(C.c_float)
    L.cdiff_net_num_params.argtypes = []
    L.cdiff_net_create.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
    L.cdiff_net_destroy.argtypes = [_vp]
    L.cdiff_net_set_params.argtypes = [_vp, _fp, C.c_int]
    L.cdiff_net_get_params.argtypes = [_vp, _fp, C.c_int]
    L.cdiff_net_dataset.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.cdiff_net_capture.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_double]
    L.cdiff_net_train.argtypes = [_vp, C.c_int, C.c_int, C.c_float, C.c_uint64, _dp]
    L.cdiff_net_eval.argtypes = [_vp, C.c_int, _fp, _dp]
    L.cdiff_net_loss_grad.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_uint16), _fp, _fp, _fp, _dp]
    L.cdiff_net_predict.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_uint16), _fp]
    L.cdiff_net_last_error.argtypes = [_vp]
    L.cdiff_net_last_error.restype = C.c_char_p
    L.cdiff_set_warm_start.argtypes = [_vp, _vp]
    L.cdiff_net_set_gate.argtypes = [_vp, C.c_float]
    L.cdiff_net_set_loss.argtypes = [_vp, C.c_int]
    for name in EXPORTS:
        if name not in ("cdiff_default_opts", "cdiff_default_tuning", "cdiff_last_error", "cdiff_status_string",
                        "cdiff_net_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def ptr(a):
    """double* of a Fortran-contiguous float64 numpy array."""
    import numpy as np

    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
        raise TypeError("expected a Fortran-contiguous float64 array")
    return a.ctypes.data_as(_dp)
