"""cdiff_next_dt_device: next_dt on device arrays with its reduction partials
in a fixed per-GPU scratch array (no allocation), against the entry point it
varies, cdiff_next_dt (which reads host arrays through a solver handle), and
the oracle's next_dt (timestep_control.jl:33-62).  Two calls in a row, the
second on a smaller array, so the scratch the first call left behind must not
leak into the second; and the size limit of the scratch."""
import ctypes as C

import numpy as np
import pytest

import apt_oracle as O
from cases import gaussian_2d
from curvilineardiffusionmanuscript_amd import _lib as L
from curvilineardiffusionmanuscript_amd import solver as S

pytestmark = pytest.mark.gpu

MAX_ELEMENTS = 32768 * 4096


class _Hip:
    """Device buffers through the HIP runtime libcdiff itself links (the test
    allocates; the entry point under test does not)."""

    def __init__(self):
        L.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), x.nbytes) == 0
        assert self.h.hipMemcpy(p, x.ctypes.data, x.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n):
        y = np.empty(n, dtype=np.float64)
        assert self.h.hipMemcpy(y.ctypes.data, p, y.nbytes, 2) == 0  # device to host
        return y


def _device_next_dt(hip, a, b, dt, stream=None, **kw):
    da = hip.upload(a.ravel(order="F"))
    db = hip.upload(b.ravel(order="F"))
    out = hip.upload(np.full(1, -1.0))
    err = C.create_string_buffer(256)
    try:
        rc = L.load().cdiff_next_dt_device(
            da, db, a.size, dt, kw.get("tol", 1e-3), kw.get("maximum_u_change", 0.05),
            kw.get("timestep_growth_factor", 1.1), out, stream, err, len(err))
        assert rc == L.OK, err.value.decode()
        if stream:
            assert hip.h.hipStreamSynchronize(stream) == 0
        return float(hip.download(out, 1)[0])
    finally:
        for p in (da, db, out):
            hip.h.hipFree(p)


@pytest.mark.parametrize("kw", [{}, {"tol": 1e-2, "maximum_u_change": 0.1, "timestep_growth_factor": 1.5}])
def test_no_alloc_next_dt_matches_the_allocating_entry_point(kw):
    hip = _Hip()
    stream = C.c_void_p()
    assert hip.h.hipStreamCreate(C.byref(stream)) == 0
    rng = np.random.default_rng(4)
    for k, (n, ni) in enumerate(((300, 257), (40, 23))):  # the second call on a smaller array
        mesh, _ = gaussian_2d(n, ni=ni)
        dev = S.PseudoTransientSolver(mesh, [S.NeumannBC()] * 4)
        ora = O.OracleSolver(mesh, [("neumann", 0.0)] * 4)
        a = np.asfortranarray(1.0 + rng.random(mesh.dims_full))
        b = np.asfortranarray(a * (1.0 + 0.02 * rng.standard_normal(mesh.dims_full)))
        dt = 1e-3
        ref = S.next_dt(a, b, dt, solver=dev, **kw)
        got = _device_next_dt(hip, a, b, dt, stream if k == 0 else None, **kw)
        print(f"{mesh.dims_full}: cdiff_next_dt {ref!r}, cdiff_next_dt_device {got!r}")
        assert got == ref
        assert got == ora.next_dt(a, b, dt, **kw)
        dev.close()
    hip.h.hipStreamDestroy(stream)


def test_no_alloc_next_dt_rejects_arrays_beyond_its_scratch():
    hip = _Hip()
    t = hip.upload(np.ones(8))
    out = hip.upload(np.zeros(1))
    err = C.create_string_buffer(256)
    # the size is checked before anything is launched: no element is read
    rc = L.load().cdiff_next_dt_device(t, t, MAX_ELEMENTS + 1, 1e-3, 1e-3, 0.05, 1.1, out, None, err, len(err))
    assert rc == L.BAD_ARG
    assert "exceed the scratch" in err.value.decode()
    assert float(hip.download(out, 1)[0]) == 0.0
    hip.h.hipFree(t)
    hip.h.hipFree(out)
