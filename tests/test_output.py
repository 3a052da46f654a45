"""VTK output (src/pseudo_transient/PseudoTransientSolver.jl:370-412,
src/vtk.jl) and the timing records of the reference's benchmark drivers
(benchmark_nonlinear_2d_w_source.jl:263-275, plot_nonlinear_benchmarks.jl:5-16)."""
import struct
import xml.etree.ElementTree as ET
from types import SimpleNamespace

import numpy as np
import pytest

from cases import gaussian_2d
from curvilineardiffusionmanuscript_amd import results as R
from curvilineardiffusionmanuscript_amd import vtk as V
from curvilineardiffusionmanuscript_amd.mesh import CurvilinearGrid3D, wavy_map_3d

REF_NAMES_2D = ["rho", "u", "u_prev", "residual", "q1", "q2", "q21", "q22", "diffusivity", "dτ_ρ", "θr_dτ",
                "source_term"]


def _scheme(shape, ndim, seed=0):
    rng = np.random.default_rng(seed)
    f = lambda: np.asfortranarray(rng.random(shape))  # noqa: E731
    return SimpleNamespace(u_prev=f(), res=f(), q=tuple(f() for _ in range(ndim)),
                           q_prime=tuple(f() for _ in range(ndim)), alpha=f(), dtau_rho=f(), theta_r_dtau=f(),
                           source_term=f())


def test_filename_like_the_reference():
    assert V.get_filename(5) == "diffusion_0000005"
    assert V.get_filename(12, "blob test") == "blob_test_0000012"
    assert V.get_filename(1234567, "x_") == "x_1234567"


def test_vts_2d_layout_and_round_trip(tmp_path):
    mesh, T = gaussian_2d(12, ni=9)
    sch = _scheme(mesh.dims_full, 2)
    rho = 2.0
    path = V.to_vtk(sch, mesh, T, rho, iteration=3, t=0.25, name="blob", directory=str(tmp_path))
    assert path.endswith("blob_0000003.vts")
    raw = open(path, "rb").read()
    start = raw.index(b'<AppendedData encoding="raw">')
    root = ET.fromstring(raw[:start] + b"<AppendedData/></VTKFile>")
    assert root.get("type") == "StructuredGrid" and root.get("header_type") == "UInt64"
    sg = root.find("StructuredGrid")
    nx, ny = mesh.dims_full[0] - 2, mesh.dims_full[1] - 2
    assert sg.get("WholeExtent") == f"0 {nx} 0 {ny} 0 0"
    cell = sg.find("Piece").find("CellData").findall("DataArray")
    assert [e.get("Name") for e in cell] == REF_NAMES_2D
    # decode "u" by hand from the appended block: UInt64 byte count, then
    # Float32 values of the domain cells, first index fastest
    el = cell[1]
    base = raw.index(b"_", start) + 1 + int(el.get("offset"))
    (nbytes,) = struct.unpack("<Q", raw[base : base + 8])
    assert nbytes == 4 * nx * ny
    u = np.frombuffer(raw[base + 8 : base + 8 + nbytes], dtype="<f4").reshape((nx, ny), order="F")
    np.testing.assert_array_equal(u, T[1:-1, 1:-1].astype(np.float32))
    coords, cd, fd = V.read_vts(path)
    x, y = mesh.coords()
    np.testing.assert_array_equal(coords[0], x.astype(np.float32))
    np.testing.assert_array_equal(coords[1], y.astype(np.float32))
    assert fd["TimeValue"][0] == 0.25
    np.testing.assert_array_equal(cd["rho"], np.full((nx, ny), 2.0, np.float32))
    np.testing.assert_array_equal(cd["q21"], sch.q_prime[0][1:-1, 1:-1].astype(np.float32))
    np.testing.assert_array_equal(cd["θr_dτ"], sch.theta_r_dtau[1:-1, 1:-1].astype(np.float32))


def test_vts_3d_float64(tmp_path):
    mesh = CurvilinearGrid3D.from_mapping(wavy_map_3d(0.04), (5, 4, 3), 1)
    sch = _scheme(mesh.dims_full, 3, seed=1)
    u = np.asfortranarray(np.random.default_rng(2).random(mesh.dims_full))
    path = V.save_vtk(sch, u, 1.0, mesh, 7, 1.5, "three d", T=np.float64, directory=str(tmp_path))
    coords, cd, fd = V.read_vts(path)
    assert path.endswith("three_d_0000007.vts")
    assert [k for k in cd] == REF_NAMES_2D[:6] + ["q3", "q21", "q22", "q23"] + REF_NAMES_2D[8:]
    np.testing.assert_array_equal(cd["u"], u[1:-1, 1:-1, 1:-1])
    np.testing.assert_array_equal(cd["q3"], sch.q[2][1:-1, 1:-1, 1:-1])
    z = mesh.coords()[2]
    np.testing.assert_array_equal(coords[2], z)


def test_vts_rejects_wrong_shapes(tmp_path):
    x, y = np.meshgrid(np.arange(4.0), np.arange(3.0), indexing="ij")
    with pytest.raises(ValueError):
        V.write_vts(str(tmp_path / "a"), [x, y], {"u": np.zeros((4, 3))})


def test_timing_csv_and_timer_json(tmp_path):
    log = R.StepLog()
    for k in range(4):
        log.record(k * 5e-6, SimpleNamespace(time=0.01 * (k + 1), niter=10 + k, rel_err=1e-9 * k))
        if k == 0:
            log.reset_timer()  # the reference's loop resets the timer after cycle 0
    log.write_csv(str(tmp_path / "t.csv"))
    cols = R.read_timing_csv(str(tmp_path / "t.csv"))
    assert open(tmp_path / "t.csv").readline().strip() == "time,time_per_cycle,iter_per_cycle,error"
    assert cols["iter_per_cycle"] == [10.0, 11.0, 12.0, 13.0] and cols["time"][1] == 5e-6
    log.write_timer_json(str(tmp_path / "t.json"))
    tot_s, per_ms, n = R.runtimes(str(tmp_path / "t.json"))
    assert n == 3 and abs(tot_s - 0.09) < 1e-12 and abs(per_ms - 30.0) < 1e-9
