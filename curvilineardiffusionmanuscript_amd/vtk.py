"""VTK output of the solver state, as the reference's ``to_vtk`` / ``save_vtk``
(src/pseudo_transient/PseudoTransientSolver.jl:370-412, src/vtk.jl:1-41).

The reference writes a structured grid (``.vts``) through WriteVTK.jl: the
domain node coordinates as points, one value per domain cell for each field,
and the time as field data ``TimeValue``.  This module writes the same file
layout (VTK XML StructuredGrid, appended raw binary, UInt64 block headers,
little-endian) with the same array names, so ParaView and the reference's
notebooks read it like the reference's output, and reads it back.

File name: ``<name>_<iteration as 7 digits>.vts`` (``get_filename``,
PseudoTransientSolver.jl:371-378).
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET

import numpy as np

_VTK_TYPES = {
    np.dtype(np.float32): "Float32",
    np.dtype(np.float64): "Float64",
    np.dtype(np.int32): "Int32",
    np.dtype(np.int64): "Int64",
    np.dtype(np.uint8): "UInt8",
}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def get_filename(iteration, name="diffusion"):
    """name with spaces replaced by underscores, a trailing underscore, and the
    iteration as 7 digits (PseudoTransientSolver.jl:371-378)."""
    name = name.replace(" ", "_")
    if not name.endswith("_"):
        name = name + "_"
    return f"{name}{int(iteration):07d}"


def write_vts(path, coords, cell_data, field_data=None, dtype=np.float32):
    """Write a VTK XML structured grid.

    coords: node coordinate arrays (2 or 3 of them, shape = cells + 1 per axis,
    first index fastest as in the reference's column-major arrays);
    cell_data: {name: array over the cells}; field_data: {name: scalar or 1-D}.
    Returns the path written (".vts" appended if missing).
    """
    if not path.endswith(".vts"):
        path = path + ".vts"
    coords = [np.asarray(c, dtype=np.float64) for c in coords]
    ndim = len(coords)
    if ndim not in (2, 3):
        raise ValueError("coords must hold 2 or 3 node arrays")
    nodes = coords[0].shape
    if any(c.shape != nodes for c in coords) or len(nodes) != ndim:
        raise ValueError("node arrays must share one shape of rank len(coords)")
    cells = tuple(n - 1 for n in nodes)
    ext = [0, cells[0], 0, cells[1], 0, cells[2] if ndim == 3 else 0]
    extent = " ".join(str(e) for e in ext)
    npts = int(np.prod(nodes))
    pts = np.zeros((npts, 3), dtype=dtype)
    for a in range(ndim):
        pts[:, a] = coords[a].reshape(-1, order="F")

    blocks = []  # raw bytes of the appended section, in order

    def add(arr):
        raw = np.ascontiguousarray(arr).tobytes(order="C")
        off = sum(len(b) for b in blocks)
        blocks.append(np.uint64(len(raw)).tobytes() + raw)
        return off

    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="StructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
             f'  <StructuredGrid WholeExtent="{extent}">']
    if field_data:
        lines.append("    <FieldData>")
        for name, v in field_data.items():
            v = np.atleast_1d(np.asarray(v, dtype=np.float64))
            off = add(v)
            lines.append(f'      <DataArray type="Float64" Name="{name}" NumberOfTuples="{v.size}" '
                         f'format="appended" offset="{off}"/>')
        lines.append("    </FieldData>")
    lines.append(f'    <Piece Extent="{extent}">')
    lines.append("      <PointData>")
    lines.append("      </PointData>")
    lines.append("      <CellData>")
    for name, arr in cell_data.items():
        arr = np.asarray(arr)
        if arr.shape != cells:
            raise ValueError(f"cell field {name!r} has shape {arr.shape}, the grid has {cells} cells")
        flat = arr.astype(dtype).reshape(-1, order="F")
        off = add(flat)
        lines.append(f'        <DataArray type="{_VTK_TYPES[np.dtype(dtype)]}" Name="{name}" format="appended" '
                     f'offset="{off}"/>')
    lines.append("      </CellData>")
    lines.append("      <Points>")
    off = add(pts)
    lines.append(f'        <DataArray type="{_VTK_TYPES[np.dtype(dtype)]}" Name="Points" NumberOfComponents="3" '
                 f'format="appended" offset="{off}"/>')
    lines.append("      </Points>")
    lines.append("    </Piece>")
    lines.append("  </StructuredGrid>")
    lines.append('  <AppendedData encoding="raw">')
    head = ("\n".join(lines) + "\n   _").encode("utf-8")
    tail = b"\n  </AppendedData>\n</VTKFile>\n"
    with open(path, "wb") as f:
        f.write(head)
        for b in blocks:
            f.write(b)
        f.write(tail)
    return path


def read_vts(path):
    """Read a file written by ``write_vts``: (node coordinate arrays,
    {cell field: array}, {field data: array})."""
    with open(path, "rb") as f:
        data = f.read()
    start = data.index(b'<AppendedData encoding="raw">')
    us = data.index(b"_", start) + 1
    end = data.rindex(b"</AppendedData>")
    xml_text = data[:start] + b"<AppendedData/></VTKFile>"
    root = ET.fromstring(xml_text)
    raw = data[us:end]
    sg = root.find("StructuredGrid")
    ext = [int(v) for v in sg.get("WholeExtent").split()]
    cells = (ext[1] - ext[0], ext[3] - ext[2], ext[5] - ext[4])
    ndim = 3 if cells[2] > 0 else 2
    cells = cells[:ndim]
    nodes = tuple(c + 1 for c in cells)

    def block(el):
        off = int(el.get("offset"))
        n = int(np.frombuffer(raw[off : off + 8], dtype=np.uint64)[0])
        return np.frombuffer(raw[off + 8 : off + 8 + n], dtype=_NP_TYPES[el.get("type")])

    field = {}
    fd = sg.find("FieldData")
    if fd is not None:
        for el in fd.findall("DataArray"):
            field[el.get("Name")] = block(el)
    piece = sg.find("Piece")
    cell = {}
    for el in piece.find("CellData").findall("DataArray"):
        cell[el.get("Name")] = block(el).reshape(cells, order="F")
    pts = block(piece.find("Points").find("DataArray")).reshape(-1, 3)
    coords = [pts[:, a].reshape(nodes, order="F") for a in range(ndim)]
    return coords, cell, field


def _domain(mesh, arr):
    h = mesh.nhalo
    return arr[tuple(slice(h, -h) for _ in range(arr.ndim))]


def _full(mesh, v):
    if np.isscalar(v) or np.ndim(v) == 0:
        return np.full(mesh.dims_full, float(v), order="F")
    return np.asarray(v)


def to_vtk(scheme, mesh, u, rho, iteration=0, t=0.0, name="diffusion", T=np.float32, directory="."):
    """Write the solver state as the reference's ``to_vtk`` does
    (PseudoTransientSolver.jl:380-412): TimeValue, rho, u, u_prev, residual,
    q1..qN (damped fluxes), q21..q2N (plain fluxes q'), diffusivity, dτ_ρ,
    θr_dτ, source_term, each over the domain cells.  Returns the path."""
    cd = {"rho": _domain(mesh, _full(mesh, rho)), "u": _domain(mesh, np.asarray(u)),
          "u_prev": _domain(mesh, scheme.u_prev), "residual": _domain(mesh, scheme.res)}
    for i, q in enumerate(scheme.q, start=1):
        cd[f"q{i}"] = _domain(mesh, q)
    for i, q in enumerate(scheme.q_prime, start=1):
        cd[f"q2{i}"] = _domain(mesh, q)
    cd["diffusivity"] = _domain(mesh, scheme.alpha)
    cd["dτ_ρ"] = _domain(mesh, scheme.dtau_rho)
    cd["θr_dτ"] = _domain(mesh, scheme.theta_r_dtau)
    cd["source_term"] = _domain(mesh, scheme.source_term)
    fn = os.path.join(directory, get_filename(iteration, name))
    return write_vts(fn, mesh.coords(), cd, {"TimeValue": float(t)}, dtype=T)


def save_vtk(scheme, u, rho, mesh, iteration=0, t=0.0, name="diffusion", T=np.float32, directory="."):
    """``save_vtk(scheme::PseudoTransientSolver, ...)`` (src/vtk.jl:35-39)."""
    return to_vtk(scheme, mesh, u, rho, iteration, t, name, T, directory)
