"""VTU output of nodal fields (the reference's VTKWriter, src/vtk_class.cpp:20-429, as
called for map_type = vtu at src/poisson_boltzmann.cpp:318-351).

File layout as the reference writes it (one rank):

* ``<basename>_0000.vtu``: an UnstructuredGrid with appended raw binary data. The
  DataArrays are ``Array`` (point coordinates, Float64 x 3), ``connectivity``,
  ``offsets`` and ``types`` (Int32; every cell a VTK_VOXEL, type 11, 8 points in the
  leaf's corner order), then the field (Float64, point data). Each appended block is
  preceded by its byte count as an int32.
* Points (VTKWriter::preparingData, :48-101): the mesh nodes in node order with their
  field values, then one extra point per hanging corner occurrence, in leaf order and
  corner order, whose value is the sum of its parents' values divided by their
  number.
* ``total_<basename>.pvtu`` (createPvtuFile, :48-100): the parallel index naming the
  pieces.

The field values come from the device (``Context.get_phi`` / ``get_markers``); this
module only lays them out.
"""
from __future__ import annotations

import struct
from typing import Sequence

import numpy as np


def vtu_arrays(mesh, field: np.ndarray):
    """(points (P, 3), values (P,), connectivity (8 L,)) as the reference lays them out."""
    N = mesh.n_nodes
    lc = np.asarray(mesh.leaf_corner, dtype=np.int64)
    f = np.asarray(field, dtype=np.float64)
    hang = lc < 0
    occ = -1 - lc[hang]                      # hanging node of each occurrence, leaf-major order
    npar = np.asarray(mesh.hang_np, dtype=np.int64)[occ]
    par = np.asarray(mesh.hang_par, dtype=np.int64)[occ]
    s = np.zeros(occ.shape[0])
    for j in range(4):                       # sequential sum over the parents, as the reference
        s = s + np.where(j < npar, f[np.maximum(par[:, j], 0)], 0.0)
    hv = s / npar
    points = np.concatenate([np.asarray(mesh.node_xyz, dtype=np.float64),
                             np.asarray(mesh.hang_xyz, dtype=np.float64)[occ]])
    values = np.concatenate([f[:N], hv])
    conn = lc.copy()
    conn[hang] = N + np.arange(occ.shape[0])
    return points, values, conn.reshape(-1).astype(np.int32)


def _block(a: np.ndarray) -> bytes:
    raw = np.ascontiguousarray(a).tobytes()
    return struct.pack("<i", len(raw)) + raw


def write_vtu(path: str, mesh, field: np.ndarray, fieldname: str) -> None:
    points, values, conn = vtu_arrays(mesh, field)
    nelems = mesh.n_leaves
    blocks = [_block(points.reshape(-1)), _block(conn), _block(8 * np.arange(1, nelems + 1, dtype=np.int32)),
              _block(np.full(nelems, 11, dtype=np.int32)), _block(values)]
    offs = np.concatenate([[0], np.cumsum([len(b) for b in blocks])[:-1]])
    head = ['<?xml version="1.0"?>\n',
            '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
            '<UnstructuredGrid>\n',
            f'<Piece NumberOfPoints="{points.shape[0]}" NumberOfCells="{nelems}">\n',
            '<Points>\n',
            f'<DataArray type="Float64" Name="Array" NumberOfComponents="3" format="appended" offset="{offs[0]}">\n',
            '</DataArray>\n', '</Points>\n', '<Cells>\n',
            f'<DataArray type="Int32" Name="connectivity" format="appended" offset="{offs[1]}">\n', '</DataArray>\n',
            f'<DataArray type="Int32" Name="offsets" format="appended" offset="{offs[2]}">\n', '</DataArray>\n',
            f'<DataArray type="Int32" Name="types" format="appended" offset="{offs[3]}">\n', '</DataArray>\n',
            '</Cells>\n', '<PointData>\n',
            f'<DataArray type="Float64" Name="{fieldname}" NumberOfComponents="1" format="appended" '
            f'offset="{offs[4]}">\n', '</DataArray>\n', '</PointData>\n',
            '</Piece>\n', '</UnstructuredGrid>\n', ' <AppendedData encoding="raw">\n', '_']
    with open(path, "wb") as fh:
        fh.write("".join(head).encode())
        for b in blocks:
            fh.write(b)
        fh.write(b" </AppendedData>\n</VTKFile>\n")


def write_pvtu(fieldnames: Sequence[str], basenames: Sequence[str], n_ranks: int = 1, directory: str = ".") -> None:
    for fname, base in zip(fieldnames, basenames):
        lines = ['<?xml version="1.0"?>\n',
                 '<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
                 '  <PUnstructuredGrid>\n', '    <PPoints>\n',
                 '      <PDataArray type="Float64" NumberOfComponents="3" Name="Array"/>\n', '    </PPoints>\n',
                 '    <PCells>\n', '      <PDataArray type="Int32" Name="connectivity"/>\n',
                 '      <PDataArray type="Int32" Name="offsets"/>\n', '      <PDataArray type="Int32" Name="types"/>\n',
                 '    </PCells>\n', '    <PPointData>\n', f'      <PDataArray type="Float64" Name="{fname}"/>\n',
                 '    </PPointData>\n']
        lines += [f'    <Piece Source="{base}_{r:04d}.vtu"/>\n' for r in range(n_ranks)]
        lines += ['  </PUnstructuredGrid>\n', '</VTKFile>\n']
        with open(f"{directory}/total_{base}.pvtu", "w") as fh:
            fh.write("".join(lines))


def read_vtu(path: str) -> dict:
    """Reads back a file written by write_vtu (for tests and checks)."""
    data = open(path, "rb").read()
    head, raw = data.split(b'<AppendedData encoding="raw">\n_', 1)
    text = head.decode()
    import re
    arrays = re.findall(r'<DataArray type="(\w+)" Name="(\w+)"[^>]*offset="(\d+)"', text)
    out = {}
    for typ, name, off in arrays:
        off = int(off)
        n = struct.unpack("<i", raw[off:off + 4])[0]
        dt = np.float64 if typ == "Float64" else np.int32
        out[name] = np.frombuffer(raw[off + 4:off + 4 + n], dtype=dt)
    m = re.search(r'NumberOfPoints="(\d+)" NumberOfCells="(\d+)"', text)
    out["n_points"], out["n_cells"] = int(m.group(1)), int(m.group(2))
    return out
