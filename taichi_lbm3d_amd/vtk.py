"""VTK XML rectilinear-grid (.vtr) writer and reader.

``grid_to_vtk`` takes the arguments the reference passes to pyevtk's
``gridToVTK`` in export_VTK (SP:462-475): a path without extension, the x, y,
z coordinate vectors and ``pointData`` whose values are 3-D arrays (scalars)
or 3-tuples of 3-D arrays (vectors).  It writes a VTK XML RectilinearGrid with
raw appended little-endian data, UInt64 block headers, points ordered x
fastest as VTK requires.  ``read_vtr`` parses such a file back.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

_VTK_TYPES = {
    np.dtype(np.int8): "Int8", np.dtype(np.uint8): "UInt8", np.dtype(np.int16): "Int16",
    np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64", np.dtype(np.float32): "Float32",
    np.dtype(np.float64): "Float64",
}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def _fortran_points(a: np.ndarray) -> np.ndarray:
    """(nx,ny,nz[,c]) -> flat array with x fastest over points, components innermost."""
    if a.ndim == 3:
        return np.ascontiguousarray(a.transpose(2, 1, 0)).reshape(-1)
    return np.ascontiguousarray(a.transpose(2, 1, 0, 3)).reshape(-1)


def grid_to_vtk(path: str, x, y, z, pointData: dict) -> str:
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    nx, ny, nz = len(x), len(y), len(z)
    extent = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    blocks = []
    coords_xml = []
    offset = 0
    for name, arr in (("x_coordinates", x), ("y_coordinates", y), ("z_coordinates", z)):
        coords_xml.append(f'<DataArray type="Float64" Name="{name}" NumberOfComponents="1" '
                          f'format="appended" offset="{offset}"/>')
        blocks.append(arr.astype("<f8").tobytes())
        offset += 8 + len(blocks[-1])
    point_xml = []
    for name, val in pointData.items():
        if isinstance(val, (tuple, list)):
            comps = [np.asarray(c) for c in val]
            for c in comps:
                if c.shape != (nx, ny, nz):
                    raise ValueError(f"point data {name!r}: shape {c.shape} != {(nx, ny, nz)}")
            arr = np.stack(comps, axis=-1)
            ncomp = len(comps)
        else:
            arr = np.asarray(val)
            if arr.shape != (nx, ny, nz):
                raise ValueError(f"point data {name!r}: shape {arr.shape} != {(nx, ny, nz)}")
            ncomp = 1
        dt = arr.dtype
        if dt not in _VTK_TYPES:
            arr = arr.astype(np.float64)
            dt = arr.dtype
        data = _fortran_points(arr).astype(dt.newbyteorder("<")).tobytes()
        point_xml.append(f'<DataArray type="{_VTK_TYPES[dt]}" Name="{name}" NumberOfComponents="{ncomp}" '
                         f'format="appended" offset="{offset}"/>')
        blocks.append(data)
        offset += 8 + len(data)
    header = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        f'<RectilinearGrid WholeExtent="{extent}">\n'
        f'<Piece Extent="{extent}">\n'
        "<Coordinates>\n" + "\n".join(coords_xml) + "\n</Coordinates>\n"
        "<PointData>\n" + "\n".join(point_xml) + "\n</PointData>\n"
        "</Piece>\n</RectilinearGrid>\n"
        '<AppendedData encoding="raw">\n_'
    )
    fname = path + ".vtr"
    with open(fname, "wb") as fh:
        fh.write(header.encode("ascii"))
        for b in blocks:
            fh.write(np.uint64(len(b)).astype("<u8").tobytes())
            fh.write(b)
        fh.write(b"\n</AppendedData>\n</VTKFile>\n")
    return fname


def read_vtr(fname: str) -> dict:
    """Parse a file written by :func:`grid_to_vtk`: returns {"x","y","z", name: array}."""
    raw = open(fname, "rb").read()
    start = raw.index(b'<AppendedData encoding="raw">')
    us = raw.index(b"_", start) + 1
    root = ET.fromstring(raw[:start].decode("ascii") + "</VTKFile>")
    grid = root.find("RectilinearGrid")
    ext = [int(v) for v in grid.get("WholeExtent").split()]
    shape = (ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1)
    out = {}
    for da in root.iter("DataArray"):
        off = us + int(da.get("offset"))
        n = int(np.frombuffer(raw[off:off + 8], "<u8")[0])
        dt = np.dtype(_NP_TYPES[da.get("type")]).newbyteorder("<")
        arr = np.frombuffer(raw[off + 8:off + 8 + n], dt)
        name = da.get("Name")
        if name.endswith("_coordinates"):
            out[name[0]] = arr.copy()
            continue
        ncomp = int(da.get("NumberOfComponents", "1"))
        if ncomp == 1:
            out[name] = arr.reshape(shape[::-1]).transpose(2, 1, 0).copy()
        else:
            out[name] = arr.reshape(shape[::-1] + (ncomp,)).transpose(2, 1, 0, 3).copy()
    return out
