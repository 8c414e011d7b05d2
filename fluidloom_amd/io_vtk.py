"""VTK XML PolyData snapshots of the leaf cells (the reference's
io::VTKWriter, src/io/VTKWriter.cpp:20-130: one vertex per cell, a Level
array, file snapshot_<step:06d>.vtp in the output directory).

Points are cell centres in level-0 units ((x + 0.5) / 2^level ...).  Data
arrays are written as base64 "binary" DataArrays (UInt64 length header), so
multi-million-leaf snapshots stay compact; the reference writes ASCII and no
fields.
"""
from __future__ import annotations

import base64
import os
import struct

import numpy as np

_VTK_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32",
              np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}


def _data_array(name, a, ncomp=1):
    a = np.ascontiguousarray(a)
    raw = a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
    payload = base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")
    return (f'        <DataArray type="{_VTK_TYPES[a.dtype]}" Name="{name}" NumberOfComponents="{ncomp}" '
            f'format="binary">\n          {payload}\n        </DataArray>\n')


def write_vtp(path, x, y, z, level, point_data=None):
    """Writes one PolyData piece.  point_data: name -> array of shape (n,) or
    (n, 3)."""
    level = np.asarray(level, np.int64)
    scale = np.ldexp(1.0, -level)
    pts = np.stack([(np.asarray(x) + 0.5) * scale, (np.asarray(y) + 0.5) * scale,
                    (np.asarray(z) + 0.5) * scale], axis=1).astype(np.float32)
    n = len(pts)
    parts = ['<?xml version="1.0"?>\n',
             '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
             '  <PolyData>\n', f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n}">\n',
             '      <Points>\n', _data_array("Points", pts, 3), '      </Points>\n', '      <Verts>\n',
             _data_array("connectivity", np.arange(n, dtype=np.int64)),
             _data_array("offsets", np.arange(1, n + 1, dtype=np.int64)), '      </Verts>\n',
             '      <PointData>\n', _data_array("Level", level.astype(np.int32))]
    for name, a in (point_data or {}).items():
        a = np.asarray(a)
        parts.append(_data_array(name, a.astype(np.float32) if a.dtype.kind == "f" else a,
                                 1 if a.ndim == 1 else a.shape[1]))
    parts += ['      </PointData>\n', '    </Piece>\n', '  </PolyData>\n', '</VTKFile>\n']
    with open(path, "w") as f:
        f.writelines(parts)
    return path


def write_snapshot(output_dir, step, solver, inst=0):
    """Density and velocity of every fluid leaf of one instance."""
    os.makedirs(output_dir, exist_ok=True)
    x, y, z, lv = solver.fluid[inst]
    rho, u = solver.moments(inst)
    return write_vtp(os.path.join(output_dir, f"snapshot_{step:06d}.vtp"), x, y, z, lv,
                     {"rho": rho, "velocity": u})


def read_vtp(path):
    """Arrays of a file written by write_vtp (for tests and tools)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    out = {}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text.strip())
        (nbytes,) = struct.unpack("<Q", raw[:8])
        dt = {v: k for k, v in _VTK_TYPES.items()}[da.get("Type") or da.get("type")]
        a = np.frombuffer(raw[8:8 + nbytes], dtype=dt)
        nc = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = a.reshape(-1, nc) if nc > 1 else a
    return out
