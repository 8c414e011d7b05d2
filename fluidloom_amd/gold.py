"""Gold data: binary field snapshots for cross-run result comparison, in the
reference test harness's file format (tools/test_harness/GoldDataManager.h:
85-107, GoldDataManager.cpp:11-90), so files written here load in the
reference's reader and the other way round.

File `<test_name>.gold`, little-endian:

    offset  size  field
    0       4     magic 0x474F4C44 ("GOLD")
    4       4     version (3)
    8       8     num_cells
    16      4     num_fields
    20      4     (padding of the C struct)
    24      8     metadata_offset (byte offset of the JSON block)
    32      8     checksum
    40      32    reserved (zero)
    72      ...   num_fields float32 arrays of num_cells values each (SoA)
    metadata_offset  JSON metadata (version, test_name, num_cells, timestep,
                     simulation_time, field_names, timestamp) to end of file

The header is the 72-byte C struct the reference writes with
`sizeof(header)`.  Differences, all readable by the reference:
  * checksum holds the CRC-32 (zlib) of the field bytes; the reference leaves
    it 0 (its TODO).  `verify` checks it when it is non-zero.
  * the JSON also carries `field_norms` (L2 norm of every field), which the
    reference's GoldMetadata declares for tolerance scaling but never writes.
  * `load_field` selects the field by name from the JSON; the reference
    returns the first field whatever the name (its TODO).
"""
from __future__ import annotations

import json
import os
import struct
import time
import zlib

import numpy as np

MAGIC = 0x474F4C44
VERSION = 3
_HEADER = struct.Struct("<IIQI4xQQ32x")
HEADER_BYTES = _HEADER.size  # 72


class GoldError(RuntimeError):
    pass


def save(output_dir, test_name, field_names, fields, timestep, sim_time, timestamp=None):
    """Writes `<output_dir>/<test_name>.gold`; fields: equal-length 1-D
    arrays (stored as float32).  Returns the path."""
    if len(field_names) != len(fields):
        raise GoldError("Field names and data size mismatch")
    if not fields:
        raise GoldError("No field data to save")
    arrays = [np.ascontiguousarray(np.asarray(f, dtype="<f4")).ravel() for f in fields]
    n = arrays[0].size
    if any(a.size != n for a in arrays):
        raise GoldError("Inconsistent field sizes")
    body = b"".join(a.tobytes() for a in arrays)
    meta = {
        "version": VERSION,
        "test_name": test_name,
        "num_cells": n,
        "timestep": int(timestep),
        "simulation_time": float(sim_time),
        "field_names": list(field_names),
        "field_norms": {name: float(np.linalg.norm(a.astype(np.float64))) for name, a in zip(field_names, arrays)},
        "timestamp": int(time.time_ns() if timestamp is None else timestamp),
    }
    meta_bytes = (json.dumps(meta, indent=2) + "\n").encode("utf-8")
    offset = HEADER_BYTES + len(body)
    header = _HEADER.pack(MAGIC, VERSION, n, len(arrays), offset, zlib.crc32(body))
    path = os.path.join(output_dir, f"{test_name}.gold")
    with open(path, "wb") as fh:
        fh.write(header)
        fh.write(body)
        fh.write(meta_bytes)
    return path


def read_header(path):
    with open(path, "rb") as fh:
        raw = fh.read(HEADER_BYTES)
    if len(raw) != HEADER_BYTES:
        raise GoldError("Invalid gold data file: short header")
    magic, version, n, nf, offset, checksum = _HEADER.unpack(raw)
    return dict(magic=magic, version=version, num_cells=n, num_fields=nf, metadata_offset=offset, checksum=checksum)


def load(path):
    """Returns (metadata dict, {field name: float32 array})."""
    h = read_header(path)
    if h["magic"] != MAGIC:
        raise GoldError("Invalid gold data file: bad magic number")
    if h["version"] != VERSION:
        raise GoldError("Gold data version mismatch")
    n, nf = h["num_cells"], h["num_fields"]
    with open(path, "rb") as fh:
        data = fh.read()
    end = HEADER_BYTES + 4 * n * nf
    if len(data) < end or h["metadata_offset"] < end:
        raise GoldError("Invalid gold data file: truncated field data")
    meta = json.loads(data[h["metadata_offset"]:].decode("utf-8")) if h["metadata_offset"] else {}
    names = meta.get("field_names") or [f"field{k}" for k in range(nf)]
    if len(names) != nf:
        raise GoldError("Invalid gold data file: field names do not match the header")
    fields = {name: np.frombuffer(data, "<f4", n, HEADER_BYTES + 4 * n * k).copy() for k, name in enumerate(names)}
    return meta, fields


def load_field(path, field_name):
    meta, fields = load(path)
    if field_name not in fields:
        raise GoldError(f"field '{field_name}' not in {path}")
    return fields[field_name]


def verify(path):
    """(ok, message): magic, version, sizes and, when present, the CRC-32."""
    try:
        h = read_header(path)
        if h["magic"] != MAGIC:
            return False, "Invalid magic number"
        if h["version"] != VERSION:
            return False, "Version mismatch"
        with open(path, "rb") as fh:
            fh.seek(HEADER_BYTES)
            body = fh.read(4 * h["num_cells"] * h["num_fields"])
        if len(body) != 4 * h["num_cells"] * h["num_fields"]:
            return False, "Truncated field data"
        if h["checksum"] and zlib.crc32(body) != h["checksum"]:
            return False, "Checksum mismatch"
        return True, ""
    except (OSError, GoldError) as e:
        return False, str(e)


def compare(path_a, path_b, rtol=1e-5):
    """Per-field relative L2 difference of two gold files (same cells, same
    fields); returns {name: rel_l2} and raises GoldError when a field
    exceeds rtol (scaled by the reference file's stored L2 norm)."""
    meta_a, fa = load(path_a)
    meta_b, fb = load(path_b)
    if set(fa) != set(fb):
        raise GoldError("field sets differ")
    out = {}
    for name, a in fa.items():
        b = fb[name]
        if a.size != b.size:
            raise GoldError(f"field '{name}' sizes differ")
        norm = meta_a.get("field_norms", {}).get(name) or float(np.linalg.norm(a.astype(np.float64)))
        out[name] = float(np.linalg.norm(a.astype(np.float64) - b.astype(np.float64)) / max(norm, 1e-300))
    bad = {k: v for k, v in out.items() if v > rtol}
    if bad:
        raise GoldError(f"fields beyond rtol {rtol}: {bad}")
    return out


def analytical(output_dir, test_name, fn, nx, ny, nz, timestep, t):
    """Gold file of fn(x, y, z, t) on an nx*ny*nz grid of unit-cube points
    (i/nx, j/ny, k/nz), x fastest (the reference's generateAnalyticalGold,
    GoldDataManager.cpp:160-188); fn takes numpy arrays."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vals = fn(i.ravel() / nx, j.ravel() / ny, k.ravel() / nz, t)
    return save(output_dir, test_name, ["analytical_field"], [np.asarray(vals, np.float32)], timestep, t)


def solver_snapshot(output_dir, test_name, solver, inst=0, timestep=0, sim_time=0.0):
    """Density and velocity of every fluid leaf of one instance of a solver
    (fluidloom_amd.solver.LbmBatch), in the instance's leaf order."""
    rho, u = solver.moments(inst)
    return save(output_dir, test_name, ["rho", "ux", "uy", "uz"], [rho, u[:, 0], u[:, 1], u[:, 2]], timestep,
                sim_time)
