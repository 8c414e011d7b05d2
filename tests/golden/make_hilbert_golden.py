"""Writes tests/golden/hilbert_golden.npz from the reference's own Hilbert
codec sources (src/core/hilbert/HilbertCodec.cpp and
src/adaptation/utils/HilbertCodec3D.cpp), compiled in place by
`make -C oracle ref` into oracle/_ref/libref_hilbert.so.

Run where the reference tree is present (the fixture it writes is what the
tests read elsewhere):
    python tests/golden/make_hilbert_golden.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

O.build(ref=True)
R = O.ref()
assert R is not None, "reference codec not built"

rng = np.random.default_rng(20251203)
rows = []
for level in range(0, 22):
    m = (1 << level) - 1
    n = 200 if level else 4
    xs = rng.integers(0, m + 1, n) if level else np.zeros(n, np.int64)
    ys = rng.integers(0, m + 1, n) if level else np.zeros(n, np.int64)
    zs = rng.integers(0, m + 1, n) if level else np.zeros(n, np.int64)
    for x, y, z in zip(xs, ys, zs):
        rows.append((int(x), int(y), int(z), level))
# coordinates beyond 21 bits and negative coordinates exercise the masking
for (x, y, z, level) in [(-1, 5, 7, 8), (1 << 22, 3, 1, 6), (2**31 - 1, -7, 12345, 21), (123, 456, 789, 5)]:
    rows.append((x, y, z, level))
a = np.array(rows, np.int64)
x, y, z, lv = (a[:, i] for i in range(4))
enc = np.array([R.ref_hilbert_encode(int(p), int(q), int(r), int(l)) for p, q, r, l in rows], np.uint64)
dx = np.empty(len(rows), np.int32); dy = np.empty_like(dx); dz = np.empty_like(dx)
for i, (h, l) in enumerate(zip(enc, lv)):
    ox, oy, oz = C.c_int32(), C.c_int32(), C.c_int32()
    R.ref_hilbert_decode(int(h), int(l), C.byref(ox), C.byref(oy), C.byref(oz))
    dx[i], dy[i], dz[i] = ox.value, oy.value, oz.value
valid = np.array([R.ref_is_valid(int(h), int(l)) for h, l in zip(enc, lv)], np.uint8)
rot8 = np.array([R.ref_rotdir_encode(int(p), int(q), int(r), 8) for p, q, r, _ in rows], np.uint64)
rot10 = np.array([R.ref_rotdir_encode(int(p), int(q), int(r), 10) for p, q, r, _ in rows], np.uint64)
np.savez_compressed(Path(__file__).with_name("hilbert_golden.npz"), x=x.astype(np.int32), y=y.astype(np.int32),
                    z=z.astype(np.int32), level=lv.astype(np.uint8), encode=enc, dec_x=dx, dec_y=dy, dec_z=dz,
                    valid=valid, rotdir8=rot8, rotdir10=rot10)
print("wrote", len(rows), "rows")
