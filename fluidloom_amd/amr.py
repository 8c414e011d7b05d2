"""Device-resident AMR cell arrays and the integer hot path (keys, sort,
scan, hash, neighbours, regrid) over the C ABI of libfluidloom_hip.so.

This is the Python mirror of the reference's adaptation / hashmap host layer
(src/adaptation/AdaptationEngine.cpp, src/core/hashmap/*.cpp): every call
runs the HIP kernels; nothing here computes on the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (Cells, DeviceArray, RegridConfig, RegridResult, check, lib)

_FIELDS = (("x", np.int32), ("y", np.int32), ("z", np.int32), ("level", np.uint8),
           ("state", np.uint8), ("flag", np.int32), ("material", np.uint32))


class GpuCells:
    """SoA cell arrays on the device (native per-level coordinates)."""

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self.n = 0
        for name, dt in _FIELDS:
            setattr(self, name, DeviceArray(max(self.capacity, 1), dt))

    @classmethod
    def from_host(cls, capacity=None, **arrays):
        n = len(arrays["x"])
        g = cls(capacity or n)
        g.n = n
        for name, dt in _FIELDS:
            a = arrays.get(name)
            if a is None:
                a = np.zeros(n, dt)
            a = np.ascontiguousarray(a, dt)
            if n:
                check(lib().fl_memcpy_h2d(getattr(g, name).ptr, a.ctypes.data, a.nbytes, None))
        return g

    def struct(self) -> Cells:
        return Cells(*(getattr(self, nm).p for nm, _ in _FIELDS))

    def to_host(self) -> dict:
        out = {}
        for name, dt in _FIELDS:
            a = np.empty(self.n, dt)
            if self.n:
                check(lib().fl_memcpy_d2h(a.ctypes.data, getattr(self, name).ptr, a.nbytes, None))
            out[name] = a
        return out

    def free(self):
        for name, _ in _FIELDS:
            getattr(self, name).free()


def hilbert_encode(x: DeviceArray, y, z, level, n, level_scalar=0, stream=None) -> DeviceArray:
    out = DeviceArray(max(n, 1), np.uint64)
    check(lib().fl_hilbert_encode(x.p, y.p, z.p, level.p if level is not None else None,
                                  level_scalar, out.p, n, stream))
    return out


def sort_pairs(keys: DeviceArray, vals: DeviceArray, n, begin_bit=0, end_bit=64, stream=None):
    """Stable in-place sort of (keys, vals) on the device."""
    ka = DeviceArray(max(n, 1), np.uint64)
    va = DeviceArray(max(n, 1), np.uint32)
    tb = lib().fl_sort_temp_bytes(n)
    tmp = DeviceArray(max(tb, 1), np.uint8)
    check(lib().fl_sort_pairs_u64u32(keys.p, vals.p, ka.p, va.p, n, begin_bit, end_bit, tmp.p, tb,
                                     stream))


def scan_exclusive(a: DeviceArray, n, stream=None):
    out = DeviceArray(max(n, 1), np.uint32)
    tot = DeviceArray(1, np.uint64)
    tb = lib().fl_scan_temp_bytes(n)
    tmp = DeviceArray(max(tb, 1), np.uint8)
    check(lib().fl_scan_exclusive_u32(a.p, out.p, tot.p, n, tmp.p, tb, stream))
    return out, int(tot.to_host()[0])


@dataclass
class RegridStats:
    num_cells_out: int
    num_split: int
    num_groups: int
    balance_iterations: int
    balance_violations: int
    balance_marked: int
    balance_converged: bool


def regrid_temp(n) -> DeviceArray:
    """A workspace for regrid() of up to n cells (reusable across calls)."""
    return DeviceArray(max(lib().fl_regrid_temp_bytes(n), 1), np.uint8)


def regrid(cells: GpuCells, balance_mode=1, anchor_level=10, max_split_level=8, max_iters=64,
           root=(1, 1, 1), merge_balance_check=0, fields=None, ncomp=0, avg_rule=0,
           out_capacity=None, stream=None, out=None, temp=None):
    """One device regrid pass (balance, merge, compaction, split, parents).

    `out` (GpuCells) and `temp` (from regrid_temp) may be passed in to reuse
    their memory; otherwise they are allocated here.  Returns (new GpuCells,
    RegridStats[, new fields]).  Flags are consumed."""
    n = cells.n
    if out is None:
        cap = out_capacity or (n + 7 * n if n else 1)
        out = GpuCells(cap)
    else:
        cap = out.capacity
    cfg = RegridConfig()
    cfg.balance_mode, cfg.anchor_level, cfg.max_split_level = balance_mode, anchor_level, max_split_level
    cfg.max_balance_iters = max_iters
    for a in range(3):
        cfg.root[a] = int(root[a])
    cfg.merge_balance_check, cfg.ncomp, cfg.avg_rule = merge_balance_check, ncomp, avg_rule
    fout = DeviceArray(max(cap * max(ncomp, 1), 1), np.float32) if ncomp else None
    tb = lib().fl_regrid_temp_bytes(n)
    tmp = temp if temp is not None and temp.nbytes >= tb else DeviceArray(max(tb, 1), np.uint8)
    res = RegridResult()
    cs, os_ = cells.struct(), out.struct()
    check(lib().fl_regrid(C.byref(cfg), C.byref(cs), n, C.byref(os_), cap,
                          fields.p if fields is not None else None,
                          fout.p if fout is not None else None, tmp.p, tb, C.byref(res), stream))
    out.n = int(res.num_cells_out)
    st = RegridStats(int(res.num_cells_out), int(res.num_split), int(res.num_groups),
                     int(res.balance_iterations), int(res.balance_violations),
                     int(res.balance_marked), bool(res.balance_converged))
    if ncomp:
        return out, st, fout
    return out, st


def balance(cells: GpuCells, balance_mode=1, anchor_level=10, max_split_level=8, max_iters=64, root=(1, 1, 1),
            temp=None, stream=None) -> RegridStats:
    """2:1 balance to a fixpoint on the device (fl_balance): marks split
    flags in place and returns the iteration counters (the cell count is
    unchanged)."""
    n = cells.n
    cfg = RegridConfig()
    cfg.balance_mode, cfg.anchor_level, cfg.max_split_level = balance_mode, anchor_level, max_split_level
    cfg.max_balance_iters = max_iters
    for a in range(3):
        cfg.root[a] = int(root[a])
    tb = lib().fl_regrid_temp_bytes(n)
    tmp = temp if temp is not None and temp.nbytes >= tb else DeviceArray(max(tb, 1), np.uint8)
    res = RegridResult()
    cs = cells.struct()
    check(lib().fl_balance(C.byref(cfg), C.byref(cs), n, tmp.p, tb, C.byref(res), stream))
    return RegridStats(n, 0, 0, int(res.balance_iterations), int(res.balance_violations),
                       int(res.balance_marked), bool(res.balance_converged))


def mark_ball(cells: GpuCells, center, radius, target_level, stream=None) -> int:
    m = C.c_uint64(0)
    cs = cells.struct()
    check(lib().fl_mark_ball(C.byref(cs), cells.n, float(center[0]), float(center[1]),
                             float(center[2]), float(radius), int(target_level), C.byref(m), stream))
    return m.value


def mark_solid_ball(cells: GpuCells, center, radius, stream=None) -> int:
    m = C.c_uint64(0)
    cs = cells.struct()
    check(lib().fl_mark_solid_ball(C.byref(cs), cells.n, float(center[0]), float(center[1]),
                                   float(center[2]), float(radius), C.byref(m), stream))
    return m.value


def face_neighbors(cells: GpuCells, root=(1, 1, 1), stream=None) -> DeviceArray:
    n = cells.n
    cap = int(lib().fl_hash_capacity(n))
    keys = DeviceArray(max(n, 1), np.uint64)
    tk = DeviceArray(cap, np.uint64)
    tv = DeviceArray(cap, np.uint32)
    nbr = DeviceArray(max(6 * n, 1), np.uint32)
    r = (C.c_int32 * 3)(*root)
    cs = cells.struct()
    check(lib().fl_face_neighbors(C.byref(cs), n, r, keys.p, tk.p, tv.p, cap, nbr.p, stream))
    return nbr
