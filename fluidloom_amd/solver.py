"""AMR D3Q19 lattice-Boltzmann solver over the fl_lbm_* C ABI.

Mirrors the reference's run loop (KernelNode::execute over SOAFieldManager
fields, src/runtime/KernelNode.cpp:37-106) with the build's own D3Q19 step.
All arithmetic runs in libfluidloom_hip.so.

LbmBatch holds many independent instances in one device layout (one launch
per level steps all of them); LbmSolver is the single-instance form.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (FL_BC_OUTFLOW, FL_BC_PERIODIC, FL_BC_WALL, FL_STORE_F16, FL_STORE_F16C, FL_STORE_F32,
                   FL_STORE_F64, LbmParams, LbmStats, check, lib)

STORAGE = {"f32": FL_STORE_F32, "f16": FL_STORE_F16, "f64": FL_STORE_F64, "f16c": FL_STORE_F16C}
BC = {"wall": FL_BC_WALL, "outflow": FL_BC_OUTFLOW, "periodic": FL_BC_PERIODIC}


def make_params(nx0, ny0, nz0, max_level, tau0, bc, bc_u, storage="f32", rho_init=1.0,
                u_init=(0.0, 0.0, 0.0), rho0=1.0) -> LbmParams:
    p = LbmParams()
    p.nx0, p.ny0, p.nz0, p.max_level = int(nx0), int(ny0), int(nz0), int(max_level)
    p.tau0, p.rho0, p.rho_init = float(tau0), float(rho0), float(rho_init)
    p.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
    for f in range(6):
        p.bc[f] = BC[bc[f]] if isinstance(bc[f], str) else int(bc[f])
        for a in range(3):
            p.bc_u[f][a] = float(bc_u[f][a])
    for a in range(3):
        p.u_init[a] = float(u_init[a])
    return p


def _concat(meshes):
    off = np.zeros(len(meshes) + 1, np.uint64)
    for k, m in enumerate(meshes):
        off[k + 1] = off[k] + len(m["x"])
    cat = {}
    for key, dt in (("x", np.int32), ("y", np.int32), ("z", np.int32), ("level", np.uint8), ("state", np.uint8)):
        parts = [np.asarray(m[key], dt) for m in meshes]
        cat[key] = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0, dt))
    return off, cat


def _fluid(m):
    fl = np.asarray(m["state"]) == 0
    return tuple(np.ascontiguousarray(np.asarray(m[k], dt)[fl]) for k, dt in
                 (("x", np.int32), ("y", np.int32), ("z", np.int32), ("level", np.uint8)))


def plan_bytes(params, meshes) -> int:
    """Device bytes a batch of these instances would take (host layout only)."""
    off, cat = _concat(meshes)
    arr = (LbmParams * len(params))(*params)
    out = C.c_uint64(0)
    check(lib().fl_lbm_plan_bytes(arr, len(params), off.ctypes.data, cat["x"].ctypes.data,
                                  cat["y"].ctypes.data, cat["z"].ctypes.data, cat["level"].ctypes.data,
                                  cat["state"].ctypes.data, C.byref(out)))
    return int(out.value)


class LbmBatch:
    def __init__(self, params, meshes, stream=None, graph=False):
        """params: list of LbmParams; meshes: list of dicts with x, y, z, level, state."""
        assert len(params) == len(meshes) and len(params) >= 1
        self.params = list(params)
        self._meshes = [dict(m) for m in meshes]
        self._dev = None
        off, cat = _concat(meshes)
        self._arr = (LbmParams * len(params))(*params)
        self.h = C.c_void_p()
        check(lib().fl_lbm_create_batch(self._arr, len(params), off.ctypes.data, cat["x"].ctypes.data,
                                        cat["y"].ctypes.data, cat["z"].ctypes.data, cat["level"].ctypes.data,
                                        cat["state"].ctypes.data, stream, C.byref(self.h)))
        self._fluid = [_fluid(m) for m in meshes]
        if graph:
            self.set_graph(True)

    @property
    def ninst(self):
        return len(self.params)

    def _download(self):
        """Host copies of the leaves after a regrid from device leaves."""
        cells, off = self._dev
        h = cells.to_host()
        self._meshes = [{k: h[k][off[i]:off[i + 1]].copy() for k in ("x", "y", "z", "level", "state")}
                        for i in range(self.ninst)]
        self._fluid = [_fluid(m) for m in self._meshes]
        self._dev = None

    @property
    def meshes(self):
        if self._meshes is None:
            self._download()
        return self._meshes

    @property
    def fluid(self):
        if self._fluid is None:
            self._download()
        return self._fluid

    def step(self, n=1):
        check(lib().fl_lbm_step(self.h, int(n)))

    def stream(self):
        s = C.c_void_p()
        check(lib().fl_lbm_stream(self.h, C.byref(s)))
        return s

    def sync(self):
        check(lib().fl_stream_sync(self.stream()))

    def stats(self) -> LbmStats:
        st = LbmStats()
        check(lib().fl_lbm_stats_get(self.h, C.byref(st)))
        return st

    def inst_updates(self, inst):
        lv = (C.c_uint64 * 24)()
        up = C.c_uint64(0)
        check(lib().fl_lbm_inst_leaves(self.h, int(inst), lv, C.byref(up)))
        return int(up.value), [int(v) for v in lv]

    def _q(self, inst, x, y, z, lv):
        if x is None:
            return self.fluid[inst]
        return tuple(np.ascontiguousarray(a, dt) for a, dt in
                     ((x, np.int32), (y, np.int32), (z, np.int32), (lv, np.uint8)))

    def get_f(self, inst=0, x=None, y=None, z=None, lv=None):
        x, y, z, lv = self._q(inst, x, y, z, lv)
        out = np.empty((len(x), 19), np.float64)
        check(lib().fl_lbm_get_f_inst(self.h, int(inst), x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                      lv.ctypes.data, len(x), out.ctypes.data))
        return out

    def set_f(self, f, inst=0, x=None, y=None, z=None, lv=None):
        x, y, z, lv = self._q(inst, x, y, z, lv)
        f = np.ascontiguousarray(f, np.float64)
        check(lib().fl_lbm_set_f_inst(self.h, int(inst), x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                      lv.ctypes.data, len(x), f.ctypes.data))

    def moments(self, inst=0, x=None, y=None, z=None, lv=None):
        x, y, z, lv = self._q(inst, x, y, z, lv)
        rho = np.empty(len(x))
        u = np.empty((len(x), 3))
        check(lib().fl_lbm_moments_inst(self.h, int(inst), x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                        lv.ctypes.data, len(x), rho.ctypes.data, u.ctypes.data))
        return rho, u

    def vorticity(self, inst=0):
        x, y, z, lv = self.fluid[inst]
        out = np.empty(len(x))
        check(lib().fl_lbm_vorticity_inst(self.h, int(inst), x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                          lv.ctypes.data, len(x), out.ctypes.data))
        return out

    def totals_batch(self):
        m = np.zeros(self.ninst)
        j = np.zeros((self.ninst, 3))
        check(lib().fl_lbm_totals_batch(self.h, m.ctypes.data, j.ctypes.data))
        return m, j

    def force_batch(self):
        F = np.zeros((self.ninst, 3))
        check(lib().fl_lbm_force_batch(self.h, F.ctypes.data))
        return F

    def velocity_change(self):
        du = np.zeros(self.ninst)
        check(lib().fl_lbm_velocity_change(self.h, du.ctypes.data))
        return du

    def velocity_sample(self):
        check(lib().fl_lbm_velocity_sample(self.h))

    def set_live(self, inst, live):
        check(lib().fl_lbm_set_live(self.h, int(inst), int(bool(live))))

    def set_graph(self, enable):
        check(lib().fl_lbm_set_graph(self.h, int(bool(enable))))

    def set_variant(self, level, variant):
        check(lib().fl_lbm_set_variant(self.h, int(level), int(variant)))

    def autotune(self, iters=10):
        n = self.stats().nlevels
        out = (C.c_int * max(n, 1))()
        check(lib().fl_lbm_autotune(self.h, int(iters), out))
        return [int(out[l]) for l in range(n)]

    def time_collide(self, level, iters, variant=None):
        ms = C.c_float(0)
        if variant is None:
            check(lib().fl_lbm_time_collide(self.h, int(level), int(iters), C.byref(ms)))
        else:
            check(lib().fl_lbm_time_collide_variant(self.h, int(level), int(variant), int(iters), C.byref(ms)))
        return ms.value

    def regrid_batch(self, meshes):
        """Switch every instance to a new leaf list, carrying the populations over."""
        assert len(meshes) == self.ninst
        off, cat = _concat(meshes)
        check(lib().fl_lbm_regrid_batch(self.h, off.ctypes.data, cat["x"].ctypes.data, cat["y"].ctypes.data,
                                        cat["z"].ctypes.data, cat["level"].ctypes.data, cat["state"].ctypes.data))
        self._meshes = [dict(m) for m in meshes]
        self._fluid = [_fluid(m) for m in meshes]
        self._dev = None

    def regrid_batch_device(self, cells, offsets):
        """Switch every instance to new leaves held on the device (an
        amr.GpuCells, instance k owning leaves [offsets[k], offsets[k+1])):
        the layout is built by kernels and the populations carried over
        without the leaves leaving the device.  Host copies of the leaves
        are made only if a later call needs them (queries by default leaf
        list, mesh properties), from `cells`, which the batch keeps
        referenced: read `meshes` first if `cells` is to be freed or
        overwritten."""
        off = np.ascontiguousarray(offsets, np.uint64)
        assert len(off) == self.ninst + 1 and int(off[-1]) <= cells.n
        check(lib().fl_lbm_regrid_batch_device(self.h, off.ctypes.data, cells.x.p, cells.y.p, cells.z.p,
                                               cells.level.p, cells.state.p))
        self._meshes = self._fluid = None
        self._dev = (cells, [int(v) for v in off])

    def close(self):
        if self.h:
            lib().fl_lbm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def variant_names():
    return [lib().fl_lbm_variant_name(v).decode() for v in range(lib().fl_lbm_num_variants())]


class LbmSolver(LbmBatch):
    """Single-instance solver (a batch of one)."""

    def __init__(self, nx0, ny0, nz0, max_level, tau0, bc, bc_u, x, y, z, level, state,
                 storage="f32", rho_init=1.0, u_init=(0.0, 0.0, 0.0), rho0=1.0, stream=None, graph=False):
        p = make_params(nx0, ny0, nz0, max_level, tau0, bc, bc_u, storage, rho_init, u_init, rho0)
        super().__init__([p], [dict(x=x, y=y, z=z, level=level, state=state)], stream=stream, graph=graph)

    @property
    def x(self):
        return np.asarray(self.meshes[0]["x"], np.int32)

    @property
    def y(self):
        return np.asarray(self.meshes[0]["y"], np.int32)

    @property
    def z(self):
        return np.asarray(self.meshes[0]["z"], np.int32)

    @property
    def level(self):
        return np.asarray(self.meshes[0]["level"], np.uint8)

    @property
    def state(self):
        return np.asarray(self.meshes[0]["state"], np.uint8)

    @property
    def fx(self):
        return self.fluid[0][0]

    @property
    def fy(self):
        return self.fluid[0][1]

    @property
    def fz(self):
        return self.fluid[0][2]

    @property
    def flv(self):
        return self.fluid[0][3]

    def get_f(self, x=None, y=None, z=None, lv=None):
        return super().get_f(0, x, y, z, lv)

    def set_f(self, f, x=None, y=None, z=None, lv=None):
        return super().set_f(f, 0, x, y, z, lv)

    def moments(self, x=None, y=None, z=None, lv=None):
        return super().moments(0, x, y, z, lv)

    def vorticity(self):
        return super().vorticity(0)

    def totals(self):
        m, j = self.totals_batch()
        return float(m[0]), j[0]

    def force(self):
        return self.force_batch()[0]

    def regrid(self, x, y, z, level, state):
        self.regrid_batch([dict(x=x, y=y, z=z, level=level, state=state)])

    def regrid_device(self, cells):
        """regrid from the n leaves of an amr.GpuCells on the device."""
        self.regrid_batch_device(cells, [0, cells.n])
