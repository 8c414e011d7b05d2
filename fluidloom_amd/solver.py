"""AMR D3Q19 lattice-Boltzmann solver over the fl_lbm_* C ABI.

Mirrors the reference's run loop (KernelNode::execute over SOAFieldManager
fields, src/runtime/KernelNode.cpp:37-106) with the build's own D3Q19 step.
All arithmetic runs in libfluidloom_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (FL_BC_OUTFLOW, FL_BC_PERIODIC, FL_BC_WALL, FL_STORE_F16, FL_STORE_F32,
                   FL_STORE_F64, LbmParams, LbmStats, check, lib)

STORAGE = {"f32": FL_STORE_F32, "f16": FL_STORE_F16, "f64": FL_STORE_F64}
BC = {"wall": FL_BC_WALL, "outflow": FL_BC_OUTFLOW, "periodic": FL_BC_PERIODIC}


class LbmSolver:
    def __init__(self, nx0, ny0, nz0, max_level, tau0, bc, bc_u, x, y, z, level, state,
                 storage="f32", rho_init=1.0, u_init=(0.0, 0.0, 0.0), rho0=1.0, stream=None):
        p = LbmParams()
        p.nx0, p.ny0, p.nz0, p.max_level = nx0, ny0, nz0, max_level
        p.tau0, p.rho0, p.rho_init = tau0, rho0, rho_init
        p.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
        for f in range(6):
            p.bc[f] = BC[bc[f]] if isinstance(bc[f], str) else int(bc[f])
            for a in range(3):
                p.bc_u[f][a] = float(bc_u[f][a])
        for a in range(3):
            p.u_init[a] = float(u_init[a])
        self.params = p
        self.x, self.y, self.z = (np.ascontiguousarray(a, np.int32) for a in (x, y, z))
        self.level = np.ascontiguousarray(level, np.uint8)
        self.state = np.ascontiguousarray(state, np.uint8)
        self.h = C.c_void_p()
        check(lib().fl_lbm_create(C.byref(p), self.x.ctypes.data, self.y.ctypes.data,
                                  self.z.ctypes.data, self.level.ctypes.data,
                                  self.state.ctypes.data, len(self.x), stream, C.byref(self.h)))
        fl = self.state == 0
        self.fx, self.fy, self.fz, self.flv = (np.ascontiguousarray(a) for a in
                                               (self.x[fl], self.y[fl], self.z[fl], self.level[fl]))

    def step(self, n=1):
        check(lib().fl_lbm_step(self.h, int(n)))

    def sync(self):
        s = C.c_void_p()
        check(lib().fl_lbm_stream(self.h, C.byref(s)))
        check(lib().fl_stream_sync(s))

    def stats(self) -> LbmStats:
        st = LbmStats()
        check(lib().fl_lbm_stats_get(self.h, C.byref(st)))
        return st

    def _q(self, x, y, z, lv):
        if x is None:
            return self.fx, self.fy, self.fz, self.flv
        return tuple(np.ascontiguousarray(a, dt) for a, dt in
                     ((x, np.int32), (y, np.int32), (z, np.int32), (lv, np.uint8)))

    def get_f(self, x=None, y=None, z=None, lv=None):
        x, y, z, lv = self._q(x, y, z, lv)
        out = np.empty((len(x), 19), np.float64)
        check(lib().fl_lbm_get_f(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data, lv.ctypes.data,
                                 len(x), out.ctypes.data))
        return out

    def set_f(self, f, x=None, y=None, z=None, lv=None):
        x, y, z, lv = self._q(x, y, z, lv)
        f = np.ascontiguousarray(f, np.float64)
        check(lib().fl_lbm_set_f(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data, lv.ctypes.data,
                                 len(x), f.ctypes.data))

    def moments(self, x=None, y=None, z=None, lv=None):
        x, y, z, lv = self._q(x, y, z, lv)
        rho = np.empty(len(x))
        u = np.empty((len(x), 3))
        check(lib().fl_lbm_moments(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                   lv.ctypes.data, len(x), rho.ctypes.data, u.ctypes.data))
        return rho, u

    def totals(self):
        m = np.zeros(1)
        j = np.zeros(3)
        check(lib().fl_lbm_totals(self.h, m.ctypes.data, j.ctypes.data))
        return float(m[0]), j

    def force(self):
        F = np.zeros(3)
        check(lib().fl_lbm_force(self.h, F.ctypes.data))
        return F

    def vorticity(self):
        out = np.empty(len(self.fx))
        check(lib().fl_lbm_vorticity(self.h, self.fx.ctypes.data, self.fy.ctypes.data, self.fz.ctypes.data,
                                     self.flv.ctypes.data, len(self.fx), out.ctypes.data))
        return out

    def regrid(self, x, y, z, level, state):
        """Switch to a new leaf list, carrying the populations over."""
        x, y, z = (np.ascontiguousarray(a, np.int32) for a in (x, y, z))
        level = np.ascontiguousarray(level, np.uint8)
        state = np.ascontiguousarray(state, np.uint8)
        check(lib().fl_lbm_regrid(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data, level.ctypes.data,
                                  state.ctypes.data, len(x)))
        self.x, self.y, self.z, self.level, self.state = x, y, z, level, state
        fl = state == 0
        self.fx, self.fy, self.fz, self.flv = (np.ascontiguousarray(a) for a in (x[fl], y[fl], z[fl], level[fl]))

    def time_collide(self, level, iters):
        ms = C.c_float(0)
        check(lib().fl_lbm_time_collide(self.h, int(level), int(iters), C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            lib().fl_lbm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
