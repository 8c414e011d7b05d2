"""Lid-driven cavity with vorticity-driven regrid: the step loop with a regrid
every K steps (the AdaptMeshNode / ExecutionGraph role of the reference,
src/runtime/AdaptMeshNode.cpp:9-64, which runs once and has no time loop).

Regrid: |vorticity| at every fluid leaf (device), flags +1 where it exceeds
frac * max and the level is below max_level, -1 where it does not and the
level is above 0; one device regrid pass (D3Q19 neighbour-cube balance,
merges that would break it are dropped); the solver carries its
populations over to the new leaves.
"""
from __future__ import annotations

import numpy as np

from . import amr, meshes
from .solver import LbmSolver


def cavity_params(n0, lid):
    bc = ["wall"] * 6
    bcu = [(0.0, 0.0, 0.0)] * 5 + [(lid, 0.0, 0.0)]
    return bc, bcu


def regrid_flags(level, state, w, frac, max_level):
    fl = state == 0
    flags = np.zeros(len(level), np.int32)
    if not fl.any():
        return flags
    wmax = w.max() if len(w) else 0.0
    if wmax <= 0:
        return flags
    lv = level[fl]
    f = np.where((w > frac * wmax) & (lv < max_level), 1, np.where((w <= frac * wmax) & (lv > 0), -1, 0))
    flags[fl] = f
    return flags


class Cavity:
    def __init__(self, n0=32, max_level=1, tau=0.6, lid=0.05, frac=0.5, storage="f32"):
        self.n0, self.max_level, self.frac = n0, max_level, frac
        bc, bcu = cavity_params(n0, lid)
        m = meshes.uniform(n0, n0, n0)
        self.solver = LbmSolver(n0, n0, n0, max_level, tau, bc, bcu, m["x"], m["y"], m["z"], m["level"],
                                m["state"], storage=storage)
        self.regrids = 0

    def regrid(self):
        s = self.solver
        w = s.vorticity()
        flags = regrid_flags(s.level, s.state, w, self.frac, self.max_level)
        cells = amr.GpuCells.from_host(x=s.x, y=s.y, z=s.z, level=s.level, state=s.state, flag=flags)
        new, st = amr.regrid(cells, balance_mode=1, max_split_level=self.max_level, root=(self.n0,) * 3,
                             merge_balance_check=1)
        h = new.to_host()
        cells.free()
        new.free()
        s.regrid(h["x"], h["y"], h["z"], h["level"], h["state"])
        self.regrids += 1
        return st

    def run(self, steps, regrid_every=50):
        for k in range(steps):
            if regrid_every and k > 0 and k % regrid_every == 0:
                self.regrid()
            self.solver.step(1)
