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


def main():
    """The reference's CPU-runnable case: python -m fluidloom_amd.cavity
    (32^3, one extra level where |w| > 0.5 max, regrid every 50, 500 steps)."""
    import argparse
    import json
    import sys
    import time

    from ._lib import check, lib

    ap = argparse.ArgumentParser()
    ap.add_argument("--n0", type=int, default=32)
    ap.add_argument("--levels", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--regrid-every", type=int, default=50)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "f16c", "f64"])
    a = ap.parse_args()
    check(lib().fl_init(0))
    c = Cavity(n0=a.n0, max_level=a.levels, storage=a.storage)
    upd = 0
    t_step = t_regrid = 0.0
    for k in range(a.steps):
        if a.regrid_every and k > 0 and k % a.regrid_every == 0:
            c.solver.sync()
            t0 = time.perf_counter()
            c.regrid()
            t_regrid += time.perf_counter() - t0
        t0 = time.perf_counter()
        c.solver.step(1)
        c.solver.sync()
        t_step += time.perf_counter() - t0
        upd += int(c.solver.stats().updates_per_step)
    m, _ = c.solver.totals()
    out = dict(workload=f"lid-driven cavity {a.n0}^3, lid 0.05, tau 0.6, {a.levels} extra level(s) where |w| > 0.5 max, "
                        f"regrid every {a.regrid_every}, {a.steps} steps, {a.storage}",
               mlups_stepping=round(upd / t_step / 1e6, 1), step_seconds=round(t_step, 3),
               regrids=c.regrids, regrid_seconds=round(t_regrid, 3),
               leaves_per_level=[int((c.solver.level[c.solver.state == 0] == l).sum()) for l in range(a.levels + 1)],
               mass=m)
    print(json.dumps(out), flush=True)
    print(f"[cavity] {out['mlups_stepping']} MLUPS stepping, {c.regrids} regrids in {t_regrid:.2f} s", file=sys.stderr)


if __name__ == "__main__":
    main()
