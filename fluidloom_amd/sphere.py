"""Flow past a sphere with the drag coefficient sampled along the run.

The configuration of the benchmark: base grid 256x128x128 / scale, sphere of
radius 16 / scale cells centred at 64 / scale on each axis, 3 refinement
levels within 2 radii, velocity inlet u_in on -x, outflow on +x, no-slip
walls on the four sides, tau0 for the Reynolds number on the diameter,
started from the uniform inlet velocity.  Every `report_every` coarse steps
the momentum-exchange force on the solid links gives
C_d = F_x / (1/2 rho u_in^2 pi r^2), printed with the total mass, so that a
run shows how far the drag has settled.

    python -m fluidloom_amd.sphere --steps 20000 --report-every 1000 --scale 2
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import meshes
from ._lib import check, lib
from .solver import LbmSolver


def log(msg):
    print(f"[sphere] {msg}", file=sys.stderr, flush=True)


def sphere_case(scale=1, u_in=0.05, re=100.0, levels=3):
    nx0, ny0, nz0 = 256 // scale, 128 // scale, 128 // scale
    r = 16.0 / scale
    ctr = (64.0 / scale,) * 3
    nu = u_in * 2 * r / re
    tau0 = 3 * nu + 0.5
    bc = ["wall", "outflow", "wall", "wall", "wall", "wall"]
    bcu = [(u_in, 0.0, 0.0)] + [(0.0, 0.0, 0.0)] * 5
    return dict(nx0=nx0, ny0=ny0, nz0=nz0, r=r, ctr=ctr, tau0=tau0, bc=bc, bcu=bcu, u_in=u_in, levels=levels)


def run(steps=5000, report_every=500, scale=1, storage="f32", re=100.0):
    check(lib().fl_init(0))
    c = sphere_case(scale, re=re)
    m = meshes.sphere(c["nx0"], c["ny0"], c["nz0"], c["ctr"], c["r"], 2 * c["r"], c["levels"], log=log)
    s = LbmSolver(c["nx0"], c["ny0"], c["nz0"], c["levels"], c["tau0"], c["bc"], c["bcu"], m["x"], m["y"], m["z"],
                  m["level"], m["state"], storage=storage, u_init=(c["u_in"], 0.0, 0.0))
    s.autotune(3)
    s.set_graph(True)
    area = np.pi * c["r"] ** 2
    hist = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(report_every, steps - done)
        s.step(n)
        done += n
        f = s.force()
        mass, _ = s.totals()
        cd = float(f[0] / (0.5 * c["u_in"] ** 2 * area))
        hist.append(dict(step=done, cd=round(cd, 5), mass=mass))
        log(f"step {done}: C_d {cd:.5f}, mass {mass:.6e}, {time.perf_counter() - t0:.1f} s")
    # flow-throughs: coarse steps x u_in / domain length (coarse cells)
    return dict(workload=f"flow past a sphere: base {c['nx0']}x{c['ny0']}x{c['nz0']}, r={c['r']:g}, "
                         f"{c['levels']} levels within 2r, Re={re:g}, u_in={c['u_in']}, {storage}",
                steps=steps, flow_throughs=round(steps * c["u_in"] / c["nx0"], 3), history=hist)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--report-every", type=int, default=500)
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "f16c", "f64"])
    ap.add_argument("--re", type=float, default=100.0)
    a = ap.parse_args()
    print(json.dumps(run(a.steps, a.report_every, a.scale, a.storage, a.re)), flush=True)


if __name__ == "__main__":
    main()
