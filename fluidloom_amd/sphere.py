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

With --compare the same sphere runs in FP64 storage (the GPU oracle run)
and in each listed storage form side by side, from the same start: C_d per
form along the run, and at the end each form's velocity relative L2, density
error, C_d error and mass drift against the FP64 run.

    python -m fluidloom_amd.sphere --steps 5000 --compare f32,f16,f16c
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
    s = _solver(c, m, storage)
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


def _solver(c, m, storage):
    return LbmSolver(c["nx0"], c["ny0"], c["nz0"], c["levels"], c["tau0"], c["bc"], c["bcu"], m["x"], m["y"],
                     m["z"], m["level"], m["state"], storage=storage, u_init=(c["u_in"], 0.0, 0.0))


def precision(steps=5000, report_every=500, scale=1, forms=("f32", "f16", "f16c"), re=100.0, tune=True):
    """Every storage form against the FP64 run of the same sphere.  All the
    solvers are resident at once (full size: about 47 GiB for FP64, FP32,
    FP16 and FP16C) and step in turns of report_every coarse steps, replayed
    from a hipGraph and, with tune, on the autotuner's variants."""
    check(lib().fl_init(0))
    c = sphere_case(scale, re=re)
    m = meshes.sphere(c["nx0"], c["ny0"], c["nz0"], c["ctr"], c["r"], 2 * c["r"], c["levels"], log=log)
    sols = {}
    for st in ("f64",) + tuple(forms):
        s = _solver(c, m, st)
        if tune:
            s.autotune(3)
        s.set_graph(True)
        sols[st] = s
    upd = int(sols["f64"].stats().updates_per_step)
    mass0 = {st: s.totals()[0] for st, s in sols.items()}
    area = np.pi * c["r"] ** 2
    secs = {st: 0.0 for st in sols}
    hist = []
    done = 0
    while done < steps:
        n = min(report_every, steps - done)
        row = dict(step=done + n)
        for st, s in sols.items():
            s.sync()
            t0 = time.perf_counter()
            s.step(n)
            s.sync()
            secs[st] += time.perf_counter() - t0
            row[st] = round(float(s.force()[0] / (0.5 * c["u_in"] ** 2 * area)), 6)
        done += n
        hist.append(row)
        log(f"step {done}: C_d " + ", ".join(f"{st} {row[st]:.5f}" for st in sols))
    rho64, u64 = sols["f64"].moments()
    nu64 = np.linalg.norm(u64)
    res = {}
    for st, s in sols.items():
        mass = s.totals()[0]
        r = dict(mlups=round(upd * steps / secs[st] / 1e6, 1), cd=hist[-1][st],
                 mass_drift=float((mass - mass0[st]) / mass0[st]))
        if st != "f64":
            rho, u = s.moments()
            r.update(vel_l2=float(np.linalg.norm(u - u64) / nu64),
                     rho_l2=float(np.linalg.norm(rho - rho64) / np.linalg.norm(rho64)),
                     cd_rel_err=float(abs(hist[-1][st] - hist[-1]["f64"]) / abs(hist[-1]["f64"])))
            del rho, u
        res[st] = r
        log(f"{st}: {r}")
    for s in sols.values():
        s.close()
    return dict(workload=f"flow past a sphere: base {c['nx0']}x{c['ny0']}x{c['nz0']}, r={c['r']:g}, "
                         f"{c['levels']} levels within 2r, Re={re:g}, u_in={c['u_in']}; every form against the "
                         f"FP64 run",
                steps=steps, leaves=int(len(m["x"])), results=res, history=hist)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--report-every", type=int, default=500)
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "f16c", "f64"])
    ap.add_argument("--re", type=float, default=100.0)
    ap.add_argument("--compare", default=None,
                    help="comma-separated storage forms to run beside the FP64 run (e.g. f32,f16,f16c)")
    a = ap.parse_args()
    if a.compare:
        out = precision(a.steps, a.report_every, a.scale, tuple(a.compare.split(",")), a.re)
    else:
        out = run(a.steps, a.report_every, a.scale, a.storage, a.re)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
