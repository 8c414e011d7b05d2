#!/usr/bin/env python3
"""AMR D3Q19 benchmark: flow past a sphere on one MI355X.

Workload (the single-GPU configuration of the metric): base grid 256x128x128,
sphere of radius 16 cells centred at (64,64,64), 3 refinement levels within
2 radii of the centre (D3Q19 neighbour-cube 2:1 balance), velocity inlet
u = 0.05 on -x, outflow on +x, bounce-back walls elsewhere, tau0 for Re = 100
on the diameter, a device regrid pass every 100 steps.  A "step" is one coarse
(level-0) time step with all level sub-cycles; MLUPS counts every leaf
update (level l leaves update 2^l times per step).

Every regrid_every steps the timed loop runs a full regrid on the device:
refinement marks, the regrid pass (fl_regrid: D3Q19 balance, merge,
compaction, split), then the solver's lattice layout rebuilt by kernels
from the pass's leaves and the populations carried over
(fl_lbm_regrid_batch_device).  The sphere's mesh is static, so the pass
returns the same leaves (the bench checks that it split and merged
nothing) and the rebuild is done anyway, as a regrid that changed the mesh
would do it; --regrid-keep-layout keeps the layout when the pass changed
nothing.  The pass and the rebuild are also timed alone, before the loop.

Prints one JSON line.  `value` is whole-job MLUPS; `roofline` is the level-3
collide-stream kernel (the dominant kernel) measured with HIP events on the
solver's stream, with `traffic` (HBM bytes per launch) taken from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this command;
`cpu_baseline` is the CPU port of the lattice step in FP32 (the oracle's
FP32 build, OpenMP) on a bounded quarter-size instance of the same
configuration.

`amr_integer` (after the timed loop; --no-amr skips it) times the AMR integer
path on the device -- the 64M-cell micro-benchmark's stages (Hilbert and
rot/dir keys, the 64-bit sort with the digit passes it ran, scan, regrid
pass, face-neighbour rebuild, hash build / query) and a cascading D3Q19
balance on this sphere mesh -- with the oracle's single-threaded C
restatement of each stage timed beside it on bounded samples.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "AMR D3Q19 MLUPS + HBM GB/s; batched solves/s & p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"f32": 152, "f16": 76, "f16c": 76, "f64": 304}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sphere_params(nx0, ny0, nz0, r, u_in, re):
    nu = u_in * 2 * r / re
    tau0 = 3 * nu + 0.5
    bc = ["wall", "outflow", "wall", "wall", "wall", "wall"]
    bcu = [(u_in, 0.0, 0.0)] + [(0.0, 0.0, 0.0)] * 5
    return tau0, bc, bcu


def cpu_baseline(budget_s=15.0):
    """The lattice step on the host in FP32 (the oracle's FP32 build, all
    host cores) on the quarter-size sphere: base 64x32x32, r = 4, 3 levels
    within 2r, Re 100 -- the same workload at 1/64 of the leaves.  Its leaves
    come from the device mesh builder; only the stepping is timed."""
    from fluidloom_amd import meshes
    from oracle import oracle as O
    nx0, ny0, nz0, r = 64, 32, 32, 4.0
    ctr = (16.0, 16.0, 16.0)
    m = meshes.sphere(nx0, ny0, nz0, ctr, r, 2 * r, 3)
    tau0, bc, bcu = sphere_params(nx0, ny0, nz0, r, 0.05, 100.0)
    bci = [{"wall": 0, "outflow": 1}[b] for b in bc]
    s = O.Lbm(nx0, ny0, nz0, 3, tau0, bci, bcu, m["x"], m["y"], m["z"], m["level"], m["state"],
              u_init=(0.05, 0, 0), single=True)
    fluid = np.asarray(m["state"]) == 0
    lv = np.asarray(m["level"])
    upd = int(sum(int((fluid & (lv == l)).sum()) << l for l in range(4)))
    s.step(1)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        s.step(1)
        steps += 1
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    return {"value": round(upd * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"FP32 CPU port of the lattice step (oracle FP32 build, OpenMP), sphere "
                      f"{nx0}x{ny0}x{nz0} r={r:g} 3 levels ({upd} leaf updates/step), {steps} steps in {dt:.1f} s"}


def cpu_amr_baseline():
    """The AMR integer stages on the host: the oracle's C restatement
    (single-threaded) on bounded samples of the 64M-cell workload's
    generator -- 4M cells for the streaming stages, 1M for the face
    neighbours, 256K for the regrid pass, whose reference-rule balance keeps
    every holder of an anchor in one probe chain (as the reference's index
    table does) and grows faster than linearly on these seeded cells."""
    from oracle import oracle as O

    def timed(f):
        t0 = time.perf_counter()
        r = f()
        return r, time.perf_counter() - t0

    out = {"cores": 1, "kind": "port", "unit": "M cells/s"}
    rates = {}
    n = 1 << 22
    x, y, z, lv = O.gen_cells(42, n, 10)
    _, dt = timed(lambda: O.rotdir_encode(x, y, z, 8))
    rates["rotdir"] = n / dt / 1e6
    k, dt = timed(lambda: O.key63(x, y, z, lv))
    rates["keygen"] = n / dt / 1e6
    _, dt = timed(lambda: O.sort_pairs(k, np.arange(n, dtype=np.uint32)))
    rates["sort"] = n / dt / 1e6
    _, dt = timed(lambda: O.scan_exclusive(np.arange(n, dtype=np.uint32)))
    rates["scan"] = n / dt / 1e6
    rk = np.random.default_rng(42).integers(0, 1 << 63, n, dtype=np.uint64)
    cap = O.hash_capacity(n)
    (tk, tv, _), dt = timed(lambda: O.hash_build(rk, np.arange(n, dtype=np.uint32), cap))
    rates["hash_build"] = n / dt / 1e6
    _, dt = timed(lambda: O.hash_query(tk, tv, cap, rk))
    rates["hash_query"] = n / dt / 1e6
    m = 1 << 20
    c = O.CellSet(x=x[:m], y=y[:m], z=z[:m], level=lv[:m])
    _, dt = timed(lambda: O.face_neighbors(c))
    rates["neighbors"] = m / dt / 1e6
    r = 1 << 18
    c = O.gen_flags(O.CellSet(x=x[:r], y=y[:r], z=z[:r], level=lv[:r]), 42, 0.05, 0.03)
    _, dt = timed(lambda: O.regrid(c, 0, 10, 10))
    rates["regrid"] = r / dt / 1e6
    out["rates"] = {k2: round(v, 2) for k2, v in rates.items()}
    out["sample"] = (f"oracle C restatement, 1 thread: {n} seeded cells (levels 0-10, 2^10 root) for the codecs, "
                     f"sort, scan and hash; {m} for the face neighbours; {r} for the regrid pass")
    return out


def amr_stages(mesh, ctr, r, root, log):
    """The AMR integer path on the device: the 64M-cell micro-benchmark's
    stages (best of two runs) and a cascading D3Q19 balance on this sphere
    mesh."""
    from fluidloom_amd import amr, amr_bench
    runs = [amr_bench.run() for _ in range(2)]  # the first also loads each kernel's code
    res = dict(runs[-1])
    for k in res:
        if k.endswith("_ms"):
            res[k] = min(r[k] for r in runs)
    res["sort_gbs"] = max(r["sort_gbs"] for r in runs)
    res["scan_gbs"] = max(r["scan_gbs"] for r in runs)
    cells = amr.GpuCells.from_host(**{k: mesh[k] for k in ("x", "y", "z", "level", "state")})
    casc = amr_bench.cascade(cells, ctr, 2 * r, 4, root)
    cells.free()
    log(f"amr stages: {res}; cascade {casc}")
    keep = ("rotdir_ms", "keygen_ms", "sort_ms", "sort_passes", "sort_gbs", "scan_ms", "scan_gbs", "regrid_ms",
            "cells_out", "neighbors_ms", "hash_build_ms", "hash_query_ms")
    out = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in res.items() if k in keep}
    out["rates_mcells_s"] = {
        "rotdir": round(res["cells"] / res["rotdir_ms"] / 1e3, 1),
        "keygen": round(res["cells"] / res["keygen_ms"] / 1e3, 1),
        "sort": round(res["cells"] / res["sort_ms"] / 1e3, 1),
        "scan": round(res["cells"] / res["scan_ms"] / 1e3, 1),
        "hash_build": round(res["cells"] / res["hash_build_ms"] / 1e3, 1),
        "hash_query": round(res["cells"] / res["hash_query_ms"] / 1e3, 1),
        "neighbors": round(res["cells_out"] / res["neighbors_ms"] / 1e3, 1),
        "regrid": round(res["cells"] / res["regrid_ms"] / 1e3, 1)}
    out["cascade_balance"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in casc.items()}
    return out


def copy_bandwidth(_lib, L, stream, nbytes=2 << 30, iters=10):
    C = _lib.C
    a, b = _lib.DeviceArray(nbytes, np.uint8), _lib.DeviceArray(nbytes, np.uint8)
    e0, e1 = C.c_void_p(), C.c_void_p()
    _lib.check(L.fl_event_create(C.byref(e0)))
    _lib.check(L.fl_event_create(C.byref(e1)))
    _lib.check(L.fl_copy_kernel(b.ptr, a.ptr, nbytes, stream))
    _lib.check(L.fl_event_record(e0, stream))
    for _ in range(iters):
        _lib.check(L.fl_copy_kernel(b.ptr, a.ptr, nbytes, stream))
    _lib.check(L.fl_event_record(e1, stream))
    _lib.check(L.fl_stream_sync(stream))
    ms = C.c_float(0)
    _lib.check(L.fl_event_elapsed_ms(e0, e1, C.byref(ms)))
    a.free()
    b.free()
    return 2 * nbytes * iters / (ms.value * 1e-3) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "f16c", "f64"])
    ap.add_argument("--scale", type=int, default=1, help="divide the base grid and radii by this")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-amr", action="store_true", help="skip the AMR integer stages (64M cells, cascade)")
    ap.add_argument("--regrid-every", type=int, default=100)
    ap.add_argument("--regrid-keep-layout", action="store_true",
                    help="keep the solver's layout when the regrid pass changed no leaf")
    ap.add_argument("--no-graph", action="store_true", help="step eagerly instead of replaying a hipGraph")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default collide variant")
    ap.add_argument("--variant", default=None, help="force this collide variant on every level (no autotune)")
    args = ap.parse_args()
    if args.steps < 1:
        print("bench.py: --steps must be at least 1", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    from fluidloom_amd import _lib, amr, meshes
    from fluidloom_amd.solver import LbmSolver

    L = _lib.lib()
    _lib.check(L.fl_init(0))
    sc = args.scale
    nx0, ny0, nz0 = 256 // sc, 128 // sc, 128 // sc
    r = 16.0 / sc
    ctr = (64.0 / sc, 64.0 / sc, 64.0 / sc)
    u_in, re, levels = 0.05, 100.0, 3
    tau0, bc, bcu = sphere_params(nx0, ny0, nz0, r, u_in, re)

    t0 = time.perf_counter()
    mesh = meshes.sphere(nx0, ny0, nz0, ctr, r, 2 * r, levels, log=log)
    t1 = time.perf_counter()
    log(f"mesh built on device in {t1 - t0:.1f} s")
    solver = LbmSolver(nx0, ny0, nz0, levels, tau0, bc, bcu, mesh["x"], mesh["y"], mesh["z"], mesh["level"],
                       mesh["state"], storage=args.storage, u_init=(u_in, 0.0, 0.0))
    t2 = time.perf_counter()
    st = solver.stats()
    leaves = [int(st.leaves[l]) for l in range(levels + 1)]
    upd = int(st.updates_per_step)
    log(f"solver layout built in {t2 - t1:.1f} s: leaves/level {leaves}, bricks/level "
        f"{[int(st.bricks[l]) for l in range(levels + 1)]}, {upd} updates/step, "
        f"{st.device_bytes / 2**30:.2f} GiB")
    from fluidloom_amd.solver import variant_names
    if args.variant:
        for l in range(levels + 1):
            solver.set_variant(l, variant_names().index(args.variant))
    elif not args.no_autotune:
        chosen = solver.autotune(6)
        log(f"autotuned collide variants per level: {[variant_names()[v] for v in chosen]}")
    if not args.no_graph:
        solver.set_graph(True)
    # leaves on the device for the periodic regrid pass
    cells = amr.GpuCells.from_host(**{k: mesh[k] for k in ("x", "y", "z", "level", "state")})
    # the pass's output and workspace, allocated once
    regrid_out = amr.GpuCells(cells.n + cells.n // 8 + 64)
    regrid_tmp = amr.regrid_temp(cells.n)

    def regrid_pass():
        """Refinement marks + one device regrid pass; the sphere's mesh is
        static, so the pass must return the leaves unchanged."""
        marked = amr.mark_ball(cells, ctr, 2 * r, levels)
        if marked:
            raise RuntimeError("the sphere's mesh changed at a regrid; a static mesh is expected")
        new, rst = amr.regrid(cells, balance_mode=1, max_split_level=levels, root=(nx0, ny0, nz0),
                              merge_balance_check=1, out=regrid_out, temp=regrid_tmp)
        same = rst.num_split == 0 and rst.num_groups == 0 and new.n == cells.n
        if not same:
            raise RuntimeError("the sphere's mesh changed at a regrid; a static mesh is expected")
        return new

    def full_regrid():
        """The regrid pass, then the solver's layout rebuilt on the device
        from its leaves with the populations carried over."""
        new = regrid_pass()
        if not args.regrid_keep_layout:
            solver.regrid_device(new)
            return 1
        return 0

    # untimed: one regrid pass alone, and one layout rebuild + population
    # transfer from the pass's leaves alone
    _lib.check(L.fl_device_sync())
    t_r0 = time.perf_counter()
    new = regrid_pass()
    _lib.check(L.fl_device_sync())
    regrid_pass_s = time.perf_counter() - t_r0
    rebuild_ms = []
    for _ in range(3):  # the first call also allocates what later ones reuse
        t_r0 = time.perf_counter()
        solver.regrid_device(new)
        solver.sync()
        rebuild_ms.append((time.perf_counter() - t_r0) * 1e3)
    log(f"regrid pass {regrid_pass_s * 1e3:.1f} ms; solver layout rebuild + population transfer from "
        f"the pass's leaves on the device: {', '.join(f'{t:.1f}' for t in rebuild_ms)} ms")

    solver.step(args.warmup)
    if not args.no_graph:
        solver.step(2)  # untimed: captures the graph for the timed loop's buffer parity
    solver.sync()
    ev0, ev1 = _lib.C.c_void_p(), _lib.C.c_void_p()
    stream = _lib.C.c_void_p()
    _lib.check(L.fl_lbm_stream(solver.h, _lib.C.byref(stream)))
    _lib.check(L.fl_event_create(_lib.C.byref(ev0)))
    _lib.check(L.fl_event_create(_lib.C.byref(ev1)))
    _lib.check(L.fl_device_sync())
    regrids = rebuilds = 0
    w0 = time.perf_counter()
    _lib.check(L.fl_event_record(ev0, stream))
    k = 0
    while k < args.steps:
        n = min(10, args.steps - k)
        if args.regrid_every:
            n = min(n, args.regrid_every - k % args.regrid_every)
        solver.step(n)
        k += n
        # the full regrid after every regrid_every-th step, inside the timed region
        if args.regrid_every and k % args.regrid_every == 0:
            rebuilds += full_regrid()
            regrids += 1
        if k % 50 == 0:
            log(f"step {k}/{args.steps}")
    _lib.check(L.fl_event_record(ev1, stream))
    _lib.check(L.fl_device_sync())
    w1 = time.perf_counter()
    ms = _lib.C.c_float(0)
    _lib.check(L.fl_event_elapsed_ms(ev0, ev1, _lib.C.byref(ms)))
    wall = w1 - w0
    mlups = upd * args.steps / wall / 1e6
    bpu = BYTES_PER_UPDATE[args.storage]
    F = solver.force()
    cd = F[0] / (0.5 * 1.0 * u_in ** 2 * np.pi * r ** 2)
    mass, mom = solver.totals()
    st2 = solver.stats()

    # roofline of the dominant kernel: level-3 collide-stream alone
    iters = 20
    ms_k = solver.time_collide(levels, iters)
    alg_bytes = leaves[levels] * bpu
    achieved = alg_bytes / (ms_k * 1e-3) / 1e9
    log(f"level-{levels} collide: {ms_k:.3f} ms/launch, {achieved:.0f} GB/s algorithmic")
    variant_ms = {}
    for v, name in enumerate(variant_names()):
        variant_ms[name] = round(solver.time_collide(levels, 10, v), 4)
    log("finest-level collide variants (ms/launch): " + ", ".join(f"{k} {t:.3f}" for k, t in variant_ms.items()))
    finest_variant = variant_names()[int(solver.stats().variant[levels])]
    traffic = None
    tf = ROOT / "profiles" / f"k_collide_sphere_{args.storage}_traffic.json"
    if tf.exists() and sc == 1:
        t = json.loads(tf.read_text())
        if t.get("variant") == finest_variant:
            traffic = int(t["traffic_bytes"])
        elif finest_variant in t.get("per_variant_traffic_bytes", {}):
            # the same counter runs cover every variant (the sweep below runs them all)
            traffic = int(t["per_variant_traffic_bytes"][finest_variant])

    # bandwidth reference: a plain 16 B/lane copy kernel over 2 x 2 GiB
    copy_gbs = copy_bandwidth(_lib, L, stream)
    log(f"plain copy kernel: {copy_gbs:.0f} GB/s")

    amr_out = None
    if not args.no_amr:
        log("AMR integer stages (64M seeded cells; cascading balance on the sphere mesh)...")
        amr_out = amr_stages(mesh, ctr, r, (nx0, ny0, nz0), log)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, bounded sample)...")
        cpu = cpu_baseline()
        if amr_out is not None:
            amr_out["cpu_baseline"] = cpu_amr_baseline()

    out = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic",
        "config": {"workload": f"flow past a sphere: base {nx0}x{ny0}x{nz0}, r={r:g} at {ctr}, "
                               f"{levels} levels within {2 * r:g} cells, Re={re:g}, u_in={u_in}, "
                               f"tau0={tau0:.4f}, device regrid (pass"
                               f"{'' if args.regrid_keep_layout else ' + layout rebuild + population transfer'}"
                               f") every {args.regrid_every} steps",
                   "leaves_per_level": leaves, "updates_per_step": upd, "storage": args.storage,
                   "parallelism": "single GPU"},
        "hbm_gbps_algorithmic": round(mlups * 1e6 * bpu / 1e9, 1),
        "gpu_event_ms": round(ms.value, 3),
        "regrid_passes": regrids,
        "layout_rebuilds": rebuilds,
        "regrid_pass_ms": round(regrid_pass_s * 1e3, 2),
        "layout_rebuild_ms": round(float(np.median(rebuild_ms)), 2),
        "layout_rebuild_first_ms": round(rebuild_ms[0], 2),
        "collide_variants": [variant_names()[int(st2.variant[l])] for l in range(levels + 1)],
        "graph": bool(st2.graph),
        "drag_coefficient": round(float(cd), 4),
        "mass": mass,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": (f"profiles/{tf.name}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of "
                                        f"this command, scaled by the copy-kernel calibration")
                                       if traffic is not None else None,
                     "kernel": f"k_collide {finest_variant} level {levels} ({leaves[levels]} leaves x {bpu} B)"},
        "copy_kernel_gbs": round(copy_gbs, 1),
        "finest_variant_ms": variant_ms,
        "amr_integer": amr_out,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
