"""Batched + precision workload: many independent lid-driven cavities of
different sizes, lid speeds, relaxation times and refinement depths held in
one device layout and stepped together (one collide-stream launch per level
for the whole batch, replayed from a hipGraph, collide variant per level
chosen by the autotuner).  The same batch runs with FP32 storage, FP16
(f - w) storage and FP64 storage; the FP64 run is the oracle the other two
are measured against (velocity relative L2 per instance, mass drift).

"Plain" is the unbatched form: the same instances solved one after another
as single-instance FP32 solvers, stepped eagerly with the default variant.
"Plain with graph" gives each single-instance solver the autotuner and
hipGraph replay as well, so that the gain of batching alone shows.

    python -m fluidloom_amd.batched --instances 256 --steps 2000
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import amr, meshes
from ._lib import check, lib
from .cavity import regrid_flags
from .solver import LbmBatch, LbmSolver, make_params, variant_names


def log(msg):
    print(f"[batched] {msg}", file=sys.stderr, flush=True)


def instance_specs(n, seed=42, nmin=16, nmax=96):
    """Seeded instance draws: base size log-uniform in nmin..nmax (rounded to
    a multiple of 4, the brick edge), lid speed U(0.02, 0.10), tau U(0.55,
    0.9), refinement levels U{0, 1, 2}.  Generator: numpy MT19937(seed)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    specs = []
    for _ in range(n):
        size = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        n0 = int(max(nmin, 4 * round(size / 4)))
        specs.append(dict(n0=n0, lid=float(rng.uniform(0.02, 0.10)), tau=float(rng.uniform(0.55, 0.9)),
                          levels=int(rng.integers(0, 3))))
    return specs


def cavity_params(spec, storage):
    bcu = [(0.0, 0.0, 0.0)] * 5 + [(spec["lid"], 0.0, 0.0)]
    n0 = spec["n0"]
    return make_params(n0, n0, n0, spec["levels"], spec["tau"], ["wall"] * 6, bcu, storage)


def build_meshes(specs):
    return [meshes.cavity(s["n0"], s["levels"]) for s in specs]


def regrid_meshes(b64: LbmBatch, specs, frac=0.5):
    """New leaf sets from the FP64 batch's vorticity (flag cells above
    frac * max, unflag below; one device regrid pass per instance)."""
    out = []
    for k, s in enumerate(specs):
        m = b64.meshes[k]
        if s["levels"] == 0:
            out.append(m)
            continue
        w = b64.vorticity(k)
        flags = regrid_flags(np.asarray(m["level"]), np.asarray(m["state"]), w, frac, s["levels"])
        cells = amr.GpuCells.from_host(x=m["x"], y=m["y"], z=m["z"], level=m["level"], state=m["state"], flag=flags)
        new, _ = amr.regrid(cells, balance_mode=1, max_split_level=s["levels"], root=(s["n0"],) * 3,
                            merge_balance_check=1)
        h = new.to_host()
        cells.free()
        new.free()
        out.append(h)
    return out


def velocity_error(b: LbmBatch, ref: LbmBatch):
    errs = []
    for k in range(b.ninst):
        u = b.moments(k)[1]
        ur = ref.moments(k)[1]
        errs.append(float(np.linalg.norm(u - ur) / max(np.linalg.norm(ur), 1e-300)))
    return np.array(errs)


def timed_steps(b: LbmBatch, steps, chunk=10):
    b.sync()
    t0 = time.perf_counter()
    k = 0
    while k < steps:
        n = min(chunk, steps - k)
        b.step(n)
        k += n
    b.sync()
    return time.perf_counter() - t0


def plain_rate(specs, mesh_list, steps, sample, tuned=False):
    """Unbatched FP32: single-instance solvers one after another.  Plain
    (tuned=False) steps eagerly with the default collide variant; tuned
    gives each solver what the batch has -- the autotuner's variants and
    hipGraph replay -- so that batching is compared alone."""
    upd, t = 0, 0.0
    for k in range(min(sample, len(specs))):
        s = specs[k]
        m = mesh_list[k]
        sol = LbmSolver(s["n0"], s["n0"], s["n0"], s["levels"], s["tau"], ["wall"] * 6,
                        [(0.0, 0.0, 0.0)] * 5 + [(s["lid"], 0.0, 0.0)], m["x"], m["y"], m["z"], m["level"],
                        m["state"], storage="f32")
        if tuned:
            sol.autotune(3)
            sol.set_graph(True)
        sol.step(2)
        t += timed_steps(sol, steps, chunk=10 if tuned else 1)
        upd += int(sol.stats().updates_per_step) * steps
        sol.close()
    return upd / t / 1e6, upd, t


def run(instances=256, steps=2000, regrid_every=0, plain_sample=256, plain_steps=200, seed=42, nmax=96,
        forms=("f32", "f16", "f16c")):
    check(lib().fl_init(0))
    specs = instance_specs(instances, seed, nmax=nmax)
    t0 = time.perf_counter()
    mesh_list = build_meshes(specs)
    log(f"{instances} meshes built in {time.perf_counter() - t0:.1f} s; "
        f"{sum(len(m['x']) for m in mesh_list)} leaves")
    batches = {}
    for st in ("f64",) + tuple(forms):
        t1 = time.perf_counter()
        b = LbmBatch([cavity_params(s, st) for s in specs], mesh_list)
        if st != "f64":
            chosen = b.autotune(6)
            log(f"{st}: autotuned {[variant_names()[v] for v in chosen]}")
        b.set_graph(True)
        batches[st] = b
        log(f"{st} batch built in {time.perf_counter() - t1:.1f} s, "
            f"{b.stats().device_bytes / 2**30:.2f} GiB")
    upd_step = int(batches["f64"].stats().updates_per_step)
    m0 = {st: b.totals_batch()[0] for st, b in batches.items()}
    times = {st: 0.0 for st in batches}
    done = 0
    regrids = 0
    upd_done = 0
    regrid_s = 0.0
    while done < steps:
        n = steps - done
        if regrid_every:
            n = min(n, regrid_every - done % regrid_every)
        for st, b in batches.items():
            times[st] += timed_steps(b, n)
        done += n
        upd_done += upd_step * n
        if regrid_every and done % regrid_every == 0 and done < steps:
            r0 = time.perf_counter()
            new = regrid_meshes(batches["f64"], specs)
            for b in batches.values():
                b.regrid_batch(new)
            regrid_s += time.perf_counter() - r0
            regrids += 1
            upd_step = int(batches["f64"].stats().updates_per_step)
        log(f"step {done}/{steps}: " + ", ".join(f"{st} {upd_done / max(t, 1e-9) / 1e6:.0f} MLUPS"
                                                 for st, t in times.items()))
    res = {}
    errs = {}
    for st in forms:
        e = velocity_error(batches[st], batches["f64"])
        errs[st] = e
        m = batches[st].totals_batch()[0]
        drift = np.abs(m - m0[st]) / m0[st]
        res[st] = dict(mlups=round(upd_done / times[st] / 1e6, 1), seconds=round(times[st], 3),
                       solves_per_s=round(instances / times[st], 3),
                       vel_l2_max=float(e.max()), vel_l2_median=float(np.median(e)),
                       mass_drift_max=float(drift.max()))
    for st in forms:
        if st == "f32":
            continue
        worst = np.argsort(-errs[st])[:5]
        res[f"{st}_worst"] = [dict(instance=int(k), **specs[k], vel_l2=float(errs[st][k])) for k in worst]
        res[f"{st}_frac_below_1e-3"] = float((errs[st] <= 1e-3).mean())
    m64 = batches["f64"].totals_batch()[0]
    res["f64"] = dict(mlups=round(upd_done / times["f64"] / 1e6, 1), seconds=round(times["f64"], 3),
                      mass_drift_max=float((np.abs(m64 - m0["f64"]) / m0["f64"]).max()))
    for b in batches.values():
        b.close()
    plain, pu, pt = plain_rate(specs, mesh_list, plain_steps, plain_sample)
    tuned, _, _ = plain_rate(specs, mesh_list, plain_steps, plain_sample, tuned=True)
    out = dict(workload=f"{instances} lid-driven cavities, base 16^3..{nmax}^3 log-uniform, lid U(0.02,0.10), "
                        f"tau U(0.55,0.9), 0-2 levels, {steps} steps" +
                        (f", regrid every {regrid_every}" if regrid_every else ""),
               updates_per_step=upd_step, regrids=regrids, regrid_seconds_all_forms=round(regrid_s, 2),
               results=res,
               plain_f32=dict(mlups=round(plain, 1), sample=f"first {min(plain_sample, instances)} instances x "
                                                            f"{plain_steps} steps, single-instance eager"),
               plain_f32_graph=dict(mlups=round(tuned, 1),
                                    sample=f"the same instances, single-instance, autotuned variants and hipGraph "
                                           f"replay"),
               over_plain={st: round(res[st]["mlups"] / plain, 2) for st in forms},
               over_plain_graph={st: round(res[st]["mlups"] / tuned, 2) for st in forms})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--regrid-every", type=int, default=0)
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--nmax", type=int, default=96)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--forms", default="f32,f16,f16c", help="storage forms measured against FP64")
    a = ap.parse_args()
    print(json.dumps(run(a.instances, a.steps, a.regrid_every, a.plain_sample, a.plain_steps, a.seed, a.nmax,
                         tuple(a.forms.split(",")))), flush=True)


if __name__ == "__main__":
    main()
