"""Serving with continuous batching.

Requests (cavity and sphere solves of different sizes) arrive as a Poisson
process.  At every scheduling tick the server

  1. admits the requests that have arrived, in arrival order, while their
     device footprint (fl_lbm_plan_bytes of their own layout) fits the
     memory budget (free HBM at start-up, capped at 288 GB, times a safety
     fraction); the admitted ones form a new cohort -- one LbmBatch whose
     layout holds all of them, with its own stream and hipGraph;
  2. enqueues K steps on every cohort (each on its own stream, so the
     cohorts' work overlaps on the device);
  3. checks each live request for steady state or its step limit; a
     finished request is retired (fl_lbm_set_live 0: its bricks are skipped
     from then on) and its time-to-solution recorded.  Steady state: over
     every fluid leaf, the largest change of the window-mean velocity between
     consecutive windows of K steps, per step, is below the tolerance
     (|du| < 1e-6).  The window mean averages velocity samples taken every
     `sample_every` steps: 16-bit storage rounds stochastically, and a
     single-step snapshot of its velocity carries that noise (on the refined
     64x32x32 sphere at Re 100 the snapshot test's floor is 1.5-2e-6 per step
     with K = 100, so FP16 requests would never retire; window means over
     K = 400 with a sample every 10 steps give 4.8e-7 for FP16, 4e-7 for
     FP16C and 3.8e-7 for FP32, whose value there is the flow's own slow
     drift, and all three retire at step 2800);
  4. frees a cohort once none of its requests is live.

A request's trajectory does not depend on what it is batched with (bricks of
different instances never interact, reductions are per instance), so its
result and its step count are the same alone and in the mix.

The plain server is the comparison: the same trace served one request at a
time (one cohort of one request, FP32 storage, eager steps), i.e. what a
single-instance solver loop does.  Requests still unserved when a run stops
at its wall limit enter the latency percentiles with the time they had
waited so far (a lower bound), so an overloaded server is not flattered by
counting only what it finished.

    python -m fluidloom_amd.serving --rate 20 --duration 600
    python -m fluidloom_amd.serving --rate 4 --duration 60 --scale 2 --mode both
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import meshes
from ._lib import check, lib
from .solver import LbmBatch, make_params, plan_bytes

HBM_CAP = 288 * 10 ** 9


def log(msg):
    print(f"[serving] {msg}", file=sys.stderr, flush=True)


@dataclass
class Request:
    rid: int
    arrival: float
    kind: str
    spec: dict
    params: object = None
    mesh: dict = None
    footprint: int = 0
    admitted: float = -1.0
    finished: float = -1.0
    steps: int = 0
    updates_per_step: int = 0
    converged: bool = False
    result: dict = field(default_factory=dict)


def request_specs(rate, duration, seed=42, scale=1, levels_max=3):
    """Poisson arrivals (rate per second over `duration` seconds): 70 %
    cavities, base n0 log-uniform in 16..128; 30 % spheres, base from
    128x64x64 to 256x128x128, radius U(8, 16), Re U(50, 200), inlet 0.05.
    1..levels_max grid levels.  Sizes and radii are divided by `scale`
    (bricks need multiples of 4).  Generator: numpy MT19937(seed)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    t, rid, out = 0.0, 0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        nlev = int(rng.integers(1, levels_max + 1))
        if rng.uniform() < 0.7:
            n0 = float(np.exp(rng.uniform(np.log(16), np.log(128)))) / scale
            n0 = int(max(8, 4 * round(n0 / 4)))
            spec = dict(n0=n0, lid=float(rng.uniform(0.02, 0.10)), tau=float(rng.uniform(0.55, 0.9)),
                        max_level=nlev - 1)
            out.append(Request(rid, t, "cavity", spec))
        else:
            f = float(rng.uniform(1.0, 2.0))
            ny = int(max(8, 4 * round(64 * f / scale / 4)))
            r = float(rng.uniform(8.0, 16.0)) / scale
            re = float(rng.uniform(50.0, 200.0))
            spec = dict(nx=2 * ny, ny=ny, nz=ny, r=r, re=re, u_in=0.05, max_level=nlev - 1)
            out.append(Request(rid, t, "sphere", spec))
        rid += 1
    return out


def prepare(req: Request, storage):
    """Mesh (device regrid), lattice parameters and device footprint."""
    s = req.spec
    if req.kind == "cavity":
        n0 = s["n0"]
        req.mesh = meshes.cavity(n0, s["max_level"])
        bcu = [(0.0, 0.0, 0.0)] * 5 + [(s["lid"], 0.0, 0.0)]
        req.params = make_params(n0, n0, n0, s["max_level"], s["tau"], ["wall"] * 6, bcu, storage)
    else:
        nx, ny, nz, r = s["nx"], s["ny"], s["nz"], s["r"]
        ctr = (nx / 4.0, ny / 2.0, nz / 2.0)
        req.mesh = meshes.sphere(nx, ny, nz, ctr, r, 2 * r, s["max_level"])
        tau0 = 3 * s["u_in"] * 2 * r / s["re"] + 0.5
        bc = ["wall", "outflow", "wall", "wall", "wall", "wall"]
        bcu = [(s["u_in"], 0.0, 0.0)] + [(0.0, 0.0, 0.0)] * 5
        req.params = make_params(nx, ny, nz, s["max_level"], tau0, bc, bcu, storage, u_init=(s["u_in"], 0.0, 0.0))
    req.footprint = plan_bytes([req.params], [req.mesh])


class Cohort:
    def __init__(self, reqs, graph=True):
        self.reqs = reqs
        self.batch = LbmBatch([r.params for r in reqs], [r.mesh for r in reqs], graph=graph)
        self.live = [True] * len(reqs)
        self.bytes = int(self.batch.stats().device_bytes)
        for k, r in enumerate(reqs):
            r.updates_per_step = self.batch.inst_updates(k)[0]
        self.batch.velocity_change()  # snapshot at step 0

    def any_live(self):
        return any(self.live)


class Server:
    def __init__(self, storage="f16", check_every=400, tol=1e-6, max_steps=20000, mem_fraction=0.9,
                 max_cohort=16, graph=True, keep_results=False, max_cohorts=None, sample_every=10):
        if check_every % sample_every:
            raise ValueError("check_every must be a multiple of sample_every")
        self.storage, self.K, self.tol, self.max_steps = storage, check_every, tol, max_steps
        self.sample_every = sample_every
        self.max_cohort, self.graph, self.keep_results = max_cohort, graph, keep_results
        self.max_cohorts = max_cohorts  # None: as many concurrent cohorts as memory admits
        free, total = C_mem_info()
        self.budget = int(min(free, HBM_CAP) * mem_fraction)
        self.used = 0
        self.cohorts: list[Cohort] = []
        self.done: list[Request] = []
        self.t_log = time.perf_counter()
        self.prep_s = 0.0  # host + device time spent preparing meshes and layouts

    def _admit(self, waiting, now):
        if self.max_cohorts is not None and len(self.cohorts) >= self.max_cohorts:
            return 0
        batch = []
        t0 = time.perf_counter()
        while waiting and waiting[0].arrival <= now and len(batch) < self.max_cohort:
            r = waiting[0]
            if r.mesh is None:
                prepare(r, self.storage)
            if r.footprint > self.budget:
                raise RuntimeError(f"request {r.rid} needs {r.footprint} B, more than the whole budget")
            if self.used + sum(q.footprint for q in batch) + r.footprint > self.budget:
                break
            batch.append(waiting.pop(0))
        if batch:
            c = Cohort(batch, self.graph)
            t = time.perf_counter()
            for r in batch:
                r.admitted = t
            self.used += c.bytes
            self.cohorts.append(c)
            self.prep_s += t - t0
            log(f"admitted a cohort of {len(batch)} ({sum(r.kind == 'sphere' for r in batch)} spheres, "
                f"{c.bytes / 2**30:.2f} GiB) in {t - t0:.1f} s; {len(waiting)} waiting")
            self.t_log = t
        return len(batch)

    def _retire(self, c: Cohort, k, now, converged):
        r = c.reqs[k]
        r.finished = now
        r.converged = converged
        if self.keep_results:
            r.result["f"] = c.batch.get_f(k)
            r.result["fluid"] = tuple(np.asarray(a).copy() for a in c.batch.fluid[k])
        c.batch.set_live(k, False)
        c.live[k] = False
        self.done.append(r)

    def run(self, requests, wall_limit=None, t0=None):
        """Serve `requests` (arrival times in seconds from start).  Returns
        the list of finished requests; arrival/admit/finish are wall-clock
        seconds from the start."""
        waiting = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter() if t0 is None else t0
        for r in waiting:
            r.arrival += t0
        ticks = 0
        while waiting or self.cohorts:
            now = time.perf_counter()
            if wall_limit is not None and now - t0 > wall_limit:
                break
            self._admit(waiting, now)
            if not self.cohorts:
                if waiting:
                    time.sleep(max(0.0, min(waiting[0].arrival - time.perf_counter(), 0.05)))
                continue
            # K steps per cohort, each on its own stream, with a velocity
            # sample every sample_every steps (stream-ordered, no sync)
            for _ in range(self.K // self.sample_every):
                for c in self.cohorts:
                    c.batch.step(self.sample_every)
                    c.batch.velocity_sample()
            for c in self.cohorts:
                du = c.batch.velocity_change()
                now = time.perf_counter()
                for k, r in enumerate(c.reqs):
                    if not c.live[k]:
                        continue
                    r.steps += self.K
                    if du[k] / self.K < self.tol:
                        self._retire(c, k, now, True)
                    elif r.steps >= self.max_steps:
                        self._retire(c, k, now, False)
            keep = []
            for c in self.cohorts:
                if c.any_live():
                    keep.append(c)
                else:
                    self.used -= c.bytes
                    c.batch.close()
            self.cohorts = keep
            ticks += 1
            if ticks % 20 == 0 or time.perf_counter() - self.t_log > 30:
                self.t_log = time.perf_counter()
                log(f"t={time.perf_counter() - t0:.1f}s done={len(self.done)} waiting={len(waiting)} "
                    f"cohorts={len(self.cohorts)} used={self.used / 2**30:.2f} GiB")
        t_end = time.perf_counter()
        for r in self.done:
            r.arrival -= t0
            r.admitted -= t0
            r.finished -= t0
        # arrived but unfinished: how long each has waited so far
        pending = [r for r in waiting if r.arrival <= t_end]
        pending += [r for c in self.cohorts for k, r in enumerate(c.reqs) if c.live[k]]
        self.pending_wait = [t_end - r.arrival for r in pending]
        self.unserved = len(waiting) + sum(sum(c.live) for c in self.cohorts)
        return self.done


def C_mem_info():
    import ctypes as C
    free, total = C.c_size_t(0), C.c_size_t(0)
    check(lib().fl_mem_info(C.byref(free), C.byref(total)))
    return int(free.value), int(total.value)


def summary(done, wall, unserved, pending_wait=()):
    """Latency percentiles over every request that arrived: finished ones
    with their time to solution, unfinished ones with their wait so far
    (counted in `censored`; the percentiles are then lower bounds)."""
    lat = [r.finished - r.arrival for r in done] + list(pending_wait)
    lat = np.array(lat) if lat else np.zeros(1)
    upd = sum(r.updates_per_step * r.steps for r in done)
    return dict(served=len(done), unserved=unserved, censored=len(pending_wait),
                converged=sum(r.converged for r in done),
                p50_s=round(float(np.percentile(lat, 50)), 3), p99_s=round(float(np.percentile(lat, 99)), 3),
                mean_queue_s=round(float(np.mean([r.admitted - r.arrival for r in done])) if done else 0.0, 3),
                mlups=round(upd / wall / 1e6, 1), solves_per_s=round(len(done) / wall, 3), wall_s=round(wall, 2))


def plain_server(check_every=400, tol=1e-6, max_steps=20000, sample_every=10, graph=False):
    """One request at a time, FP32: the single-instance baseline (eager
    steps unless `graph`)."""
    return Server("f32", check_every, tol, max_steps, max_cohort=1, graph=graph, max_cohorts=1,
                  sample_every=sample_every)


def serve(srv, reqs, wall_limit):
    t0 = time.perf_counter()
    done = srv.run(reqs, wall_limit, t0)
    wall = time.perf_counter() - t0
    return dict(storage=srv.storage, budget_gib=round(srv.budget / 2**30, 1), prepare_s=round(srv.prep_s, 1),
                **summary(done, wall, srv.unserved, srv.pending_wait))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=600.0)
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--max-steps", type=int, default=20000)
    ap.add_argument("--check-every", type=int, default=400)
    ap.add_argument("--sample-every", type=int, default=10)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--storage", default="f16")
    ap.add_argument("--wall-limit", type=float, default=None)
    ap.add_argument("--mode", default="batched", choices=["batched", "plain", "both"])
    ap.add_argument("--plain-wall-limit", type=float, default=None,
                    help="wall limit of the plain run (default: twice the trace duration)")
    ap.add_argument("--seed", type=int, default=42)
    a = ap.parse_args()
    check(lib().fl_init(0))
    out = dict(workload=f"Poisson {a.rate:g} req/s for {a.duration:g} s, 70% cavities / 30% spheres, "
                        f"sizes / {a.scale}, steady state |du|/step < {a.tol:g} or {a.max_steps} steps")
    if a.mode in ("batched", "both"):
        reqs = request_specs(a.rate, a.duration, a.seed, a.scale)
        log(f"continuous batching: {len(reqs)} requests over {a.duration:g} s")
        out["requests"] = len(reqs)
        out["batched"] = serve(Server(a.storage, a.check_every, a.tol, a.max_steps, sample_every=a.sample_every), reqs,
                               a.wall_limit)
    if a.mode in ("plain", "both"):
        reqs = request_specs(a.rate, a.duration, a.seed, a.scale)
        limit = a.plain_wall_limit if a.plain_wall_limit is not None else 2 * a.duration
        log(f"plain (one request at a time, FP32, eager): {len(reqs)} requests, wall limit {limit:g} s")
        out["requests"] = len(reqs)
        out["plain"] = serve(plain_server(a.check_every, a.tol, a.max_steps, a.sample_every), reqs, limit)
    if a.mode == "both":
        b, p = out["batched"], out["plain"]
        out["p99_plain_over_batched"] = round(p["p99_s"] / max(b["p99_s"], 1e-9), 2)
        out["p50_plain_over_batched"] = round(p["p50_s"] / max(b["p50_s"], 1e-9), 2)
        out["solves_per_s_batched_over_plain"] = round(b["solves_per_s"] / max(p["solves_per_s"], 1e-9), 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
