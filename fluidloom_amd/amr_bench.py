"""AMR integer micro-benchmark on the device: 64M seeded leaf cells at levels
0-10 in a 2^10 root domain, split Bernoulli(0.05), merge Bernoulli(0.03) on
complete sibling groups; the reference's rot/dir keys, Hilbert keys, 64-bit
key sort, a 64M-element exclusive scan, one regrid pass
(reference balance rule to fixpoint, split, merge, scan compaction) and the
hash face-neighbour rebuild.  Each stage is timed with HIP events; GB/s is
computed from the algorithmic bytes of that stage (for the sort: the digit
passes it actually ran).

The seeded cells overlap (they are not an octree), and the reference balance
rule converges on them in one pass, so cascade() times a balance that does
cascade: the D3Q19 rule (mode 1) to its fixpoint on a valid graded octree
(the 3-level sphere mesh, 76M leaves) after every fluid leaf of its finest
region is flagged for one more level.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import amr
from ._lib import DeviceArray, RegridConfig, RegridResult, check, lib


class _Timer:
    def __init__(self):
        self.a, self.b = C.c_void_p(), C.c_void_p()
        check(lib().fl_event_create(C.byref(self.a)))
        check(lib().fl_event_create(C.byref(self.b)))

    def __enter__(self):
        check(lib().fl_device_sync())
        check(lib().fl_event_record(self.a, None))
        return self

    def __exit__(self, *exc):
        check(lib().fl_event_record(self.b, None))
        ms = C.c_float(0)
        check(lib().fl_event_elapsed_ms(self.a, self.b, C.byref(ms)))
        self.ms = ms.value


def run(n=64 << 20, seed=42, p_split=0.05, p_merge=0.03, max_level=10, log=None):
    L = lib()
    cells = amr.GpuCells(n)
    cells.n = n
    cs = cells.struct()
    out = {"cells": n}
    t = _Timer()
    with t:
        check(L.fl_gen_cells(C.byref(cs), n, seed, max_level, None))
    tb = L.fl_regrid_temp_bytes(n)
    tmp = DeviceArray(tb, np.uint8)
    check(L.fl_gen_flags(C.byref(cs), n, seed, p_split, p_merge, tmp.p, tb, None))
    keys = DeviceArray(n, np.uint64)
    # rot/dir codec, the reference's merge / balance hash key at max_level 8
    # (20 B/cell: 12 coords in, 8 key out)
    with t:
        check(L.fl_rotdir_encode(cells.x.p, cells.y.p, cells.z.p, 8, keys.p, n, None))
    out["rotdir_ms"] = t.ms
    out["rotdir_gbs"] = 20 * n / t.ms / 1e6
    # keygen: key63 | level (21 B/cell: 12 coords + 1 level + 8 key)
    with t:
        check(L.fl_hilbert_key63(cells.x.p, cells.y.p, cells.z.p, cells.level.p, 1, keys.p, n, None))
    out["keygen_ms"] = t.ms
    out["keygen_gbs"] = 21 * n / t.ms / 1e6
    # 64-bit key sort with values = original index (skips constant digits)
    vals = DeviceArray(n, np.uint32)
    check(L.fl_iota_u32(vals.p, n, None))
    ka, va = DeviceArray(n, np.uint64), DeviceArray(n, np.uint32)
    stb = L.fl_sort_temp_bytes(n)
    st = DeviceArray(stb, np.uint8)
    with t:
        check(L.fl_sort_pairs_u64u32(keys.p, vals.p, ka.p, va.p, n, 0, 64, st.p, stb, None))
    out["sort_ms"] = t.ms
    out["sort_gkeys_s"] = n / t.ms / 1e6
    # the sort skips a one-byte digit that every key shares (found by its
    # upfront histogram): count the digits that vary over these keys
    kh = keys.to_host()
    var = int(np.bitwise_or.reduce(kh ^ kh[0])) if n else 0
    del kh
    passes = sum(1 for d in range(8) if (var >> (8 * d)) & 0xFF)
    out["sort_passes"] = passes
    # bytes moved: the upfront histogram reads the keys (8 B), each pass reads
    # and writes u64 key + u32 value (24 B)
    out["sort_gbs"] = (8 + passes * 24) * n / t.ms / 1e6
    del ka, va, st
    # exclusive scan of n u32 (4 B read + 4 B written per element)
    sin, sout, stot = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32), DeviceArray(1, np.uint32)
    check(L.fl_iota_u32(sin.p, n, None))
    sctb = int(L.fl_scan_temp_bytes(n))
    sct = DeviceArray(max(sctb, 1), np.uint8)
    with t:
        check(L.fl_scan_exclusive_u32(sin.p, sout.p, stot.p, n, sct.p, sctb, None))
    out["scan_ms"] = t.ms
    out["scan_gbs"] = 8 * n / t.ms / 1e6
    del sin, sout, stot, sct
    # regrid pass
    cap = n + n // 2
    new = amr.GpuCells(cap)
    cfg = RegridConfig()
    cfg.balance_mode, cfg.anchor_level, cfg.max_split_level, cfg.max_balance_iters = 0, max_level, max_level, 64
    cfg.root[0] = cfg.root[1] = cfg.root[2] = 1
    res = RegridResult()
    ns = new.struct()
    w0 = time.perf_counter()
    with t:
        check(L.fl_regrid(C.byref(cfg), C.byref(cs), n, C.byref(ns), cap, None, None, tmp.p, tb, C.byref(res),
                          None))
    out["regrid_ms"] = t.ms
    out["regrid_mcells_s"] = n / t.ms / 1e3
    out["regrid_wall_ms"] = (time.perf_counter() - w0) * 1e3
    out.update(cells_out=int(res.num_cells_out), split=int(res.num_split), merged_groups=int(res.num_groups),
               balance_iterations=int(res.balance_iterations), balance_converged=bool(res.balance_converged),
               balance_marked=int(res.balance_marked))
    new.n = int(res.num_cells_out)
    cells.free()
    del tmp
    # hash face-neighbour rebuild over the new cells
    m = new.n
    hcap = int(L.fl_hash_capacity(m))
    hk, tk, tv = DeviceArray(m, np.uint64), DeviceArray(hcap, np.uint64), DeviceArray(hcap, np.uint32)
    nbr = DeviceArray(6 * m, np.uint32)
    root = (C.c_int32 * 3)(1, 1, 1)
    nsn = new.struct()
    with t:
        check(L.fl_face_neighbors(C.byref(nsn), m, root, hk.p, tk.p, tv.p, hcap, nbr.p, None))
    out["neighbors_ms"] = t.ms
    out["neighbors_mcells_s"] = m / t.ms / 1e3
    # reference-parity hash build (fl_hash_build: the reference's low-bit slot
    # function, 32 probes) on uniformly random 63-bit keys, the distribution
    # that slot function is made for (24 B/key: key + value read, key + value
    # written)
    rk = np.random.default_rng(seed).integers(0, 1 << 63, n, dtype=np.uint64)
    rkeys = DeviceArray.from_host(rk)
    del rk
    hcap = int(L.fl_hash_capacity(n))
    tk, tv = DeviceArray(hcap, np.uint64), DeviceArray(hcap, np.uint32)
    with t:
        check(L.fl_hash_clear(tk.p, tv.p, hcap, None))
        check(L.fl_hash_build(tk.p, tv.p, hcap, rkeys.p, vals.p, n, None, None))
    out["hash_build_ms"] = t.ms
    out["hash_build_gkeys_s"] = n / t.ms / 1e6
    out["hash_build_gbs"] = 24 * n / t.ms / 1e6
    qout = DeviceArray(n, np.uint32)
    with t:
        check(L.fl_hash_query(tk.p, tv.p, hcap, rkeys.p, qout.p, n, None, None))
    out["hash_query_ms"] = t.ms
    out["hash_query_gkeys_s"] = n / t.ms / 1e6
    new.free()
    if log:
        log(f"amr micro-benchmark: {out}")
    return out


def cascade(cells: amr.GpuCells, center, radius, target_level, root, temp=None):
    """Flags every fluid leaf below target_level whose cell lies within
    `radius` of `center`, then runs the D3Q19 2:1 balance (mode 1) to its
    fixpoint on the device.  `cells` is modified (flags).  Returns the flag
    count and the balance's counters and time (HIP events around the whole
    fl_balance call, which reads its counters back after each iteration)."""
    flagged = amr.mark_ball(cells, center, radius, target_level)
    if temp is None:
        temp = amr.regrid_temp(cells.n)
    t = _Timer()
    w0 = time.perf_counter()
    with t:
        st = amr.balance(cells, balance_mode=1, max_split_level=target_level, root=root, temp=temp)
    wall = (time.perf_counter() - w0) * 1e3
    return {"cells": cells.n, "flagged": flagged, "iterations": st.balance_iterations,
            "violations": st.balance_violations, "marked": st.balance_marked, "converged": st.balance_converged,
            "ms": t.ms, "wall_ms": wall, "mcells_iter_s": cells.n * st.balance_iterations / t.ms / 1e3}


def sphere_cascade(scale=1, log=None):
    """cascade() on the 3-level sphere mesh (base 256x128x128 / scale, r = 16 /
    scale): the whole finest region (radius 2r) is flagged for a fourth
    level, so the balance must refine a ring of each coarser level in turn."""
    from . import meshes
    nx0, ny0, nz0, r = 256 // scale, 128 // scale, 128 // scale, 16.0 / scale
    ctr = (64.0 / scale,) * 3
    m = meshes.sphere(nx0, ny0, nz0, ctr, r, 2 * r, 3, log=log)
    cells = amr.GpuCells.from_host(**{k: m[k] for k in ("x", "y", "z", "level", "state")})
    out = cascade(cells, ctr, 2 * r, 4, (nx0, ny0, nz0))
    cells.free()
    return out


def main():
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=64 << 20)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--cascade-scale", type=int, default=1, help="sphere scale of the cascade stage (0: skip)")
    a = ap.parse_args()
    check(lib().fl_init(0))
    # one plain copy of 1 GiB: the reference point that scales counter runs
    # (FETCH_SIZE / WRITE_SIZE) of this command to bytes
    src, dst = DeviceArray(1 << 30, np.uint8), DeviceArray(1 << 30, np.uint8)
    check(lib().fl_copy_kernel(dst.ptr, src.ptr, 1 << 30, None))
    check(lib().fl_device_sync())
    del src, dst
    if a.cascade_scale:
        casc = sphere_cascade(a.cascade_scale)
        print(f"[amr_bench] cascade {casc}", file=sys.stderr, flush=True)
    runs = []
    for _ in range(a.repeat):
        runs.append(run(a.cells, a.seed))
        print(f"[amr_bench] {runs[-1]}", file=sys.stderr, flush=True)
    best = {k: (min(r[k] for r in runs) if k.endswith("_ms") else runs[-1][k]) for k in runs[-1]}
    for k in list(best):
        if k.endswith("_gbs") or k.endswith("_s"):
            best[k] = max(r[k] for r in runs)
    if a.cascade_scale:
        best["cascade"] = casc
    print(json.dumps(best), flush=True)


if __name__ == "__main__":
    main()
