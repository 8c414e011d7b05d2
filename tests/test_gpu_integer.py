"""GPU parity of the integer hot path (keys, sort, scan, hash, neighbours,
adaptation) against the oracle: bit-exact, through the C ABI."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "hilbert_golden.npz"


def dev(fl, a):
    return fl.DeviceArray.from_host(np.ascontiguousarray(a))


def test_hilbert_encode_decode_golden(fl):
    g = np.load(GOLD)
    n = len(g["x"])
    x, y, z, lv = dev(fl, g["x"]), dev(fl, g["y"]), dev(fl, g["z"]), dev(fl, g["level"])
    out = fl.DeviceArray(n, np.uint64)
    fl.check(fl.lib().fl_hilbert_encode(x.p, y.p, z.p, lv.p, 0, out.p, n, None))
    assert np.array_equal(out.to_host(), g["encode"])
    k = dev(fl, g["encode"])
    ox, oy, oz = (fl.DeviceArray(n, np.int32) for _ in range(3))
    fl.check(fl.lib().fl_hilbert_decode(k.p, lv.p, 0, ox.p, oy.p, oz.p, n, None))
    assert np.array_equal(ox.to_host(), g["dec_x"]) and np.array_equal(oy.to_host(), g["dec_y"])
    assert np.array_equal(oz.to_host(), g["dec_z"])
    for ml, name in [(8, "rotdir8"), (10, "rotdir10")]:
        fl.check(fl.lib().fl_rotdir_encode(x.p, y.p, z.p, ml, out.p, n, None))
        assert np.array_equal(out.to_host(), g[name])


def test_hilbert_scalar_level_and_key63(fl, oracle):
    rng = np.random.default_rng(1)
    for n in [1, 63, 1000, 100003]:
        for level in [0, 1, 7, 10, 21]:
            x, y, z = (rng.integers(0, 1 << max(level, 1), n).astype(np.int32) for _ in range(3))
            out = fl.DeviceArray(n, np.uint64)
            dx, dy, dz = dev(fl, x), dev(fl, y), dev(fl, z)
            fl.check(fl.lib().fl_hilbert_encode(dx.p, dy.p, dz.p, None, level, out.p, n, None))
            assert np.array_equal(out.to_host(), oracle.hilbert_encode(x, y, z, level))
    n = 50000
    lv = rng.integers(0, 20, n).astype(np.uint8)
    x, y, z = (np.array([rng.integers(0, 1 << int(l)) for l in lv], np.int32) for _ in range(3))
    out = fl.DeviceArray(n, np.uint64)
    dx, dy, dz, dl = dev(fl, x), dev(fl, y), dev(fl, z), dev(fl, lv)
    fl.check(fl.lib().fl_hilbert_key63(dx.p, dy.p, dz.p, dl.p, 0, out.p, n, None))
    k63 = oracle.key63(x, y, z, lv)
    assert np.array_equal(out.to_host(), k63)
    fl.check(fl.lib().fl_hilbert_key63(dx.p, dy.p, dz.p, dl.p, 1, out.p, n, None))
    assert np.array_equal(out.to_host(), k63 | lv.astype(np.uint64))


def gpu_sort(fl, keys, vals, b, e):
    n = len(keys)
    k, v = dev(fl, keys), dev(fl, vals)
    ka, va = fl.DeviceArray(max(n, 1), np.uint64), fl.DeviceArray(max(n, 1), np.uint32)
    tb = fl.lib().fl_sort_temp_bytes(n)
    t = fl.DeviceArray(max(tb, 1), np.uint8)
    fl.check(fl.lib().fl_sort_pairs_u64u32(k.p, v.p, ka.p, va.p, n, b, e, t.p, tb, None))
    return k.to_host(), v.to_host()


@pytest.mark.parametrize("n", [0, 1, 2, 1000, 4096, 4097, 100003, (1 << 20) + 3])
def test_sort_stable_matches_oracle(fl, oracle, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    if n > 10:
        m = len(keys[1::5])
        keys[0:5 * m:5] = keys[1::5]
        keys[3::17] &= np.uint64(0xFF00FF)
    vals = np.arange(n, dtype=np.uint32)
    for b, e in [(0, 64), (8, 40), (33, 63)]:
        gk, gv = gpu_sort(fl, keys, vals, b, e)
        ok, ov = oracle.sort_pairs(keys, vals, b, e)
        assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


def test_sort_64m_key63_properties(fl):
    """Full-size AMR micro-benchmark keys: 64M packed key63|level keys."""
    n = 64 << 20
    rng = np.random.default_rng(42)
    keys = (rng.integers(0, 1 << 30, n, dtype=np.uint64) << np.uint64(33)) | rng.integers(0, 11, n).astype(np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = gpu_sort(fl, keys, vals, 0, 64)
    assert np.all(gk[1:] >= gk[:-1])
    assert np.array_equal(keys[gv], gk)
    same = gk[1:] == gk[:-1]
    assert np.all(gv[1:][same] > gv[:-1][same])  # stability
    assert np.array_equal(np.bincount(gv, minlength=n), np.ones(n))


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 65537, 10_000_019])
def test_scan_matches_oracle(fl, oracle, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 7, n).astype(np.uint32)
    d = dev(fl, a) if n else fl.DeviceArray(1, np.uint32)
    out = fl.DeviceArray(max(n, 1), np.uint32)
    tot = fl.DeviceArray(1, np.uint64)
    tb = fl.lib().fl_scan_temp_bytes(n)
    t = fl.DeviceArray(max(tb, 8), np.uint8)
    fl.check(fl.lib().fl_scan_exclusive_u32(d.p, out.p, tot.p, n, t.p, tb, None))
    ref, rt = oracle.scan_exclusive(a)
    assert int(tot.to_host()[0]) == rt
    if n:
        assert np.array_equal(out.to_host()[:n], ref)


@pytest.mark.parametrize("off_in,off_out", [(1, 0), (0, 3), (2, 2)])
def test_scan_unaligned_pointers(fl, oracle, off_in, off_out):
    # the C ABI takes any element pointer: runs that are not 16 B aligned
    # take the element-wise path and must give the same result
    n = 3 * 4096 + 77
    a = np.random.default_rng(off_in + 7 * off_out).integers(0, 9, n).astype(np.uint32)
    d = dev(fl, np.concatenate([np.zeros(off_in, np.uint32), a]))
    out = fl.DeviceArray(n + off_out, np.uint32)
    tot = fl.DeviceArray(1, np.uint64)
    tb = fl.lib().fl_scan_temp_bytes(n)
    t = fl.DeviceArray(max(tb, 8), np.uint8)
    fl.check(fl.lib().fl_scan_exclusive_u32(d.p + 4 * off_in, out.p + 4 * off_out, tot.p, n, t.p, tb, None))
    ref, rt = oracle.scan_exclusive(a)
    assert int(tot.to_host()[0]) == rt
    assert np.array_equal(out.to_host()[off_out:off_out + n], ref)


@pytest.mark.parametrize("n", [1, 1000, 300001])
def test_hash_matches_oracle(fl, oracle, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 50, n, dtype=np.uint64)
    if n > 10:
        keys[n // 2:n // 2 + n // 10] = keys[: n // 10]
        keys[7] = np.uint64(0xFFFFFFFFFFFFFFFF)
    vals = rng.permutation(n).astype(np.uint32)
    cap = int(fl.lib().fl_hash_capacity(n))
    assert cap == oracle.hash_capacity(n)
    tk, tv = fl.DeviceArray(cap, np.uint64), fl.DeviceArray(cap, np.uint32)
    pc = fl.DeviceArray(n, np.uint32)
    fl.check(fl.lib().fl_hash_clear(tk.p, tv.p, cap, None))
    dk, dv = dev(fl, keys), dev(fl, vals)
    fl.check(fl.lib().fl_hash_build(tk.p, tv.p, cap, dk.p, dv.p, n, pc.p, None))
    q = np.concatenate([keys, rng.integers(1 << 51, 1 << 52, 5000, dtype=np.uint64)])
    out = fl.DeviceArray(len(q), np.uint32)
    dq = dev(fl, q)
    fl.check(fl.lib().fl_hash_query(tk.p, tv.p, cap, dq.p, out.p, len(q), None, None))
    otk, otv, _ = oracle.hash_build(keys, vals, cap)
    assert np.array_equal(out.to_host(), oracle.hash_query(otk, otv, cap, q))
    assert pc.to_host().max() < 32


def _gpu_cells(cs):
    from fluidloom_amd import amr
    return amr.GpuCells.from_host(x=cs.x, y=cs.y, z=cs.z, level=cs.level, state=cs.state, flag=cs.flag,
                                  material=cs.material)


def test_face_neighbors_match_oracle(fl, oracle):
    from fluidloom_amd import amr
    from tests.amr_util import random_octree
    for seed, root in [(1, (2, 2, 2)), (5, (3, 2, 1))]:
        c = random_octree(seed=seed, root=root, depth=4)
        g = _gpu_cells(c)
        nbr = amr.face_neighbors(g, root).to_host()[: 6 * c.n]
        assert np.array_equal(nbr, oracle.face_neighbors(c, root))
    x, y, z, lv = oracle.gen_cells(42, 200000, 10)
    c = oracle.CellSet(len(x), x=x, y=y, z=z, level=lv)
    g = _gpu_cells(c)
    assert np.array_equal(amr.face_neighbors(g, (1, 1, 1)).to_host()[: 6 * c.n], oracle.face_neighbors(c, (1, 1, 1)))


def test_face_neighbors_dense_blocks_at_high_load(fl, oracle):
    """The cell map's worst case: every cell in a full 2x2x2 block (a
    uniform 84^3 region at level 7, a refined 32^3 patch inside it, and a
    shuffled copy of the same cells), so whole blocks collide as units, at
    the public capacity's load (the uniform region: 593K cells in 2^20
    slots, 0.57; with the patch: 822K cells in 2^21 slots, 0.39).  Face neighbours must equal the
    oracle's, whose own map probes per cell."""
    from fluidloom_amd import amr
    g84 = np.stack(np.meshgrid(*(np.arange(84),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
    c = oracle.CellSet(len(g84), x=g84[:, 0], y=g84[:, 1], z=g84[:, 2], level=np.full(len(g84), 7, np.uint8))
    assert 0.55 < c.n / float(fl.lib().fl_hash_capacity(c.n)) < 0.6
    assert np.array_equal(amr.face_neighbors(_gpu_cells(c), (1, 1, 1)).to_host()[: 6 * c.n],
                          oracle.face_neighbors(c, (1, 1, 1)))
    # replace the level-7 cells of [16, 48)^3 by their level-8 children
    inside = np.all((g84 >= 16) & (g84 < 48), axis=1)
    kids = (2 * g84[inside][:, None, :] + np.array([[(q >> a) & 1 for a in range(3)] for q in range(8)],
                                                   np.int32)[None]).reshape(-1, 3)
    xyz = np.concatenate([g84[~inside], kids])
    lv = np.concatenate([np.full((~inside).sum(), 7, np.uint8), np.full(len(kids), 8, np.uint8)])
    perm = np.random.default_rng(3).permutation(len(xyz))
    for order in (np.arange(len(xyz)), perm):
        c = oracle.CellSet(len(xyz), x=xyz[order, 0], y=xyz[order, 1], z=xyz[order, 2], level=lv[order])
        assert np.array_equal(amr.face_neighbors(_gpu_cells(c), (1, 1, 1)).to_host()[: 6 * c.n],
                              oracle.face_neighbors(c, (1, 1, 1)))


def _same_cells(gh, oc):
    for nm in ("x", "y", "z", "level", "state", "material"):
        assert np.array_equal(gh[nm], getattr(oc, nm)), nm


def test_regrid_reference_mode_matches_oracle(fl, oracle):
    """Micro-benchmark semantics at reduced size: random cells, levels 0-10,
    split Bernoulli(0.05), merge -1 on complete sibling groups (0.03)."""
    from fluidloom_amd import amr
    from tests.c3_util import c3_cells
    c = c3_cells(oracle, 300000, seed=42)
    g = _gpu_cells(c)
    out, st = amr.regrid(g, balance_mode=0, anchor_level=10, max_split_level=10, max_iters=64)
    ref, res, gid, lead = oracle.regrid(c, 0, 10, 10, 64)
    _same_cells(out.to_host(), ref)
    assert st.num_groups == len(lead) and st.balance_iterations == res[0]
    assert st.balance_converged == bool(res[1]) and st.balance_violations == res[2] and st.balance_marked == res[3]


def test_balance_reference_mode_preflagged_on_gpu(fl, oracle):
    """The reference's balance loop starts from the stored levels and counts
    split flags set before the call only after its first marking pass
    (BalanceEnforcer.cpp:165-167, :248-254).  The two hand-derived cases of
    the oracle tests, then random complete octrees with 15 % pre-set flags,
    whose balance cascades over several passes."""
    from fluidloom_amd import amr
    from tests.amr_util import random_octree
    cases = [(dict(x=[0, 2], y=[0, 0], level=[1, 2], flag=[0, 1]), (1, 1, 0, 0), [0, 1]),
             (dict(x=[0, 2, 0, 4], y=[0, 0, 1, 4], level=[1, 2, 1, 3], flag=[0, 1, 0, 0]), (3, 1, 2, 2),
              [1, 1, 1, 0])]
    for geo, exp, flags in cases:
        n = len(geo["x"])
        c = oracle.CellSet(n, **geo)
        g = _gpu_cells(c)
        out, st = amr.regrid(g, balance_mode=0, anchor_level=8, max_split_level=8, max_iters=10)
        assert (st.balance_iterations, int(st.balance_converged), st.balance_violations, st.balance_marked) == exp
        assert st.num_split == sum(flags)
        ref, res, _, _ = oracle.regrid(c, 0, 8, 8, 10)
        assert res == exp
        _same_cells(out.to_host(), ref)
    multi = 0
    for seed in (5, 6, 7, 21, 22):
        c = random_octree(seed=seed, root=(2, 2, 2), depth=3, p_split=0.5)
        c.level[:] += 1
        rng = np.random.default_rng(seed)
        c.flag[:] = np.where(rng.random(c.n) < 0.15, 1, 0)
        g = _gpu_cells(c)
        out, st = amr.regrid(g, balance_mode=0, anchor_level=6, max_split_level=6, max_iters=64)
        ref, res, _, _ = oracle.regrid(c, 0, 6, 6, 64)
        _same_cells(out.to_host(), ref)
        assert (st.balance_iterations, int(st.balance_converged), st.balance_violations, st.balance_marked) == res
        multi += res[0] > 1
    assert multi >= 2


def test_regrid_lbm_mode_matches_oracle(fl, oracle):
    from fluidloom_amd import amr
    from tests.amr_util import random_octree
    c = random_octree(seed=11, root=(2, 2, 2), depth=4, p_split=0.35)
    rng = np.random.default_rng(3)
    c.flag[:] = np.where(rng.random(c.n) < 0.1, 1, np.where(rng.random(c.n) < 0.5, -1, 0))
    g = _gpu_cells(c)
    out, st = amr.regrid(g, balance_mode=1, max_split_level=6, root=(2, 2, 2), merge_balance_check=1)
    ref, res, gid, lead = oracle.regrid(c, 1, 19, 6, 64, (2, 2, 2), merge_balance_check=1)
    _same_cells(out.to_host(), ref)
    assert st.num_groups == len(lead)


def test_balance_fixture_on_gpu(fl):
    from fluidloom_amd import amr
    g = amr.GpuCells.from_host(x=np.array([0, 4]), y=np.zeros(2), z=np.zeros(2), level=np.array([1, 3]))
    out, st = amr.regrid(g, balance_mode=0, anchor_level=8, max_split_level=8, max_iters=10)
    assert st.balance_converged and st.balance_violations > 0 and st.balance_marked > 0 and st.num_split == 1


def test_synthetic_generators_match_oracle(fl, oracle):
    from fluidloom_amd import amr
    from tests.c3_util import c3_cells
    n = 300000
    g = amr.GpuCells(n)
    g.n = n
    cs = g.struct()
    fl.check(fl.lib().fl_gen_cells(C.byref(cs), n, 42, 10, None))
    tb = fl.lib().fl_regrid_temp_bytes(n)
    t = fl.DeviceArray(tb, np.uint8)
    fl.check(fl.lib().fl_gen_flags(C.byref(cs), n, 42, 0.05, 0.03, t.p, tb, None))
    h = g.to_host()
    ref = c3_cells(oracle, n, seed=42)
    for nm in ("x", "y", "z", "level", "flag"):
        assert np.array_equal(h[nm], getattr(ref, nm)), nm


def test_amr_microbenchmark_full_size_properties(fl):
    """64M cells: the regrid output is a permutation-free function of the
    input counts (survivors + 8 x split + groups) and every balance pass
    converged."""
    from fluidloom_amd import amr_bench
    r = amr_bench.run()
    assert r["balance_converged"]
    assert r["cells_out"] <= r["cells"] + 7 * r["split"]
    assert r["merged_groups"] > 0 and r["split"] > 0


def test_hash_empty_build_and_queries(fl, oracle):
    # an empty table answers INVALID for every key, including the EMPTY sentinel
    cap = int(fl.lib().fl_hash_capacity(0))
    assert cap == oracle.hash_capacity(0)
    tk, tv = fl.DeviceArray(cap, np.uint64), fl.DeviceArray(cap, np.uint32)
    fl.check(fl.lib().fl_hash_clear(tk.p, tv.p, cap, None))
    fl.check(fl.lib().fl_hash_build(tk.p, tv.p, cap, None, None, 0, None, None))
    q = np.array([0, 1, 12345, 0xFFFFFFFFFFFFFFFF], np.uint64)
    dq = dev(fl, q)
    out = fl.DeviceArray(len(q), np.uint32)
    fl.check(fl.lib().fl_hash_query(tk.p, tv.p, cap, dq.p, out.p, len(q), None, None))
    assert np.all(out.to_host() == 0xFFFFFFFF)


def test_regrid_empty_and_single_cell(fl, oracle):
    from fluidloom_amd import amr
    g = amr.GpuCells(0)
    out, st = amr.regrid(g, balance_mode=1, max_split_level=4, root=(1, 1, 1))
    assert out.n == 0 and st.num_cells_out == 0 and st.balance_converged
    # one root cell flagged for refinement: eight children in child order
    one = amr.GpuCells.from_host(x=np.array([0]), y=np.array([0]), z=np.array([0]), level=np.array([0]),
                                 flag=np.array([1]))
    out, st = amr.regrid(one, balance_mode=1, max_split_level=4, root=(1, 1, 1))
    h = out.to_host()
    assert st.num_split == 1 and out.n == 8
    q = np.arange(8)
    assert np.array_equal(h["x"], q & 1) and np.array_equal(h["y"], (q >> 1) & 1)
    assert np.array_equal(h["z"], (q >> 2) & 1) and np.all(h["level"] == 1)
    # the eight children flagged -1 merge back into the root cell
    back = amr.GpuCells.from_host(x=h["x"], y=h["y"], z=h["z"], level=h["level"], flag=-np.ones(8, np.int32))
    out2, st2 = amr.regrid(back, balance_mode=1, max_split_level=4, root=(1, 1, 1))
    h2 = out2.to_host()
    assert st2.num_groups == 1 and out2.n == 1 and h2["level"][0] == 0 and h2["x"][0] == 0
