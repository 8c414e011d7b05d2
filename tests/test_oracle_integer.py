"""CPU: the oracle's integer restatements pinned against the reference's
golden vectors, its unit-test fixtures and independent implementations."""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "hilbert_golden.npz"


def test_hilbert_golden_vectors(oracle):
    g = np.load(GOLD)
    enc = oracle.hilbert_encode(g["x"], g["y"], g["z"], g["level"])
    assert np.array_equal(enc, g["encode"])
    dx, dy, dz = oracle.hilbert_decode(g["encode"], g["level"])
    assert np.array_equal(dx, g["dec_x"]) and np.array_equal(dy, g["dec_y"]) and np.array_equal(dz, g["dec_z"])
    assert np.array_equal(oracle.rotdir_encode(g["x"], g["y"], g["z"], 8), g["rotdir8"])
    assert np.array_equal(oracle.rotdir_encode(g["x"], g["y"], g["z"], 10), g["rotdir10"])


def test_hilbert_reference_known_answers(oracle):
    # tests/unit/hilbert/test_hilbert_codec.cpp:33-57
    assert oracle.hilbert_encode([0], [0], [0], 1)[0] == 0
    xs, ys, zs = np.meshgrid([0, 1], [0, 1], [0, 1], indexing="ij")
    h = oracle.hilbert_encode(xs.ravel(), ys.ravel(), zs.ravel(), 1)
    assert sorted(h.tolist()) == list(range(8))
    # level-1 curve order restated from the table (survey section 8c)
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]
    o = np.array(order)
    assert oracle.hilbert_encode(o[:, 0], o[:, 1], o[:, 2], 1).tolist() == list(range(8))
    assert oracle.lib().or_hilbert_encode(1, 2, 3, 21) == 22
    assert oracle.lib().or_hilbert_encode(5, 0, 7, 8) == 152
    assert oracle.lib().or_rotdir_encode(5, 0, 7, 8) == 357
    # canonical form (test_hilbert_codec.cpp:94-106)
    assert oracle.lib().or_hilbert_encode(123, 456, 789, 5) >> 15 == 0


def test_hilbert_round_trip_and_parent(oracle):
    rng = np.random.default_rng(42)
    for level in range(1, 22):
        n = 1000
        x, y, z = (rng.integers(0, 1 << level, n).astype(np.int32) for _ in range(3))
        h = oracle.hilbert_encode(x, y, z, level)
        assert np.all(h < (1 << (3 * level)))
        dx, dy, dz = oracle.hilbert_decode(h, level)
        assert np.array_equal(dx, x) and np.array_equal(dy, y) and np.array_equal(dz, z)
        if level < 21:  # getParent(child) == parent (test_hilbert_codec.cpp:108-125)
            q = rng.integers(0, 8, n)
            cx, cy, cz = 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1)
            hc = oracle.hilbert_encode(cx, cy, cz, level + 1)
            assert np.array_equal(hc >> np.uint64(3), h)
    # bijection at level 4 (all 4096 cells)
    g = np.arange(4096)
    h = oracle.hilbert_encode(g & 15, (g >> 4) & 15, g >> 8, 4)
    assert len(np.unique(h)) == 4096


def test_hilbert_oracle_matches_reference_build(oracle):
    R = oracle.ref()
    if R is None:
        pytest.skip("reference codec not built here (oracle/_ref absent)")
    rng = np.random.default_rng(7)
    for _ in range(2000):
        l = int(rng.integers(0, 22))
        x, y, z = (int(v) for v in rng.integers(-(1 << 22), 1 << 22, 3))
        assert oracle.lib().or_hilbert_encode(x, y, z, l) == R.ref_hilbert_encode(x, y, z, l)
        ml = int(rng.integers(1, 12))
        assert oracle.lib().or_rotdir_encode(x, y, z, ml) == R.ref_rotdir_encode(x, y, z, ml)


def test_key63_prefix_aligned(oracle):
    rng = np.random.default_rng(3)
    lv = rng.integers(0, 11, 500).astype(np.uint8)
    x, y, z = (np.array([rng.integers(0, 1 << int(l)) for l in lv], np.int32) for _ in range(3))
    k = oracle.key63(x, y, z, lv)
    for i in range(0, 500, 25):
        l = int(lv[i])
        # descendants lie in [k, k + 8^(21-L))
        if l < 21:
            ck = oracle.key63([2 * x[i] + 1], [2 * y[i]], [2 * z[i] + 1], [l + 1])[0]
            assert k[i] <= ck < k[i] + (1 << (3 * (21 - l)))


@pytest.mark.parametrize("bits", [(0, 64), (0, 40), (8, 40), (33, 63)])
def test_sort_stable_vs_numpy(oracle, bits):
    rng = np.random.default_rng(11)
    n = 20000
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    keys[::7] = keys[::3][: len(keys[::7])]  # duplicates
    keys[5::11] &= np.uint64(0xFFFF)
    vals = np.arange(n, dtype=np.uint32)
    b, e = bits
    k, v = oracle.sort_pairs(keys, vals, b, e)
    digit = (keys >> np.uint64(b)) & np.uint64((1 << (e - b)) - 1) if e - b < 64 else keys
    order = np.argsort(digit, kind="stable")
    assert np.array_equal(v, vals[order]) and np.array_equal(k, keys[order])


def test_scan_vs_cumsum(oracle):
    rng = np.random.default_rng(5)
    for n in [0, 1, 255, 4096, 65537, 300001]:
        a = rng.integers(0, 9, n).astype(np.uint32)
        out, tot = oracle.scan_exclusive(a)
        ref = np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]) if n else np.zeros(0)
        assert np.array_equal(out, ref.astype(np.uint32)) and tot == int(a.sum())


def test_hash_capacity_rule(oracle):
    # HashTableManager.cpp:47-57; test_hash_table.cpp:149-175
    for n in [0, 5, 10, 100, 1000, 10000, 700000, 1 << 20, 64 << 20]:
        cap = oracle.hash_capacity(n)
        assert cap & (cap - 1) == 0 and cap >= n / 0.6 and cap >= 1 << 20
    assert oracle.hash_capacity(64 << 20) == 1 << 27
    assert oracle.hash_capacity(700000) == 1 << 21


def test_hash_query_vs_dict(oracle):
    rng = np.random.default_rng(9)
    n = 50000
    keys = rng.integers(0, 1 << 40, n, dtype=np.uint64)
    keys[100:200] = keys[0:100]  # duplicates keep the smallest value
    keys[300] = np.uint64(0xFFFFFFFFFFFFFFFF)  # EMPTY is skipped
    vals = rng.permutation(n).astype(np.uint32)
    cap = oracle.hash_capacity(n)
    tk, tv, pc = oracle.hash_build(keys, vals, cap)
    ref = {}
    for k, v in zip(keys.tolist(), vals.tolist()):
        if k == 0xFFFFFFFFFFFFFFFF:
            continue
        ref[k] = min(v, ref.get(k, 1 << 33))
    q = np.concatenate([keys, rng.integers(1 << 41, 1 << 42, 1000, dtype=np.uint64)])
    got = oracle.hash_query(tk, tv, cap, q)
    exp = np.array([ref.get(k, 0xFFFFFFFF) for k in q.tolist()], np.uint32)
    assert np.array_equal(got, exp)
    assert pc.max() < 32
