"""CPU: the oracle's adaptation restatements against the reference's unit
test fixtures and brute-force checkers."""
import numpy as np

from tests.amr_util import (balanced_mesh, brute_face_neighbors, brute_merge_groups,
                            check_balance_d3q19, random_octree)


def test_split_fixture(oracle):
    # SplitEngineTest.cpp:35-63 / AdaptationIntegrationTest.cpp:62-100
    c = oracle.CellSet(1, flag=[1])
    cs, ch = oracle.split(c, 8)
    assert cs[0] == 0 and ch.n == 8
    assert np.all(ch.level == 1)
    q = np.arange(8)
    assert np.array_equal(ch.x, q & 1) and np.array_equal(ch.y, (q >> 1) & 1) and np.array_equal(ch.z, q >> 2)
    out, _, _, _ = oracle.regrid(c, -1, 8, 8)
    assert out.n == 8 and np.all(out.level == 1)


def test_merge_fixture(oracle):
    # MergeEngineTest.cpp:35-71 / AdaptationIntegrationTest.cpp:100-128
    q = np.arange(8)
    c = oracle.CellSet(8, x=q & 1, y=(q >> 1) & 1, z=q >> 2, level=np.ones(8), flag=-np.ones(8))
    gid, lead, par = oracle.merge(c)
    assert par.n == 1 and par.level[0] == 0 and par.x[0] == 0 and np.all(gid == 0)
    f = np.arange(16, dtype=np.float32)
    pf = oracle.merge_fields(c, lead, f, 2, 0)
    assert np.allclose(pf, [f[0::2].mean(), f[1::2].mean()])
    out, _, _, _ = oracle.regrid(c, -1, 8, 8)
    assert out.n == 1 and out.level[0] == 0


def test_balance_fixture(oracle):
    # BalanceEnforcerTest.cpp:37-93: cells at anchors {0, 128} (2^8 units),
    # levels {1, 3} -> native coordinates {0, 4}
    c = oracle.CellSet(2, x=[0, 4], level=[1, 3])
    it, conv, nv, nm = oracle.balance(c, 0, 8, 8, 10)
    assert conv == 1 and nv > 0 and nm > 0
    assert c.flag.tolist() == [1, 0]


def test_face_neighbors_vs_brute_force(oracle):
    c = random_octree(seed=1, root=(2, 2, 2), depth=3)
    got = oracle.face_neighbors(c, (2, 2, 2)).reshape(-1, 6)
    exp = brute_face_neighbors(c, (2, 2, 2))
    assert np.array_equal(got, exp)


def test_merge_vs_brute_force(oracle):
    c = random_octree(seed=2, root=(2, 2, 2), depth=3)
    rng = np.random.default_rng(4)
    c.flag[:] = np.where(rng.random(c.n) < 0.7, -1, 0)
    gid, lead, par = oracle.merge(c)
    assert sorted(lead.tolist()) == brute_merge_groups(c)


def test_balance_mode1_converges_to_balanced_mesh(oracle):
    c = random_octree(seed=3, root=(2, 2, 2), depth=4, p_split=0.35)
    out = balanced_mesh(oracle, c, (2, 2, 2), max_level=5)
    assert check_balance_d3q19(out, (2, 2, 2))
