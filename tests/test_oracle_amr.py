"""CPU: the oracle's adaptation restatements against the reference's unit
test fixtures and brute-force checkers."""
import numpy as np

from tests.amr_util import (D18, _inside, balanced_mesh, brute_face_neighbors, brute_merge_groups,
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


def test_balance_preflagged_hand_derived(oracle):
    """Expected values stepped through BalanceEnforcer.cpp:145-263 and
    balance_enforce.cl:57-172 by hand, anchor level 8 (MAX_REFINEMENT_LEVEL).

    A = (x 0, level 1), B = (x 2, level 2, flag 1).  Iteration 1 runs on a
    copy of the stored levels: A's +x test point 128 is first held at probe
    level 2 (B), and 2 > 1 + 1 is false; B's -x point 127 is held by A at
    level 1.  No violation: 1 iteration, converged, nothing marked."""
    c = oracle.CellSet(2, x=[0, 2], level=[1, 2], flag=[0, 1])
    assert oracle.balance(c, 0, 8, 8, 10) == (1, 1, 0, 0)
    assert c.flag.tolist() == [0, 1]


def test_balance_preflag_enters_after_first_marking(oracle):
    """Hand-derived from the same reference loop.  A = (0,0,0 L1),
    B = (2,0,0 L2, flag 1), C = (0,1,0 L1), E = (4,4,0 L3); anchors in 2^8
    units: A 0, B (128,0,0), C (0,128,0), E (128,128,0).

    it 1 (shadow = stored levels 1,2,1,3): C's +x point (128,128,0) is first
    held at probe level 3, by E, and 3 > 1 + 1: C violates and is marked.
    A's +x point (128,0,0) is held at probe level 2 by B: no violation.
    Shadow becomes 1,3,2,3 (B's own flag now counts).
    it 2: A's +x point is now held at probe level 3 by B: A violates and is
    marked.  Shadow 2,3,2,3.
    it 3: no test point finds a holder more than one level finer: converged.
    Result (3 iterations, converged, 2 violations, 2 marked), flags 1,1,1,0.
    Seeding the shadow with the flags from the start (the rule of the lattice
    mode) would mark A and C together in iteration 1 and stop after 2."""
    geo = dict(x=[0, 2, 0, 4], y=[0, 0, 1, 4], level=[1, 2, 1, 3], flag=[0, 1, 0, 0])
    c = oracle.CellSet(4, **geo)
    assert oracle.balance(c, 0, 8, 8, 10) == (3, 1, 2, 2)
    assert c.flag.tolist() == [1, 1, 1, 0]
    c = oracle.CellSet(4, **geo)
    assert oracle.balance(c, 0, 8, 8, 1) == (1, 0, 1, 1)
    assert c.flag.tolist() == [0, 1, 1, 0]


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


def _brute_balance(c, mode, M, max_split, max_iters=64, root=(1, 1, 1)):
    """Independent restatement of the balance loop by exhaustive search.

    mode 0 follows detect_balance_violations / mark_cascading_refinement
    (src/adaptation/kernels/balance_enforce.cl:57-172) on shadow levels
    (BalanceEnforcer.cpp:53-62): six face test points in 2^M units, levels
    probed coarse to fine, the first holder (anchor equal, shadow level equal)
    decides, violation when its level exceeds the cell's shadow + 1.
    mode 1 is the lattice rule: a leaf overlapping any of the 18 D3Q19
    neighbour cubes may be at most one shadow level finer."""
    n = c.n
    X, Y, Z, L = (c.x.astype(np.int64), c.y.astype(np.int64), c.z.astype(np.int64), c.level.astype(np.int64))
    flag = c.flag.astype(np.int64).copy()
    st = c.state
    anchors = [(int(X[j]) << (M - int(L[j])), int(Y[j]) << (M - int(L[j])), int(Z[j]) << (M - int(L[j])))
               if L[j] <= M else None for j in range(n)]
    it = conv = tv = tm = 0
    # the reference starts from a copy of the stored levels
    # (BalanceEnforcer.cpp:165-167) and adds the split flags only after a
    # marking pass (:248-254); the lattice rule counts them from the start
    sh = L.copy() if mode == 0 else L + (flag > 0)
    while it < max_iters:
        viol = np.zeros(n, bool)
        for i in range(n):
            if mode == 0:
                if st[i] != 0 or L[i] > M:
                    continue
                a = anchors[i]
                size = 1 << (M - int(sh[i])) if sh[i] <= M else 0
                tps = [(a[0] - 1, a[1], a[2]), (a[0] + size, a[1], a[2]), (a[0], a[1] - 1, a[2]),
                       (a[0], a[1] + size, a[2]), (a[0], a[1], a[2] - 1), (a[0], a[1], a[2] + size)]
                for p in tps:
                    for l in range(M + 1):
                        m = ~((1 << (M - l)) - 1)
                        q = (p[0] & m, p[1] & m, p[2] & m)
                        if min(q) < 0 or max(q) >= (1 << M):
                            continue
                        if any(anchors[j] == q and sh[j] == l for j in range(n)):
                            viol[i] = viol[i] or l > sh[i] + 1
                            break
                    if viol[i]:
                        break
            else:
                li = int(L[i])
                for d in D18:
                    P = (int(X[i]) + d[0], int(Y[i]) + d[1], int(Z[i]) + d[2])
                    if not _inside(root, li, P):
                        continue
                    for j in range(n):
                        lj = int(L[j])
                        if lj <= li:
                            continue
                        s = lj - li
                        if (int(X[j]) >> s, int(Y[j]) >> s, int(Z[j]) >> s) == P and sh[j] > sh[i] + 1:
                            viol[i] = True
                            break
                    if viol[i]:
                        break
        it += 1
        nv = int(viol.sum())
        tv += nv
        if nv == 0:
            conv = 1
            break
        mk = viol & (st == 0) & (L < max_split) & (flag <= 0)
        flag[mk] = 1
        tm += int(mk.sum())
        if mk.sum() == 0:
            break
        sh = L + (flag > 0)
    return it, conv, tv, tm, flag


def test_balance_reference_rule_vs_brute_force(oracle):
    from tests.amr_util import random_octree
    nontrivial = 0
    for seed in (5, 6, 7):
        # a complete octree of the unit root: random subtrees below its 8 children
        c = random_octree(seed=seed, root=(2, 2, 2), depth=3, p_split=0.5)
        c.level[:] += 1
        rng = np.random.default_rng(seed)
        c.flag[:] = np.where(rng.random(c.n) < 0.15, 1, 0)
        exp = _brute_balance(c, 0, 6, 6)
        got = oracle.balance(c, 0, 6, 6)
        assert got == exp[:4] and np.array_equal(c.flag, exp[4]), seed
        nontrivial += got[2] > 0
    assert nontrivial >= 2


def test_balance_lattice_rule_vs_brute_force(oracle):
    from tests.amr_util import random_octree
    for seed in (8, 9):
        c = random_octree(seed=seed, root=(2, 1, 1), depth=4, p_split=0.4)
        rng = np.random.default_rng(seed)
        c.flag[:] = np.where(rng.random(c.n) < 0.1, 1, 0)
        exp = _brute_balance(c, 1, 19, 6, root=(2, 1, 1))
        got = oracle.balance(c, 1, 19, 6, root=(2, 1, 1))
        assert got == exp[:4] and np.array_equal(c.flag, exp[4]), seed


def test_compact_append_closed_form(oracle):
    """compact_cells.cl:9-214: survivors = flag <= 0 and no merge group, in
    index order; then the children block; then the parents block."""
    c = random_octree(seed=12, root=(2, 2, 2), depth=3, p_split=0.4)
    rng = np.random.default_rng(1)
    c.flag[:] = rng.choice([-1, 0, 1], c.n, p=[0.5, 0.3, 0.2])
    c.material[:] = rng.integers(0, 1000, c.n)
    gid, lead, par = oracle.merge(c)
    _, ch = oracle.split(c, 8)
    out = oracle.compact_append(c, gid, ch, par)
    keep = (c.flag <= 0) & (gid == 0xFFFFFFFF)
    for nm in ("x", "y", "z", "level", "state", "material"):
        exp = np.concatenate([getattr(c, nm)[keep], getattr(ch, nm), getattr(par, nm)])
        assert np.array_equal(getattr(out, nm), exp), nm
    # split children: parent order, child q at (2x + (q&1), 2y + (q>>1&1), 2z + (q>>2))
    sp = np.nonzero((c.flag > 0) & (c.state == 0) & (c.level < 8))[0]
    q = np.tile(np.arange(8), len(sp))
    pi = np.repeat(sp, 8)
    assert np.array_equal(ch.x, 2 * c.x[pi] + (q & 1)) and np.array_equal(ch.z, 2 * c.z[pi] + (q >> 2))
    assert np.array_equal(ch.material, c.material[pi])


def test_merge_fields_rules(oracle):
    """merge_cells.cl merge_fields: sum over the 8 siblings in order 0..7,
    times 0.125 for rule 0."""
    q = np.arange(8)
    c = oracle.CellSet(16, x=np.concatenate([q & 1, 2 + (q & 1)]), y=np.concatenate([(q >> 1) & 1] * 2),
                       z=np.concatenate([q >> 2] * 2), level=np.ones(16), flag=-np.ones(16))
    gid, lead, par = oracle.merge(c, root=(2, 1, 1))
    assert par.n == 2
    f = np.random.default_rng(3).standard_normal(16 * 3).astype(np.float32)
    for rule in (0, 1):
        got = oracle.merge_fields(c, lead, f, 3, rule).reshape(-1, 3)
        for g, ld in enumerate(lead):
            members = np.nonzero(gid == g)[0]
            acc = np.zeros(3, np.float32)
            for m in sorted(members, key=lambda j: (c.z[j] & 1, c.y[j] & 1, c.x[j] & 1)):
                acc = (acc + (f.reshape(-1, 3)[m] * np.float32(0.125) if rule == 0 else f.reshape(-1, 3)[m]))
            assert np.allclose(got[g], acc, rtol=1e-6, atol=1e-7)


def test_cell_key_packing(oracle):
    rng = np.random.default_rng(0)
    n = 20000
    lv = rng.integers(0, 20, n)
    x, y, z = (rng.integers(0, 1 << 19, n) for _ in range(3))
    k = oracle.cell_hkey(x, y, z, lv)
    exp = (lv.astype(np.uint64) << 57) | (x.astype(np.uint64) << 38) | (y.astype(np.uint64) << 19) | z.astype(np.uint64)
    assert np.array_equal(k, exp)
    assert len(np.unique(k)) == len(np.unique(np.stack([lv, x, y, z]), axis=1).T)


def test_generators_follow_the_workload_definition(oracle):
    """Synthetic C3 input (no counterpart in the reference): levels uniform
    on 0..10, coordinates inside each level's grid, split flags
    Bernoulli(0.05), merge flags only on complete sibling groups."""
    n = 200000
    x, y, z, lv = oracle.gen_cells(42, n, 10)
    assert lv.max() == 10 and lv.min() == 0
    assert np.all(x < (1 << lv.astype(np.int64))) and np.all(z < (1 << lv.astype(np.int64)))
    cnt = np.bincount(lv, minlength=11) / n
    assert np.all(np.abs(cnt - 1 / 11) < 5 * np.sqrt(1 / 11 / n))
    c = oracle.CellSet(n, x=x, y=y, z=z, level=lv)
    oracle.gen_flags(c, 42, 0.05, 0.03)
    ps = (c.flag == 1).mean()
    assert abs(ps - 0.05) < 5 * np.sqrt(0.05 * 0.95 / n)
    first = {}
    for i in range(n):
        first.setdefault((int(lv[i]), int(x[i]), int(y[i]), int(z[i])), i)
    m = np.nonzero(c.flag == -1)[0]
    assert len(m) > 0
    for i in m[:2000]:
        L, X, Y, Z = int(lv[i]), int(x[i]) & ~1, int(y[i]) & ~1, int(z[i]) & ~1
        assert L > 0
        assert all((L, X | (q & 1), Y | ((q >> 1) & 1), Z | (q >> 2)) in first for q in range(8))
