"""Test helpers: small octrees, brute-force checkers and test meshes.

Brute-force functions here are independent of the oracle's hash-based
restatements (quadratic scans over all cells) so that they can pin them.
"""
import numpy as np

INVALID = 0xFFFFFFFF
FINER = 0xFFFFFFFE
D18 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
       (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1),
       (0, 1, -1), (0, -1, 1)]
F6 = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def _cellset(cells):
    from oracle import oracle as O
    a = np.array(cells, np.int64).reshape(-1, 4)
    return O.CellSet(len(a), x=a[:, 0], y=a[:, 1], z=a[:, 2], level=a[:, 3])


def random_octree(seed, root=(2, 2, 2), depth=3, p_split=0.3):
    rng = np.random.default_rng(seed)
    stack = [(x, y, z, 0) for z in range(root[2]) for y in range(root[1]) for x in range(root[0])]
    out = []
    while stack:
        x, y, z, l = stack.pop()
        if l < depth and rng.random() < p_split:
            for q in range(8):
                stack.append((2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1), l + 1))
        else:
            out.append((x, y, z, l))
    out.sort(key=lambda c: (c[3], c[2], c[1], c[0]))
    perm = np.random.default_rng(seed + 100).permutation(len(out))
    return _cellset([out[i] for i in perm])


def _inside(root, l, p):
    return all(0 <= p[a] < (root[a] << l) for a in range(3))


def brute_face_neighbors(c, root):
    n = c.n
    out = np.full((n, 6), INVALID, np.uint32)
    X, Y, Z, L = (c.x.astype(np.int64), c.y.astype(np.int64), c.z.astype(np.int64), c.level.astype(np.int64))
    for i in range(n):
        li = int(L[i])
        for f, d in enumerate(F6):
            p = (int(X[i]) + d[0], int(Y[i]) + d[1], int(Z[i]) + d[2])
            if not _inside(root, li, p):
                continue
            best = -1
            for j in np.nonzero(L <= li)[0]:
                sh = li - int(L[j])
                if (p[0] >> sh, p[1] >> sh, p[2] >> sh) == (X[j], Y[j], Z[j]):
                    if best < 0 or L[j] > L[best] or (L[j] == L[best] and j < best):
                        best = j
            if best >= 0:
                out[i, f] = best
                continue
            m = (L == li + 1) & ((X >> 1) == p[0]) & ((Y >> 1) == p[1]) & ((Z >> 1) == p[2])
            if m.any():
                out[i, f] = FINER
    return out


def brute_merge_groups(c):
    leaders = []
    key = {}
    for j in range(c.n):
        key.setdefault((int(c.x[j]), int(c.y[j]), int(c.z[j]), int(c.level[j])), j)
    for i in range(c.n):
        x, y, z, l = int(c.x[i]), int(c.y[i]), int(c.z[i]), int(c.level[i])
        if c.flag[i] != -1 or c.state[i] != 0 or (x | y | z) & 1 or l == 0:
            continue
        if key[(x, y, z, l)] != i:
            continue
        ok = True
        for q in range(8):
            j = key.get((x | (q & 1), y | ((q >> 1) & 1), z | ((q >> 2) & 1), l))
            if j is None or c.flag[j] != -1 or c.state[j] != 0:
                ok = False
                break
        if ok:
            leaders.append(i)
    return sorted(leaders)


def balanced_mesh(O, c, root, max_level, max_passes=20):
    for _ in range(max_passes):
        out, res, _, _ = O.regrid(c, 1, 19, max_level, 64, root)
        if out.n == c.n and res[3] == 0:
            return out
        c = out
    raise AssertionError("balance did not settle")


def check_balance_d3q19(c, root):
    """Every leaf's 18 neighbour cubes overlap only leaves at most one level finer."""
    X, Y, Z, L = (c.x.astype(np.float64), c.y.astype(np.float64), c.z.astype(np.float64), c.level.astype(np.int64))
    h = 2.0 ** (-L)
    lo = np.stack([X * h, Y * h, Z * h], 1)
    hi = lo + h[:, None]
    for i in range(c.n):
        hi_ = 2.0 ** (-int(L[i]))
        for d in D18:
            p = np.array([c.x[i] + d[0], c.y[i] + d[1], c.z[i] + d[2]], np.float64)
            if not _inside(root, int(L[i]), p.astype(np.int64)):
                continue
            plo, phi = p * hi_, (p + 1) * hi_
            ov = np.all((lo < phi) & (hi > plo), axis=1)
            if np.any(L[ov] > L[i] + 1):
                return False
    return True


# ---------------------------------------------------------------- LBM meshes
def refine_mesh(O, base, root, regions, max_level, max_passes=40):
    """Refine `base` (CellSet) so that cells intersecting region r (a function
    of (lo, hi) cube arrays in level-0 units returning bool) reach its level;
    regions = [(level, fn), ...].  Balanced with the D3Q19 cube rule."""
    c = base
    for _ in range(max_passes):
        h = 2.0 ** (-c.level.astype(np.float64))
        lo = np.stack([c.x * h, c.y * h, c.z * h], 1)
        hi = lo + h[:, None]
        want = np.zeros(c.n, np.int64)
        for lev, fn in regions:
            want = np.maximum(want, np.where(fn(lo, hi), lev, 0))
        c = c.copy()
        c.flag[:] = np.where((want > c.level) & (c.state == 0), 1, 0)
        out, res, _, _ = O.regrid(c, 1, 19, max_level, 64, root)
        if out.n == c.n and not c.flag.any():
            return out
        c = out
    raise AssertionError("refinement did not settle")


def box_region(lo_box, hi_box):
    lo_box = np.array(lo_box, np.float64)
    hi_box = np.array(hi_box, np.float64)
    return lambda lo, hi: np.all((lo < hi_box) & (hi > lo_box), axis=1)


def ball_region(center, r):
    c = np.array(center, np.float64)

    def fn(lo, hi):
        d = np.maximum(np.maximum(lo - c, 0), c - hi)
        return (d * d).sum(1) <= r * r
    return fn


def uniform_cells(O, n):
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    m = x.size
    return O.CellSet(m, x=x.ravel(), y=y.ravel(), z=z.ravel())


def mark_solid_ball(c, center, r):
    h = 2.0 ** (-c.level.astype(np.float64))
    ctr = np.stack([(c.x + 0.5) * h, (c.y + 0.5) * h, (c.z + 0.5) * h], 1)
    inside = ((ctr - np.array(center)) ** 2).sum(1) <= r * r
    c.state[:] = np.where(inside, 1, 0)
    return c
