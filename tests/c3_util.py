"""Synthetic inputs of the AMR integer micro-benchmark (seeded)."""
import numpy as np


def c3_cells(O, n, seed=42, p_split=0.05, p_merge=0.03):
    """Random cells at levels 0-10 in a 2^10 root domain (splitmix64 streams),
    split flags Bernoulli(p_split), merge flags -1 on every member of
    complete sibling groups chosen with probability p_merge."""
    x, y, z, lv = O.gen_cells(seed, n, 10)
    c = O.CellSet(n, x=x, y=y, z=z, level=lv)
    rng = np.random.default_rng(seed)
    c.flag[:] = np.where(rng.random(n) < p_split, 1, 0)
    key = {}
    for i in range(n):
        key.setdefault((int(x[i]), int(y[i]), int(z[i]), int(lv[i])), i)
    for i in range(n):
        if lv[i] == 0 or (x[i] | y[i] | z[i]) & 1 or key[(int(x[i]), int(y[i]), int(z[i]), int(lv[i]))] != i:
            continue
        sib = [key.get((int(x[i]) | (q & 1), int(y[i]) | ((q >> 1) & 1), int(z[i]) | ((q >> 2) & 1), int(lv[i])))
               for q in range(8)]
        if None in sib or rng.random() >= p_merge:
            continue
        for j in sib:
            if c.flag[j] == 0:
                c.flag[j] = -1
    return c
