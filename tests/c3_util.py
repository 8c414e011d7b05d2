"""Synthetic inputs of the AMR integer micro-benchmark (seeded)."""


def c3_cells(O, n, seed=42, p_split=0.05, p_merge=0.03):
    """Random cells at levels 0-10 in a 2^10 root domain (splitmix64 streams),
    split flags Bernoulli(p_split), merge flags -1 on the members of complete
    sibling groups chosen with probability p_merge -- the same rule the
    device generator fl_gen_flags applies."""
    x, y, z, lv = O.gen_cells(seed, n, 10)
    c = O.CellSet(n, x=x, y=y, z=z, level=lv)
    return O.gen_flags(c, seed, p_split, p_merge)
