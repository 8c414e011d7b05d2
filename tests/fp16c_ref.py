"""numpy restatement of the FP16C storage format (x = (f - w) / w; sign,
4-bit exponent with bias 14, 11-bit mantissa; subnormals below 2^-13 with
spacing 2^-24; |x| >= 4 saturates), used to check the device codec."""
import numpy as np


def decode(code):
    code = np.asarray(code, np.uint32)
    m = code & 0x7FFF
    e = m >> 11
    frac = (m & 0x7FF).astype(np.float64)
    a = np.where(e > 0, np.ldexp(1.0 + frac / 2048.0, e.astype(np.int64) - 14), m.astype(np.float64) * 2.0 ** -24)
    return np.where(code & 0x8000, -a, a)


def encode_rne(x):
    """Round to nearest (ties to even) onto the format's grid; saturating."""
    x = np.asarray(x, np.float64)
    a = np.abs(x)
    grid = decode(np.arange(0x8000, dtype=np.uint32))  # all non-negative values, increasing
    i = np.clip(np.searchsorted(grid, a), 1, len(grid) - 1)
    lo, hi = grid[i - 1], grid[i]
    pick_hi = (a - lo > hi - a) | ((a - lo == hi - a) & ((i % 2) == 0))
    code = np.where(pick_hi, i, i - 1).astype(np.uint32)
    code = np.where(a >= grid[-1], 0x7FFF, code)
    return np.where(x < 0, code | 0x8000, code).astype(np.uint32)
