"""The FP16C storage format's definition (CPU)."""
import numpy as np

from tests import fp16c_ref as F


def test_grid_is_monotone_and_round_trips():
    g = F.decode(np.arange(0x8000, dtype=np.uint32))
    assert g[0] == 0 and np.all(np.diff(g) > 0)
    assert g[0x800] == 2.0 ** -13 and g[0x7FF] == 2047 * 2.0 ** -24
    assert abs(g[-1] - 2 * (2 - 2 ** -11)) < 1e-15  # largest value just below 4
    codes = np.arange(0x10000, dtype=np.uint32)
    back = F.encode_rne(F.decode(codes))
    assert np.array_equal(back[codes != 0x8000], codes[codes != 0x8000])  # -0 encodes as +0 in the restatement


def test_precision_is_one_bit_better_than_binary16():
    rng = np.random.default_rng(0)
    x = rng.uniform(-0.9, 0.9, 200000) * 10.0 ** rng.uniform(-3, 0, 200000)
    err = np.abs(F.decode(F.encode_rne(x)) - x)
    normal = np.abs(x) >= 2.0 ** -13
    assert (err[normal] / np.abs(x[normal])).max() <= 2.0 ** -12 * (1 + 1e-12)   # half an 11-bit ulp
    assert err[~normal].max() <= 2.0 ** -25 * (1 + 1e-12)                      # half the subnormal spacing
    e16 = np.abs(x.astype(np.float16).astype(np.float64) - x)
    assert np.median(err[normal] / np.abs(x[normal])) < 0.6 * np.median(e16[normal] / np.abs(x[normal]))
