"""The device FP16C codec (fluidloom_amd/csrc/fp16c_codec.hpp), compiled
for the host and run on the CPU, against the numpy statement of the format
(tests/fp16c_ref.py, itself pinned by test_fp16c_format.py):
  * decode of all 65536 codes equals the format's values exactly;
  * round-to-nearest encode equals the format's nearest-even rounding;
  * stochastic encode: every input lands on its two neighbouring grid
    values, and the share of the 4096 random values that round up is the
    dropped fraction in 1/4096 (exact for |x| >= 2^-13, where the fraction
    has 12 bits; below, the fraction rounded to 2^-36 in x), so the mean
    over the random bits reproduces x;
  * inf, NaN and |x| beyond the largest code saturate with the sign kept.
The collide stores only through the stochastic form, so this is the check
of the bits a simulation writes."""
import subprocess

import numpy as np
import pytest

from fluidloom_amd import build
from tests import fp16c_ref as F


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("fp16c") / "fp16c_codec_check"
    src = build.ROOT / "tests" / "fp16c_codec_check.cpp"
    cmd = [build.HIPCC, "-std=c++17", "-O2", "-x", "hip", "--offload-host-only", f"-I{build.CSRC}", str(src),
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return str(out)


def _encode(exe, x):
    x = np.ascontiguousarray(x, np.float32)
    r = subprocess.run([exe, "encode"], input=x.tobytes(), capture_output=True)
    assert r.returncode == 0, r.returncode
    rec = np.frombuffer(r.stdout, np.uint32).reshape(-1, 5)
    assert len(rec) == len(x)
    return rec[:, 0], rec[:, 1], rec[:, 2], rec[:, 3]


def _inputs(seed=11):
    rng = np.random.default_rng(seed)
    grid = F.decode(np.arange(0x8000, dtype=np.uint32)).astype(np.float32)
    n = 200_000
    e = rng.uniform(-40, 2.2, n)
    dense = (rng.choice([-1.0, 1.0], n) * 2.0 ** e).astype(np.float32)
    mids = ((grid[:-1].astype(np.float64) + grid[1:]) / 2).astype(np.float32)   # ties for RNE
    quarter = (grid[:-1].astype(np.float64) * 0.75 + grid[1:] * 0.25).astype(np.float32)
    tiny = np.array([1e-45, -1e-45, 2.0 ** -37, 3 * 2.0 ** -38, 2.0 ** -25, -2.0 ** -25, 2.0 ** -13, -2.0 ** -13,
                     np.nextafter(np.float32(2.0 ** -13), np.float32(0)), 3.9990234, 3.9995, 4.0, -4.0, 100.0,
                     3.0e38], np.float32)
    return np.concatenate([dense, grid[1:], -grid[1:], mids, -mids, quarter, tiny])


def test_decode_every_code(exe):
    r = subprocess.run([exe, "decode"], capture_output=True)
    assert r.returncode == 0
    got = np.frombuffer(r.stdout, np.float32).astype(np.float64)
    codes = np.arange(0x10000, dtype=np.uint32)
    assert np.array_equal(got, F.decode(codes))
    assert np.array_equal(np.signbit(got), (codes & 0x8000) != 0)


def test_round_to_nearest_matches_format(exe):
    x = _inputs()
    rne, _, _, _ = _encode(exe, x)
    want = F.encode_rne(x.astype(np.float64))
    nz = x != 0
    assert np.array_equal(rne[nz], want[nz])


def test_stochastic_rounding_is_exact_fraction(exe):
    x = _inputs()
    _, up, lo, hi = _encode(exe, x)
    a = np.abs(x.astype(np.float64))
    assert np.array_equal(lo & 0x8000, np.where(np.signbit(x), 0x8000, 0))
    grid = F.decode(np.arange(0x8000, dtype=np.uint32))
    sat = a >= grid[-1]
    assert np.all((lo[sat] & 0x7FFF) == 0x7FFF) and np.all(up[sat] == 0)
    a, lo, hi, up = a[~sat], lo[~sat] & 0x7FFF, hi[~sat] & 0x7FFF, up[~sat]
    i = np.searchsorted(grid, a, side="right") - 1
    step = grid[i + 1] - grid[i]
    frac = (a - grid[i]) / step * 4096.0          # exact: dyadic values
    normal = a >= 2.0 ** -13
    assert np.all(frac[normal] == np.floor(frac[normal]))   # 12 dropped bits exactly
    total = i * 4096 + np.rint(frac).astype(np.int64)    # below 2^-13: x rounded to 2^-36, ties even
    assert np.array_equal(lo, total >> 12)
    assert np.array_equal(up, total & 0xFFF)
    moved = up > 0
    assert np.array_equal(hi[moved], lo[moved] + 1) and np.array_equal(hi[~moved], lo[~moved])
    # the mean over the random bits reproduces x exactly in the normal range
    mean = grid[lo] + up / 4096.0 * (grid[np.minimum(lo + 1, 0x7FFF)] - grid[lo])
    assert np.array_equal(mean[normal], a[normal])
    assert np.abs(mean - a).max() <= 2.0 ** -37


def test_special_values_saturate(exe):
    x = np.array([np.inf, -np.inf, np.nan, -np.nan], np.float32)
    rne, up, lo, _ = _encode(exe, x)
    assert list(lo) == [0x7FFF, 0xFFFF, 0x7FFF, 0xFFFF] and not up.any()
    assert list(rne) == [0x7FFF, 0xFFFF, 0x7FFF, 0xFFFF]
