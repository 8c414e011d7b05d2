"""GPU: the allocation-free scan entry point (fl_scan_exclusive_u32_noalloc)
against fl_scan_exclusive_u32 through the C ABI, bit for bit (integer work):
two calls in a row, the second on a smaller input, so the second must not
see the first call's look-back state left in the shared scratch array; then
a zero-element call (total 0) and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan_ref(fl, d_in, n):
    L = fl.lib()
    out = fl.DeviceArray(max(n, 1), np.uint32)
    tot = fl.DeviceArray(1, np.uint64)
    tb = L.fl_scan_temp_bytes(n)
    t = fl.DeviceArray(max(tb, 8), np.uint8)
    fl.check(L.fl_scan_exclusive_u32(d_in.p, out.p, tot.p, n, t.p, tb, None))
    return out.to_host()[:n], int(tot.to_host()[0])


def _scan_noalloc(fl, d_in, n):
    L = fl.lib()
    out = fl.DeviceArray(max(n, 1), np.uint32)
    tot = fl.DeviceArray(1, np.uint64)
    fl.check(L.fl_memset(tot.p, 0xAB, 8, None))
    fl.check(L.fl_scan_exclusive_u32_noalloc(d_in.p, out.p, tot.p, n, None))
    fl.check(L.fl_stream_sync(None))
    return out.to_host()[:n], int(tot.to_host()[0])


def test_noalloc_scan_matches_scan(fl):
    L = fl.lib()
    assert L.fl_scan_noalloc_max_elements() == 1 << 31
    rng = np.random.default_rng(17)
    # large values so that the 64-bit total passes 2^32 on the first input
    for n in (5_000_011, 70_001):
        a = rng.integers(0, 1 << 12, n, dtype=np.uint32)
        if n > 1_000_000:
            a[:1000] = 0xFFFFFFF0
        d = fl.DeviceArray.from_host(a)
        print(f"noalloc scan: n = {n}", flush=True)
        ref, rtot = _scan_ref(fl, d, n)
        got, gtot = _scan_noalloc(fl, d, n)
        assert np.array_equal(got, ref) and gtot == rtot
        exp = np.zeros(n + 1, np.uint64)
        np.cumsum(a, dtype=np.uint64, out=exp[1:])
        assert gtot == int(exp[-1]) and np.array_equal(got, (exp[:-1] & 0xFFFFFFFF).astype(np.uint32))
        d.free()
    # n = 0: the total is written as 0, nothing else is touched
    d = fl.DeviceArray(1, np.uint32)
    _, t0 = _scan_noalloc(fl, d, 0)
    assert t0 == 0
    st = L.fl_scan_exclusive_u32_noalloc(d.p, d.p, None, (1 << 31) + 1, None)
    assert st == fl.FL_E_INVALID and "2^31" in L.fl_last_error_string().decode()
