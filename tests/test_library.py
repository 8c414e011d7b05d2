"""CPU: the C-ABI library loads and exports every symbol the public header
declares (no device calls)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "fluidloom_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from fluidloom_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libfluidloom_hip.so not built (run `python -m fluidloom_amd.build`)")
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) <= set(_lib.exported_symbols())


def test_error_convention_without_device():
    from fluidloom_amd import _lib
    L = _lib.lib()
    assert L.fl_version().decode().startswith("fluidloom-hip")
    # invalid arguments are rejected before any device work
    st = L.fl_hilbert_encode(None, None, None, None, 0, None, 10, None)
    assert st == _lib.FL_E_INVALID
    assert "null" in L.fl_last_error_string().decode()
    assert L.fl_hash_capacity(700000) == 1 << 21
    # the allocation-free scan rejects sizes beyond its scratch array first
    assert L.fl_scan_noalloc_max_elements() == 1 << 31
    st = L.fl_scan_exclusive_u32_noalloc(None, None, None, (1 << 31) + 1, None)
    assert st == _lib.FL_E_INVALID and "2^31" in L.fl_last_error_string().decode()


def test_cell_map_size_bound_without_device():
    """The exact cell map needs load <= 0.6 inside its 2^30-slot ceiling:
    more than floor(0.6 * 2^30) cells are rejected before any device work
    (the pointers below are never dereferenced)."""
    import ctypes as C

    from fluidloom_amd import _lib
    L = _lib.lib()
    big = 644245095
    dummy = C.c_void_p(16)
    cells = _lib.Cells(*([dummy.value] * 7))
    cfg = _lib.RegridConfig(balance_mode=1, anchor_level=10, max_split_level=10, max_balance_iters=8,
                            root=(_lib.i32 * 3)(1, 1, 1))
    res = _lib.RegridResult()
    tb = L.fl_regrid_temp_bytes(big)
    st = L.fl_regrid(C.byref(cfg), C.byref(cells), big, C.byref(cells), big, None, None, dummy, tb, C.byref(res),
                     None)
    assert st == _lib.FL_E_INVALID and "cell map" in L.fl_last_error_string().decode()
    st = L.fl_balance(C.byref(cfg), C.byref(cells), big, dummy, tb, C.byref(res), None)
    assert st == _lib.FL_E_INVALID and "cell map" in L.fl_last_error_string().decode()
    root = (_lib.i32 * 3)(1, 1, 1)
    st = L.fl_face_neighbors(C.byref(cells), big, root, dummy, dummy, dummy, 1 << 30, dummy, None)
    assert st == _lib.FL_E_INVALID and "cell map" in L.fl_last_error_string().decode()


def test_plan_bytes_host_only(oracle):
    """fl_lbm_plan_bytes runs the layout on the host (no GPU needed) and
    reports the footprint a batch would allocate; mesh errors surface here."""
    from fluidloom_amd import _lib
    from fluidloom_amd.solver import make_params, plan_bytes
    from tests.lbm_cases import case
    cs = [case(n, oracle) for n in ("cavity", "cavity2", "sphere3")]

    def pm(c):
        nx, ny, nz = c["n0"]
        return make_params(nx, ny, nz, c["max_level"], c["tau0"], c["bc"], c["bc_u"], "f32", u_init=c["u_init"])

    def mesh(c):
        cc = c["cells"]
        return dict(x=cc.x, y=cc.y, z=cc.z, level=cc.level, state=cc.state)

    one = [plan_bytes([pm(c)], [mesh(c)]) for c in cs]
    both = plan_bytes([pm(c) for c in cs], [mesh(c) for c in cs])
    assert all(b > 0 for b in one)
    # per-instance tables add a little; the bulk is additive
    assert abs(both - sum(one)) < 0.01 * sum(one)
    p16 = make_params(8, 8, 8, 0, 0.6, ["wall"] * 6, [(0, 0, 0)] * 6, "f16")
    p32 = make_params(8, 8, 8, 0, 0.6, ["wall"] * 6, [(0, 0, 0)] * 6, "f32")
    m = mesh(cs[0])
    assert plan_bytes([p16], [m]) < plan_bytes([p32], [m])
    # leaves above the declared max_level
    c = cs[1]["cells"]
    lv = c.level.copy()
    bad = dict(x=c.x, y=c.y, z=c.z, level=lv, state=c.state)
    p = pm(cs[1])
    p.max_level = 0
    with pytest.raises(_lib.FluidLoomError):
        plan_bytes([p], [bad])
