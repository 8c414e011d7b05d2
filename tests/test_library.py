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
