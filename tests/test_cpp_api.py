"""The C++ drop-in API (IBackend/HIPBackend, SOAFieldManager, hash map,
adaptation engines, HaloExchanger) exercised by tests/cpp/test_api.cpp,
which mirrors the reference's gtest suites."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "_build" / "test_api"


def _run(mode):
    if not BIN.exists():
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    r = subprocess.run([str(BIN), mode], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_api_host_layer():
    _run("--mock")


@pytest.mark.gpu
def test_cpp_api_on_device(fl):
    _run("--hip")
