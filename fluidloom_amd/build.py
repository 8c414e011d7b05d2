"""Build libfluidloom_hip.so (every HIP source under csrc/) for gfx950.

The library is built in-tree into fluidloom_amd/_lib/ with hipcc; objects go
to fluidloom_amd/_build/.  `python -m fluidloom_amd.build` rebuilds.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libfluidloom_hip.so"
ARCH = os.environ.get("FL_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # keep packed FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fuse multiply-adds only within one source expression, so that every
    # launch variant of a kernel rounds the same way (cross-statement fusion
    # depends on the code shape around it)
    "-ffp-contract=on",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))



def _compile(src: Path) -> Path:
    out = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
    return out


def build_cpp_tests(verbose: bool = True) -> Path:
    """The C++ API test program (tests/cpp/test_api.cpp), linked against the
    library and the oracle (its checker)."""
    src = ROOT / "tests" / "cpp" / "test_api.cpp"
    out_dir = ROOT / "tests" / "cpp" / "_build"
    out_dir.mkdir(exist_ok=True)
    out = out_dir / "test_api"
    oracle_dir = ROOT / "oracle" / "_build"
    cmd = [HIPCC, "-std=c++17", "-O2", f"-I{ROOT / 'include'}", str(src), "-o", str(out),
           f"-L{LIBDIR}", "-lfluidloom_hip", f"-Wl,-rpath,{LIBDIR}", f"-L{oracle_dir}", "-loracle",
           f"-Wl,-rpath,{oracle_dir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ API tests failed to build:\n{r.stderr}")
    if verbose:
        print(f"built {out}", flush=True)
    return out


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    from .lattice import generate
    generate()  # lattice constants header from the lattice declarations
    srcs = sources()
    jobs = jobs or min(8, len(srcs), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
