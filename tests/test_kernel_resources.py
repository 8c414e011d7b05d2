"""Register budget of the lattice kernels (compiled here for gfx950, no GPU).

Every hot lattice kernel must run without scratch (no register spills), and
the collide-stream instances must keep at least four waves per SIMD: a spill
in the collide turns an HBM-bound kernel into a scratch-bound one (three to
six times slower), which no parity test notices.  The figures come from the
compiler's kernel resource-usage remarks for the same flags the library
build uses."""
import re
import subprocess

import pytest

from fluidloom_amd import build


@pytest.fixture(scope="module")
def usage(tmp_path_factory):
    out = tmp_path_factory.mktemp("res") / "lbm_kernels.o"
    src = build.CSRC / "lbm_kernels.hip"
    cmd = [build.HIPCC, *build.COMMON, "-c", str(src), "-o", str(out), "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    kernels = {}
    for block in re.split(r"remark: Function Name: ", r.stderr)[1:]:
        name = block.split()[0]

        def field(key):
            m = re.search(re.escape(key) + r": (\d+)", block)
            return int(m.group(1)) if m else None

        kernels[name] = dict(vgpr=field("VGPRs"), scratch=field("ScratchSize [bytes/lane]"),
                             occupancy=field("Occupancy [waves/SIMD]"))
    return kernels


def _lattice(usage, stem):
    return {k: v for k, v in usage.items() if stem in k}


@pytest.mark.parametrize("stem", ["k_collide", "k_fill_c2f", "k_fill_f2c", "k_capture"])
def test_no_scratch(usage, stem):
    ks = _lattice(usage, stem)
    assert ks, f"no {stem} instances found"
    spilled = {k: v for k, v in ks.items() if v["scratch"]}
    assert not spilled, spilled


def test_collide_occupancy(usage):
    ks = _lattice(usage, "k_collide")
    # 4 storage types x (21 launch variants + 3 pressure-outlet instances)
    assert len(ks) == 96
    low = {k: v for k, v in ks.items() if v["occupancy"] < 4}
    assert not low, low
    # the non-persistent FP32 / FP16 / FP16C forms (the ones the autotuner
    # picks on large levels) keep the full eight waves per SIMD
    full = {}
    for k, v in ks.items():
        m = re.search(r"k_collideI(.+?)Li(\d+)ELi(\d)E", k)
        assert m, k
        if m.group(1) != "d" and m.group(3) != "2":
            full[k] = v["occupancy"]
    assert len(full) == 63
    short = {k: o for k, o in full.items() if o < 8}
    assert not short, short


def test_device_code_keeps_fp32_denormals(tmp_path):
    """The FP16C codec maps its codes below 2^-13 onto FP32 denormals
    (fp16c_codec.hpp); a flush-to-zero build would silently zero them.  A
    kernel using the codec, built with the library's flags, must declare
    FP32 denormal mode 3 (inputs and outputs kept) in its descriptor."""
    src = tmp_path / "denorm.hip"
    src.write_text('#include "fp16c_codec.hpp"\n'
                   "__global__ void k(const unsigned short* c, float* f, unsigned short* e) {\n"
                   "    const unsigned i = threadIdx.x;\n"
                   "    f[i] = fl::lbm::fp16c_decode(c[i]);\n"
                   "    e[i] = fl::lbm::fp16c_encode_sr_scaled(f[i] * fl::lbm::kFp16cDown, i);\n"
                   "}\n")
    out = tmp_path / "denorm.s"
    cmd = [build.HIPCC, *build.COMMON, "--offload-device-only", "-S", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    modes = re.findall(r"\.amdhsa_float_denorm_mode_32\s+(\d+)", out.read_text())
    assert modes and all(m == "3" for m in modes), modes
