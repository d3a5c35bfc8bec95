"""The built gfx950 code holds no packed FP32 VALU instructions (every HIP
translation unit is compiled with the packed-fp32-ops feature off), the
convolution library runs on bf16 MFMA, and both HIP builds carry the flags
that make them reproducible: gfx950 only and FMA contraction within one
expression (-ffp-contract=on)."""
import os
import re
import shutil
import subprocess

import pytest

from fingering_dynamics_amd import build

OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"
PACKED_FP32 = re.compile(r"\bv_pk_(fma|mul|add)_f32\b")


def test_hip_flags():
    f = build.HIP_FLAGS
    assert "--offload-arch=gfx950" in f and "-ffp-contract=on" in f
    i = f.index("-packed-fp32-ops")
    assert f[i - 3:i] == ["-Xclang", "-target-feature", "-Xclang"]
    assert not any(a.startswith("--offload-arch=") and a != "--offload-arch=gfx950" for a in f)


def _disassemble(lib, tmp_path):
    if not os.path.exists(OBJDUMP):
        pytest.skip("llvm-objdump not available")
    if not os.path.exists(lib):
        pytest.skip("%s is not built" % lib)
    local = str(tmp_path / os.path.basename(lib))
    shutil.copy(lib, local)            # the bundles are extracted next to the input file
    subprocess.run([OBJDUMP, "--offloading", local], check=True, capture_output=True, cwd=str(tmp_path))
    cos = [p for p in os.listdir(tmp_path) if "amdgcn-amd-amdhsa" in p]
    assert len(cos) == 1 and cos[0].endswith("gfx950"), cos
    out = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", str(tmp_path / cos[0])], check=True,
                         capture_output=True, text=True).stdout
    assert len(out.splitlines()) > 1000
    return out


@pytest.mark.parametrize("name", ["libfdlbm.so", "libfdsurr.so"])
def test_no_packed_fp32_in_the_built_code(tmp_path, name):
    asm = _disassemble(os.path.join(build.LIBDIR, name), tmp_path)
    assert not PACKED_FP32.search(asm)
    if name == "libfdsurr.so":
        assert "v_mfma_f32_16x16x32_bf16" in asm
