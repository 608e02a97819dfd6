"""The built library's device code holds no packed-FP32 arithmetic
(v_pk_add_f32, v_pk_mul_f32, v_pk_fma_f32).

gpu-vah_amd/Makefile compiles every HIP translation unit with
``-Xclang -target-feature -Xclang -packed-fp32-ops``.  The host pass of hipcc
prints that it does not know the feature; the device pass honours it.  This
test reads the gfx950 code objects of the library built here (unbundled into a
temporary directory) and checks the instruction stream itself, so that a build
change that drops the flag fails here rather than on the device."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpu-vah_amd", "libvah.so")
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"


@pytest.mark.skipif(not os.path.exists(OBJDUMP), reason="ROCm llvm-objdump not installed")
def test_no_packed_fp32_instructions(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("libvah.so not built")
    lib = tmp_path / "libvah.so"
    shutil.copy(LIB, lib)
    # --offloading writes <lib>.<k>.<triple> next to its input
    subprocess.run([OBJDUMP, "--offloading", str(lib)], check=True, capture_output=True, cwd=tmp_path)
    objs = sorted(p for p in tmp_path.iterdir() if "amdgcn-amd-amdhsa" in p.name)
    assert objs, "no gfx950 code object in libvah.so"
    total_f64 = 0
    for obj in objs:
        asm = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", str(obj)], check=True, capture_output=True,
                             text=True).stdout
        packed = [ln for ln in asm.splitlines() if any(f"v_pk_{op}_f32" in ln for op in ("add", "mul", "fma"))]
        assert not packed, f"{obj.name}: {len(packed)} packed-FP32 instructions, e.g. {packed[0].strip()}"
        total_f64 += asm.count("v_fma_f64")
    assert total_f64 > 1000  # the disassembly really holds the stage kernels
