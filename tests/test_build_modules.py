"""Kernels built as code objects of their own (linqt_amd/csrc/modules).

The build compiles each module source to LLVM IR, gives the kernel named in
its `lq-module-kernel:` line the attributes that line asks for, links the code
object and embeds it in liblinqt_amd.so.  These tests check the IR rewrite on
a small hand-written module, that the library carries the code object, and
that the code object has the register split its marker line asks for.  No GPU.
"""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from linqt_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_IR = """define amdgpu_kernel void @k_one(ptr addrspace(1) %0) local_unnamed_addr #0 {
  ret void
}

define amdgpu_kernel void @k_two(ptr addrspace(1) %0) local_unnamed_addr #0 {
  ret void
}

define internal void @helper() #1 {
  ret void
}

attributes #0 = { convergent "amdgpu-flat-work-group-size"="1,512" "target-cpu"="gfx950" }
attributes #1 = { nounwind }
"""


def test_attribute_goes_to_a_copy_of_the_kernels_group_only():
    out = build._with_fn_attr(_IR, "k_two", '"amdgpu-agpr-alloc"="192"')
    groups = dict(re.findall(r"^attributes #(\d+) = \{ (.*) \}$", out, re.M))
    assert set(groups) == {"0", "1", "2"}
    assert groups["2"].startswith('"amdgpu-agpr-alloc"="192" ')
    assert groups["2"].endswith(groups["0"])  # the rest of the original group is kept
    assert "amdgpu-agpr-alloc" not in groups["0"]  # k_one shares #0 and keeps it unchanged
    assert re.search(r"@k_two\([^\n]*\) local_unnamed_addr #2 \{", out)
    assert re.search(r"@k_one\([^\n]*\) local_unnamed_addr #0 \{", out)
    assert re.search(r"@helper\(\) #1 \{", out)
    with pytest.raises(RuntimeError):
        build._with_fn_attr(_IR, "k_missing", '"x"="1"')


def test_every_module_source_names_its_kernel_and_split():
    mods = sorted(f for f in os.listdir(build.MODDIR) if f.endswith(".hip"))
    assert "lq_contract_t128.hip" in mods
    for f in mods:
        marks = build._MARK.findall(open(os.path.join(build.MODDIR, f)).read())
        assert marks, f
        for kernel, agpr in marks:
            assert 0 < int(agpr) <= 256 and int(agpr) % 4 == 0, (f, kernel, agpr)
    src = open(os.path.join(build.MODDIR, "lq_contract_t128.hip")).read()
    assert build._MARK.findall(src) == [("lq_contract_f64_3m_t128", "192")]
    assert 'extern "C" __global__' in src and "lq_contract_f64_3m_t128(" in src


def _embedded_code_object():
    from linqt_amd import _lib

    L = _lib.load()
    size = ctypes.c_size_t.in_dll(L, "lq_module_contract_t128_size").value
    data = bytes((ctypes.c_ubyte * size).in_dll(L, "lq_module_contract_t128"))
    return data


def test_library_embeds_the_contraction_code_object():
    data = _embedded_code_object()
    assert len(data) > 4096
    assert data[:4] == b"\x7fELF"
    assert data[18] == 224  # e_machine EM_AMDGPU (0xe0)
    assert b"lq_contract_f64_3m_t128" in data


def test_code_object_has_the_requested_register_split(tmp_path):
    tool = None
    for root in (os.environ.get("ROCM_PATH"), "/opt/rocm"):
        if root and os.path.exists(os.path.join(root, "llvm", "bin", "llvm-readelf")):
            tool = os.path.join(root, "llvm", "bin", "llvm-readelf")
            break
    if tool is None:
        tool = shutil.which("llvm-readelf")
    if tool is None:
        pytest.skip("llvm-readelf not available")
    path = tmp_path / "t128.hsaco"
    path.write_bytes(_embedded_code_object())
    notes = subprocess.run([tool, "--notes", str(path)], capture_output=True, text=True, check=True).stdout
    agpr = re.findall(r"\.agpr_count:\s+(\d+)", notes)
    vgpr = re.findall(r"\.vgpr_count:\s+(\d+)", notes)
    spill = re.findall(r"\.vgpr_spill_count:\s+(\d+)", notes)
    # .vgpr_count is the wave's whole vector register allocation (VGPRs + AGPRs):
    # 64 + 192 = 256, i.e. two waves per SIMD, with nothing spilled
    assert agpr == ["192"] and vgpr == ["256"] and spill == ["0"], (agpr, vgpr, spill)
