"""Register use of the built gfx950 kernels, read from the code objects'
metadata notes in stencilfft_amd/libstencilfft.so.

The fused-path kernels (the specialised line pass, the tiled repack, the
strip Laplacian) must not spill: a spill puts scratch traffic on the HBM
path and, for the FP64 line pass at N = 8192, cost 20% of its time.  The
kernels' register budgets depend on compiler behaviour (for example whether
loop-invariant addresses stay live across the group loop), so this guards
against a toolchain or source change that brings spills back.
"""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "stencilfft_amd", "libstencilfft.so")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"

# Kernels off the fused path that spill a little at the 256-VGPR limit: the
# generic line kernel and the two-stage pass, both FP64 forward with compute.
KNOWN = {"zm::line_kernel<double, true, true>": 2, "zm::stage_b_kernel<double, true, true>": 2}


def _tools():
    tools = {"objcopy": shutil.which("objcopy"),
             "bundler": os.path.join(LLVM, "clang-offload-bundler"),
             "readelf": os.path.join(LLVM, "llvm-readelf"),
             "demangle": shutil.which("c++filt")}
    if not all(t and os.path.exists(t) for t in tools.values()):
        return None
    return tools


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    tools = _tools()
    if tools is None or not os.path.exists(LIB):
        pytest.skip("needs the built library and the ROCm LLVM tools")
    d = tmp_path_factory.mktemp("co")
    fat = d / "fat.bin"
    subprocess.run([tools["objcopy"], f"--dump-section=.hip_fatbin={fat}", LIB, str(d / "lib.so")], check=True,
                   capture_output=True)
    data = fat.read_bytes()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    offs = [m.start() for m in re.finditer(re.escape(magic), data)] + [len(data)]
    raw = {}
    for n, (a, b) in enumerate(zip(offs, offs[1:])):
        part, co = d / f"b{n}.bin", d / f"co{n}.o"
        part.write_bytes(data[a:b])
        subprocess.run([tools["bundler"], "--unbundle", "--type=o", f"--input={part}", f"--targets={TARGET}",
                        f"--output={co}"], check=True, capture_output=True)
        notes = subprocess.run([tools["readelf"], "--notes", str(co)], check=True, capture_output=True,
                               text=True).stdout
        for entry in re.split(r"\n\s+- \.agpr_count", notes):
            nm = re.search(r"\.name:\s+(_Z\S+)", entry)
            if not nm:
                continue
            raw[nm.group(1)] = {k: int(v) for k, v in re.findall(
                r"\.(vgpr_spill_count|sgpr_spill_count|vgpr_count|private_segment_fixed_size):\s+(\d+)", entry)}
    assert raw, "no kernel metadata found"
    mangled = list(raw)
    names = subprocess.run([tools["demangle"]], input="\n".join(mangled), check=True, capture_output=True,
                           text=True).stdout.splitlines()
    assert len(names) == len(mangled)
    return {n.split("(")[0].replace("void ", ""): raw[m] for n, m in zip(names, mangled)}


def test_every_object_targets_gfx950_and_holds_the_hot_kernels(kernels):
    fams = {k.split("<")[0] for k in kernels}
    for f in ("zm::line_fast_kernel", "zm::repack_tiled_kernel", "zm::laplace_strip_kernel", "zm::deriv4_march_kernel",
              "zm::laplace_zmarch_kernel"):
        assert f in fams, f


def test_fused_path_kernels_do_not_spill(kernels):
    spilled = {k: m for k, m in kernels.items() if m.get("vgpr_spill_count", 0) or m.get("sgpr_spill_count", 0)}
    unexpected = {k: m for k, m in spilled.items() if KNOWN.get(k, -1) < m.get("vgpr_spill_count", 0)
                  or m.get("sgpr_spill_count", 0)}
    assert not unexpected, unexpected


def test_no_kernel_keeps_arrays_in_scratch(kernels):
    # A private segment without spills means a per-thread array the compiler
    # left in scratch memory: its accesses go through the cache hierarchy to
    # HBM and each load waits before the next step.  The double2 depth-march
    # Laplacian once held its halo slots there (80-336 B per thread), which
    # wrote 1.375x and read 1.64x the field at 512^3 FP64.  Only the two known
    # spilling kernels may use scratch, for their spill slots.
    scratch = {k: m["private_segment_fixed_size"] for k, m in kernels.items()
               if m.get("private_segment_fixed_size", 0) and k not in KNOWN}
    assert not scratch, scratch


def test_fp64_line_pass_at_n8192_fits_without_spills(kernels):
    # N = 8192 = VEC * M in both orders, forward and inverse, both table forms
    hot = [k for k in kernels if k.startswith("zm::line_fast_kernel<double, 6, 7, 0,")
           or k.startswith("zm::line_fast_kernel<double, 7, 6, 0,")]
    assert len(hot) == 8, hot
    for k in hot:
        assert kernels[k]["vgpr_spill_count"] == 0 and kernels[k]["vgpr_count"] <= 256, (k, kernels[k])
