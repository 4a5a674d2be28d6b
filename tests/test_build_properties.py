"""Properties of the built gfx950 code objects, read from their metadata.

The kernels' register budgets are chosen (workgroup sizes, the closure
collide's waves per SIMD); a change that makes a kernel spill to scratch or
outgrow the occupancy it is allocated for shows up here, on the CPU, before
any GPU run.  The code objects are the device parts of the objects that
`make -C tlbm_amd` builds (tlbm_amd/_build/*.o): the offload bundle in each
object's .hip_fatbin section is unbundled with clang-offload-bundler and its
AMDGPU metadata note read with llvm-readelf.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tlbm_amd", "_build")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"
OBJECTS = ("kernels_f64", "kernels_f32", "kernels_f32x", "mlp")


def _tool(name):
    path = os.path.join(LLVM, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} is not installed")
    return path


def _kernels(obj):
    """{kernel symbol: metadata fields} of one built object's gfx950 code object."""
    src = os.path.join(BUILD, obj + ".o")
    if not os.path.exists(src):
        pytest.skip(f"{src} is not built (make -C tlbm_amd all)")
    tmp = tempfile.mkdtemp()
    try:
        bundle, co = os.path.join(tmp, "b"), os.path.join(tmp, "co")
        subprocess.run([_tool("llvm-objcopy"), f"--dump-section=.hip_fatbin={bundle}", src,
                        os.path.join(tmp, "host.o")], check=True, capture_output=True)
        subprocess.run([_tool("clang-offload-bundler"), "--unbundle", "--type=o", f"--input={bundle}",
                        f"--targets={TARGET}", f"--output={co}"], check=True, capture_output=True)
        notes = subprocess.run([_tool("llvm-readelf"), "--notes", co], check=True,
                               capture_output=True, text=True).stdout
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out, cur = {}, None
    for line in notes.splitlines():
        m = re.match(r"\s+\.name:\s+(\S+)$", line)
        if m and m.group(1).startswith("_Z"):
            cur = out.setdefault(m.group(1), {})
            continue
        m = re.match(r"\s+\.(private_segment_fixed_size|vgpr_count|agpr_count|vgpr_spill_count):\s+(\d+)$", line)
        if m and cur is not None:
            cur[m.group(1)] = int(m.group(2))
    return out


@pytest.fixture(scope="module")
def kernels():
    k = {}
    for obj in OBJECTS:
        k.update(_kernels(obj))
    return k


def test_every_kernel_is_present(kernels):
    names = " ".join(kernels)
    for lattice in ("D3Q15", "D3Q19", "D3Q27"):
        assert re.search(rf"k_lbm_closureINS_5{lattice}", names), lattice
        assert re.search(rf"k_lbmINS_5{lattice}", names), lattice
    assert "k_mlp_grad" in names and "k_reduce_partial" in names


def test_no_kernel_uses_scratch(kernels):
    """No spills and no stack: every kernel keeps its state in registers."""
    spilling = {k: v for k, v in kernels.items()
                if v.get("private_segment_fixed_size", 0) or v.get("vgpr_spill_count", 0)}
    assert not spilling, sorted(spilling)[:5]


def _waves(vgprs, agprs):
    """Wavefronts per SIMD that gfx950's 512 registers per lane hold
    (allocated in blocks of 8, VGPRs and AGPRs together)."""
    regs = -(-vgprs // 8) * 8 + -(-agprs // 8) * 8
    return min(8, 512 // regs)


@pytest.mark.parametrize("lattice,storage,waves", [
    ("D3Q15", 1, 6), ("D3Q19", 1, 5), ("D3Q27", 1, 4),
    ("D3Q15", 2, 6), ("D3Q19", 2, 6), ("D3Q27", 2, 5)])
def test_closure_collide_holds_its_waves(kernels, lattice, storage, waves):
    """The closure collide runs the occupancy it is register-allocated for
    (the waves per SIMD that its step time was measured at)."""
    name = [k for k in kernels if re.match(rf"_ZN4tlbm13k_lbm_closureINS_5{lattice}ELi{storage}E", k)]
    assert len(name) == 1, name
    meta = kernels[name[0]]
    assert _waves(meta["vgpr_count"], meta.get("agpr_count", 0)) >= waves, meta


@pytest.mark.parametrize("storage,waves", [(0, 4), (1, 5), (2, 8), (3, 4)])
def test_d3q19_step_holds_its_waves(kernels, storage, waves):
    """The D3Q19 step kernels run the occupancy they were measured at
    (FP64 4 waves per SIMD, one 1024-thread workgroup per CU; FP32 5; FP32X,
    whose arithmetic is FP64, 4; FP16 8)."""
    name = [k for k in kernels if re.match(rf"_ZN4tlbm5k_lbmINS_5D3Q19ELi{storage}ELi0ELb0E", k)]
    assert len(name) == 1, name
    meta = kernels[name[0]]
    assert _waves(meta["vgpr_count"], meta.get("agpr_count", 0)) >= waves, meta
