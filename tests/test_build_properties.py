"""Properties of the built gfx950 code objects, read from their metadata
and their disassembly.

The kernels' register budgets are chosen (workgroup sizes, the closure
collide's waves per SIMD); a change that makes a kernel spill to scratch or
outgrow the occupancy it is allocated for shows up here, on the CPU, before
any GPU run.  So does a change that undoes a measured instruction order: the
kernel arguments loaded in one batch, the FP16 step's and the closure's
node-type load ahead of the gathers, the closure's direct global-to-LDS
weight staging, and the wave priority raised only where it paid.  The code objects are the device parts of the objects that
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


def _code_object_text(obj, tool, *args):
    """Output of an LLVM tool on one built object's gfx950 code object."""
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
        return subprocess.run([_tool(tool), *args, co], check=True, capture_output=True, text=True).stdout
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def _kernels(obj):
    """{kernel symbol: metadata fields} of one built object's gfx950 code object."""
    notes = _code_object_text(obj, "llvm-readelf", "--notes")
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


def _disassembly(obj):
    """{kernel symbol: its instructions in order, comments stripped}."""
    out, cur = {}, None
    for line in _code_object_text(obj, "llvm-objdump", "-d", "--no-show-raw-insn").splitlines():
        m = re.match(r"^[0-9a-f]+ <(\S+)>:", line)
        if m:
            cur = out.setdefault(m.group(1), [])
            continue
        if cur is not None and line.strip():
            cur.append(line.split("//")[0].strip())
    return out


@pytest.fixture(scope="module")
def code():
    k = {}
    for obj in OBJECTS:
        k.update(_disassembly(obj))
    return k


def _step(code, lattice, storage):
    name = [k for k in code if re.match(rf"_ZN4tlbm5k_lbmINS_5{lattice}ELi{storage}ELi0ELb", k)]
    assert len(name) == 1, name
    return code[name[0]]


def _closure(code, lattice, storage):
    name = [k for k in code if re.match(rf"_ZN4tlbm13k_lbm_closureINS_5{lattice}ELi{storage}E", k)]
    assert len(name) == 1, name
    return code[name[0]]


def _first(ins, pattern):
    """Index of the first instruction matching pattern (len(ins) if none does)."""
    return next((i for i, x in enumerate(ins) if re.match(pattern, x)), len(ins))


GATHER = r"buffer_load_(ushort|dword|dwordx2) "


@pytest.mark.parametrize("lattice", ["D3Q15", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("storage", [0, 1, 2, 3])
def test_step_prologue_waits_once_for_its_arguments(code, lattice, storage):
    """The step's kernel arguments arrive in one batch of scalar loads: at
    most two scalar-memory waits before the first gather (the FP16 step's
    second is for its node-type pointers; unbatched there were five)."""
    ins = _step(code, lattice, storage)
    waits = [x for x in ins[:_first(ins, GATHER)] if x.startswith("s_waitcnt") and "lgkmcnt" in x]
    assert len(waits) <= 2, waits


@pytest.mark.parametrize("lattice", ["D3Q15", "D3Q19", "D3Q27"])
def test_fp16_step_loads_node_type_ahead_of_gathers(code, lattice):
    ins = _step(code, lattice, 2)
    assert _first(ins, r"global_load_ubyte ") < _first(ins, GATHER)


@pytest.mark.parametrize("lattice", ["D3Q15", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("storage", [1, 2])
def test_closure_stages_weights_by_direct_lds_loads(code, lattice, storage):
    """The closure collide stages the network's weights with
    global_load_lds_dwordx4 ahead of its gathers, loads its node type ahead
    of them too, and has no register round trip for the weights."""
    ins = _closure(code, lattice, storage)
    first_gather = _first(ins, GATHER)
    assert _first(ins, r"global_load_lds_dwordx4 ") < first_gather
    assert _first(ins, r"global_load_ubyte ") < first_gather
    assert not any(x.startswith("ds_write") for x in ins[:first_gather])


@pytest.mark.parametrize("lattice,storage,raised", [
    ("D3Q27", 0, True), ("D3Q27", 1, True), ("D3Q27", 2, True), ("D3Q27", 3, True), ("D3Q19", 2, True),
    ("D3Q19", 0, False), ("D3Q19", 1, False), ("D3Q19", 3, False),
    ("D3Q15", 0, False), ("D3Q15", 1, False), ("D3Q15", 2, False), ("D3Q15", 3, False)])
def test_step_raises_wave_priority_where_measured(code, lattice, storage, raised):
    """s_setprio 3 while the gathers go out, exactly where it measured faster."""
    ins = _step(code, lattice, storage)
    assert ("s_setprio 3" in ins) == raised
