"""The C-ABI library loads, exports what include/tlbm_gpu.h declares, and
rejects bad input before touching a GPU (no compute calls here)."""
import ctypes
import os

import numpy as np
import pytest

from tlbm_amd import gpu


def test_library_exports_every_declared_symbol(gpu_lib):
    names = gpu.declared_symbols()
    assert len(names) >= 13
    for n in names:
        assert hasattr(gpu_lib, n), n


@pytest.mark.parametrize("bad", [
    dict(lattice=17), dict(nx=0), dict(storage=7), dict(omega=0.0), dict(reserved=1),
    dict(u_lbm=1.0),
])
def test_create_rejects_bad_parameters(gpu_lib, bad):
    kw = dict(lattice=19, dynamics=1, nx=4, ny=4, nz=4, storage=0, exact_weights=0,
              reserved=0, cs=0.0, rho_lbm=1.0, u_lbm=0.05, omega=1.2)
    kw.update(bad)
    p = gpu.Params(**kw)
    h = ctypes.c_void_p()
    nd = np.zeros(max(1, 4 * 4 * 4), np.uint8)
    rc = gpu_lib.tlbm_gpu_create(ctypes.byref(p), nd.ctypes.data, ctypes.byref(h))
    assert rc == -1 and not h.value
    assert gpu_lib.tlbm_gpu_last_error()


def test_create_rejects_unknown_node_type():
    nd = np.zeros(64, np.uint8)
    nd[10] = 4
    with pytest.raises(gpu.TlbmError, match="node 10 has type 4"):
        gpu.Lattice(19, 4, 4, 4, nd, omega=1.2)


def test_create_rejects_type5_before_its_uy_sources():
    nd = np.zeros(64, np.uint8)
    nd[3] = 5
    with pytest.raises(gpu.TlbmError, match="type-5 node 3"):
        gpu.Lattice(19, 4, 4, 4, nd, omega=1.2)


def test_create_rejects_short_node_array():
    with pytest.raises(ValueError):
        gpu.Lattice(19, 4, 4, 4, np.zeros(63, np.uint8), omega=1.2)


def test_params_struct_layout_matches_header():
    """The ctypes mirror of tlbm_params (tlbm_amd/gpu.py) has the size and
    field offsets the C compiler gives the header's struct."""
    import lattice_tables
    fields = [f for f, _ in gpu.Params._fields_]
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"tlbm_gpu.h\"\nint main(void) {\n"
    src += '    printf("size %zu\\n", sizeof(tlbm_params));\n'
    for f in fields:
        src += f'    printf("{f} %zu %zu\\n", offsetof(tlbm_params, {f}), sizeof(((tlbm_params *)0)->{f}));\n'
    src += "    return 0;\n}\n"
    out = lattice_tables._run_c_probe(src, ["-I", os.path.join(lattice_tables.ROOT, "include")], [])
    rows = dict((ln.split()[0], ln.split()[1:]) for ln in out.strip().splitlines())
    assert int(rows["size"][0]) == ctypes.sizeof(gpu.Params)
    for f in fields:
        off, size = (int(v) for v in rows[f])
        assert getattr(gpu.Params, f).offset == off, f
        assert getattr(gpu.Params, f).size == size, f
    # every field of the header's struct is mirrored
    header = open(gpu.HEADER_PATH).read()
    body = header[header.index("typedef struct tlbm_params {"):header.index("} tlbm_params;")]
    import re
    declared = re.findall(r"\b(?:int|double)\s+([a-z_, ]+);", body)
    names = [n.strip() for group in declared for n in group.split(",")]
    assert names == fields
