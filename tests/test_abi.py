"""The C-ABI library loads, exports what include/tlbm_gpu.h declares, and
rejects bad input before touching a GPU (no compute calls here)."""
import ctypes

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
