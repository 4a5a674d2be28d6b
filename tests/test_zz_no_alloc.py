"""tlbm_gpu_reduce_device, the entry point that allocates nothing, against
tlbm_gpu_reduce on the same state: two calls in a row (the second on a
smaller lattice, whose reduction uses fewer of the shared partial sums),
and a state with a NaN population.

The output buffer is the caller's device memory, allocated here through the
HIP runtime the library itself links (the entry point never allocates)."""
import ctypes

import numpy as np
import pytest

from tlbm_amd import geometry
from tlbm_amd.gpu import ENONFINITE, Lattice, TlbmError, load_library

pytestmark = pytest.mark.gpu

SUMS = 7  # TLBM_REDUCE_SUMS


class DeviceBuffer:
    """SUMS doubles of device memory from the HIP runtime."""

    def __init__(self, hip):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(SUMS * 8)) == 0
        # a pattern no sum can equal, so a skipped write shows
        self.write(np.full(SUMS, -7.25))

    def write(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_size_t(SUMS * 8), 1) == 0  # host to device

    def read(self):
        out = np.empty(SUMS, np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                  ctypes.c_size_t(SUMS * 8), 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.fixture(scope="module")
def hip(gpu_lib):
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _both(lat, buf):
    """(tlbm_gpu_reduce's out[5] and return code, the device sums)."""
    L = load_library()
    out = np.zeros(5, np.float64)
    rc = L.tlbm_gpu_reduce(lat._h, out.ctypes.data)
    buf.write(np.full(SUMS, -7.25))
    assert L.tlbm_gpu_reduce_device(lat._h, buf.ptr) == 0
    lat.sync()
    return out, rc, buf.read()


def _agree(out, rc, dev, first_step):
    # the same kernels on the same state in the same order: bitwise equal
    assert np.array_equal(dev[:4], out[:4]), (dev, out)
    residual = np.sqrt(dev[4] / dev[5]) if (not first_step and dev[5] > 0) else 0.0
    assert residual == out[4], (residual, out[4])
    assert dev[5] > 0
    assert (dev[6] == 0) == (rc == 0)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "f32x"])
def test_reduce_device_matches_reduce_two_calls(gpu_lib, hip, storage):
    buf = DeviceBuffer(hip)
    try:
        # first call: a lattice whose reduction fills all 1024 partials
        nx, ny, nz = 64, 64, 96
        nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(32, 32, 30, 8)])
        with Lattice(19, nx, ny, nz, nd, storage=storage, omega=1.7, u_lbm=0.05, cs=0.1) as lat:
            out, rc, dev = _both(lat, buf)  # before any step: residual 0
            _agree(out, rc, dev, True)
            assert dev[4] == 0.0
            lat.step(40)
            out, rc, dev = _both(lat, buf)
            _agree(out, rc, dev, False)
            assert rc == 0 and dev[4] > 0
        # second call, smaller input: 8 partials of the same shared array
        nx, ny, nz = 12, 10, 16
        nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(5, 5, 6, 2.5)])
        with Lattice(15, nx, ny, nz, nd, storage=storage, omega=1.6, u_lbm=0.04) as lat:
            lat.step(7)
            out, rc, dev = _both(lat, buf)
            _agree(out, rc, dev, False)
            assert rc == 0
            # the mass is the small lattice's own, not a leftover of the first
            assert abs(dev[0] / (nx * ny * nz) - 1.0) < 0.05
        # and a D3Q27 lattice after them (its own partial kernel)
        nx, ny, nz = 24, 20, 32
        nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(12, 10, 12, 4)])
        with Lattice(27, nx, ny, nz, nd, storage=storage, omega=1.5, u_lbm=0.03) as lat:
            lat.step(5)
            out, rc, dev = _both(lat, buf)
            _agree(out, rc, dev, False)
            assert rc == 0
            assert abs(dev[0] / (nx * ny * nz) - 1.0) < 0.05
    finally:
        buf.free()


def test_reduce_device_counts_nonfinite(gpu_lib, hip):
    buf = DeviceBuffer(hip)
    try:
        nx, ny, nz = 16, 12, 20
        nd = geometry.channel(nx, ny, nz)
        with Lattice(19, nx, ny, nz, nd, storage="f32", omega=1.6, u_lbm=0.05) as lat:
            lat.step(3)
            f = lat.populations()
            f[100, 3] = np.inf
            f[200, 5] = np.nan
            lat.set_populations(f)
            out, rc, dev = _both(lat, buf)
            assert rc == ENONFINITE and dev[6] == 2.0
            with pytest.raises(TlbmError):
                lat.reduce()
    finally:
        buf.free()


def test_reduce_device_rejects_null_output(gpu_lib):
    nd = geometry.channel(8, 8, 8)
    with Lattice(19, 8, 8, 8, nd, omega=1.5) as lat:
        assert load_library().tlbm_gpu_reduce_device(lat._h, None) == -1
