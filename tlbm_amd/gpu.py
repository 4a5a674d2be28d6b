"""ctypes binding of libtlbm_gpu.so (C-ABI: include/tlbm_gpu.h).

This is the binding a Python caller of the reference's time step would use:
`Lattice` wraps one tlbm_gpu handle.  It loads the in-tree HIP library and
raises if that library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtlbm_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "tlbm_gpu.h")

STORAGE = {"f64": 0, "f32": 1, "f16": 2, "f32x": 3}
BYTES_PER_VALUE = {"f64": 8, "f32": 4, "f16": 2, "f32x": 4}
LATTICE_Q = {"D3Q15": 15, "D3Q19": 19, "D3Q27": 27}
ENONFINITE = -5  # TLBM_ENONFINITE: the state holds NaN or infinite populations


class TlbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tlbm_gpu error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int), ("dynamics", ctypes.c_int), ("nx", ctypes.c_int),
        ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("storage", ctypes.c_int),
        ("exact_weights", ctypes.c_int), ("reserved", ctypes.c_int), ("cs", ctypes.c_double),
        ("rho_lbm", ctypes.c_double), ("u_lbm", ctypes.c_double), ("omega", ctypes.c_double),
    ]


_lib = None
_P = ctypes.c_void_p


def load_library(path: str = LIB_PATH):
    """Load libtlbm_gpu.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `make -C tlbm_amd lib` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    L.tlbm_gpu_create.argtypes = [ctypes.POINTER(Params), _P, ctypes.POINTER(_P)]
    L.tlbm_gpu_destroy.argtypes = [_P]
    L.tlbm_gpu_step.argtypes = [_P, ctypes.c_int]
    L.tlbm_gpu_step_timed.argtypes = [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.tlbm_gpu_sync.argtypes = [_P]
    L.tlbm_gpu_get_macros.argtypes = [_P, _P, _P, _P, _P]
    L.tlbm_gpu_get_forces.argtypes = [_P, _P, _P, _P]
    L.tlbm_gpu_get_populations.argtypes = [_P, _P]
    L.tlbm_gpu_set_populations.argtypes = [_P, _P]
    L.tlbm_gpu_init_from_macros.argtypes = [_P, _P, _P, _P, _P]
    L.tlbm_gpu_reduce.argtypes = [_P, _P]
    L.tlbm_gpu_reduce_device.argtypes = [_P, _P]
    L.tlbm_gpu_memory_bytes.argtypes = [_P, ctypes.POINTER(ctypes.c_uint64)]
    L.tlbm_gpu_time_average.argtypes = [_P, ctypes.c_int]
    L.tlbm_gpu_get_time_average.argtypes = [_P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_int64)]
    L.tlbm_gpu_set_closure.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double]
    L.tlbm_gpu_box_filter.argtypes = [_P, _P]
    L.tlbm_gpu_set_inlet_profile.argtypes = [_P, _P]
    L.tlbm_mlp_trainer_create.argtypes = [_P, _P, ctypes.c_int64, _P, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.POINTER(_P)]
    L.tlbm_mlp_trainer_step.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        ctypes.POINTER(ctypes.c_double)]
    L.tlbm_mlp_trainer_loss_grad.argtypes = [_P, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.POINTER(ctypes.c_double), _P]
    L.tlbm_mlp_trainer_params.argtypes = [_P, _P]
    L.tlbm_mlp_trainer_destroy.argtypes = [_P]
    L.tlbm_mlp_infer.argtypes = [_P, _P, ctypes.c_int64, _P]
    for name in ("tlbm_gpu_set_closure", "tlbm_gpu_box_filter", "tlbm_gpu_set_inlet_profile",
                 "tlbm_mlp_trainer_create", "tlbm_mlp_trainer_step", "tlbm_mlp_trainer_loss_grad",
                 "tlbm_mlp_trainer_params", "tlbm_mlp_trainer_destroy", "tlbm_mlp_infer"):
        getattr(L, name).restype = ctypes.c_int
    L.tlbm_gpu_last_error.restype = ctypes.c_char_p
    L.tlbm_gpu_last_error.argtypes = []
    for name in ("create", "destroy", "step", "step_timed", "sync", "get_macros", "get_forces",
                 "get_populations", "set_populations", "init_from_macros", "reduce", "reduce_device",
                 "memory_bytes", "time_average", "get_time_average"):
        getattr(L, f"tlbm_gpu_{name}").restype = ctypes.c_int
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise TlbmError(rc, load_library().tlbm_gpu_last_error().decode())


def _f64(a, n: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


FILTER_PLANES = ("rho", "ux", "uy", "uz", "uxux", "uyuy", "uzuz", "uxuy", "uxuz", "uyuz",
                 "sxx", "syy", "szz", "sxy", "sxz", "syz")

class Lattice:
    """One lattice resident on the GPU: the reference's TLBM_Partition time
    step for a single partition, driven through the C-ABI."""

    def __init__(self, lattice: int, nx: int, ny: int, nz: int, nd_type, *, omega: float,
                 u_lbm: float = 0.0, rho_lbm: float = 1.0, cs: float = 0.0, dynamics: int = 1,
                 storage: str = "f64", exact_weights: bool = False):
        L = load_library()
        if storage not in STORAGE:
            raise ValueError(f"storage must be one of {sorted(STORAGE)}")
        self.Q = int(lattice)
        self.shape = (int(nx), int(ny), int(nz))
        self.N = int(nx) * int(ny) * int(nz)
        self.storage = storage
        nd = np.ascontiguousarray(np.asarray(nd_type).ravel())
        if nd.size != self.N:
            raise ValueError(f"ndType has {nd.size} entries, lattice has {self.N}")
        if nd.min(initial=0) < 0 or nd.max(initial=0) > 255:
            raise ValueError("node types must fit in uint8")
        nd = nd.astype(np.uint8)
        p = Params(int(lattice), int(dynamics), int(nx), int(ny), int(nz), STORAGE[storage],
                   int(bool(exact_weights)), 0, float(cs), float(rho_lbm), float(u_lbm),
                   float(omega))
        h = _P()
        _check(L.tlbm_gpu_create(ctypes.byref(p), nd.ctypes.data, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            load_library().tlbm_gpu_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n: int = 1) -> None:
        _check(load_library().tlbm_gpu_step(self._h, int(n)))

    def step_timed(self, n: int) -> float:
        ms = ctypes.c_double()
        _check(load_library().tlbm_gpu_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def sync(self) -> None:
        _check(load_library().tlbm_gpu_sync(self._h))

    def macros(self) -> dict:
        out = [np.empty(self.N, np.float64) for _ in range(4)]
        _check(load_library().tlbm_gpu_get_macros(self._h, *[a.ctypes.data for a in out]))
        return dict(zip(("ux", "uy", "uz", "rho"), out))

    def forces(self) -> dict:
        out = [np.empty(self.N, np.float64) for _ in range(3)]
        _check(load_library().tlbm_gpu_get_forces(self._h, *[a.ctypes.data for a in out]))
        return dict(zip(("Fx", "Fy", "Fz"), out))

    def populations(self) -> np.ndarray:
        f = np.empty(self.N * self.Q, np.float64)
        _check(load_library().tlbm_gpu_get_populations(self._h, f.ctypes.data))
        return f.reshape(self.N, self.Q)

    def set_populations(self, f) -> None:
        a = _f64(f, self.N * self.Q)
        _check(load_library().tlbm_gpu_set_populations(self._h, a.ctypes.data))

    def init_from_macros(self, ux, uy, uz, rho) -> None:
        arrs = [_f64(v, self.N) for v in (ux, uy, uz, rho)]
        _check(load_library().tlbm_gpu_init_from_macros(self._h, *[a.ctypes.data for a in arrs]))

    def reduce(self, allow_nonfinite: bool = False) -> dict:
        """FP64 mass, momentum and residual of the state.  A state holding
        NaN or infinite populations raises TlbmError (code ENONFINITE)
        unless allow_nonfinite, which returns the sums with "finite" False."""
        L = load_library()
        out = np.zeros(5, np.float64)
        rc = L.tlbm_gpu_reduce(self._h, out.ctypes.data)
        if not (rc == ENONFINITE and allow_nonfinite):
            _check(rc)
        return {"mass": out[0], "momentum": out[1:4].copy(), "residual": out[4],
                "finite": rc == 0}

    def time_average(self, enable: bool = True) -> None:
        _check(load_library().tlbm_gpu_time_average(self._h, int(bool(enable))))

    def get_time_average(self):
        out = [np.empty(self.N, np.float64) for _ in range(4)]
        n = ctypes.c_int64()
        _check(load_library().tlbm_gpu_get_time_average(self._h, *[a.ctypes.data for a in out],
                                                          ctypes.byref(n)))
        return dict(zip(("u", "v", "w", "rho"), out)), int(n.value)

    def set_closure(self, params, feature_scale: float = 1.0, output_scale: float = 1.0) -> None:
        """Learned closure in place of Smagorinsky (params None: off)."""
        if params is None:
            _check(load_library().tlbm_gpu_set_closure(self._h, None, 0.0, 0.0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        if p.size != MLP_NPARAM:
            raise ValueError(f"the network has {MLP_NPARAM} parameters")
        _check(load_library().tlbm_gpu_set_closure(self._h, p.ctypes.data, float(feature_scale),
                                                   float(output_scale)))

    def set_inlet_profile(self, uz) -> None:
        if uz is None:
            _check(load_library().tlbm_gpu_set_inlet_profile(self._h, None))
            return
        a = _f64(uz, self.shape[0] * self.shape[1])
        _check(load_library().tlbm_gpu_set_inlet_profile(self._h, a.ctypes.data))

    def box_filter(self) -> dict:
        """2x2x2 box filter of the last step's moments and viscous stresses
        (coarse global order); keys in FILTER_PLANES."""
        nx, ny, nz = (d // 2 for d in self.shape)
        out = np.empty(len(FILTER_PLANES) * nx * ny * nz, np.float64)
        _check(load_library().tlbm_gpu_box_filter(self._h, out.ctypes.data))
        return dict(zip(FILTER_PLANES, out.reshape(len(FILTER_PLANES), -1)))

    def memory_bytes(self) -> int:
        b = ctypes.c_uint64()
        _check(load_library().tlbm_gpu_memory_bytes(self._h, ctypes.byref(b)))
        return int(b.value)


def declared_symbols(header: str = HEADER_PATH) -> list:
    """Function names the C-ABI header declares."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"\b(tlbm_gpu_[a-z_0-9]+)\s*\(", text)))


MLP_NPARAM = 1345


def mlp_infer(params, x) -> np.ndarray:
    """z of feature vectors through the collide kernel's wavefront inference."""
    p = np.ascontiguousarray(params, dtype=np.float32).ravel()
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 7)
    z = np.empty(len(xx), np.float32)
    _check(load_library().tlbm_mlp_infer(p.ctypes.data, xx.ctypes.data, len(xx), z.ctypes.data))
    return z


class MlpTrainer:
    """Adam training of the closure network on the GPU (HIP kernels)."""

    def __init__(self, x, t, init, seed: int = 0, max_batch: int = 65536):
        L = load_library()
        self._x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 7)
        self._t = np.ascontiguousarray(t, dtype=np.float32).ravel()
        if len(self._x) != len(self._t):
            raise ValueError("x and t differ in length")
        p = np.ascontiguousarray(init, dtype=np.float32).ravel()
        h = _P()
        _check(L.tlbm_mlp_trainer_create(self._x.ctypes.data, self._t.ctypes.data, len(self._t),
                                         p.ctypes.data, int(seed), int(max_batch), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load_library().tlbm_mlp_trainer_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, steps: int, batch: int, lr: float = 1e-3) -> float:
        loss = ctypes.c_double()
        _check(load_library().tlbm_mlp_trainer_step(self._h, int(steps), int(batch), float(lr),
                                                    ctypes.byref(loss)))
        return loss.value

    def loss_grad(self, first: int, count: int):
        loss = ctypes.c_double()
        g = np.empty(MLP_NPARAM, np.float32)
        _check(load_library().tlbm_mlp_trainer_loss_grad(self._h, int(first), int(count),
                                                         ctypes.byref(loss), g.ctypes.data))
        return loss.value, g

    def params(self) -> np.ndarray:
        p = np.empty(MLP_NPARAM, np.float32)
        _check(load_library().tlbm_mlp_trainer_params(self._h, p.ctypes.data))
        return p


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases (b3 = -2: softplus(z) starts near 0.13)."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (32, 7)) * np.sqrt(6.0 / 7)
    W2 = rng.uniform(-1, 1, (32, 32)) * np.sqrt(6.0 / 32)
    w3 = rng.uniform(-1, 1, 32) * np.sqrt(6.0 / 32) * 0.1
    p = np.concatenate([W1.ravel(), np.zeros(32), W2.ravel(), np.zeros(32), w3, [-2.0]])
    return p.astype(np.float32)
