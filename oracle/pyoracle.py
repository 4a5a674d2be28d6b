"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class _Params(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("nz", ctypes.c_int), ("dynamics", ctypes.c_int), ("cs", ctypes.c_double),
        ("rho_lbm", ctypes.c_double), ("u_lbm", ctypes.c_double),
        ("omega", ctypes.c_double), ("exact_weights", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(LIB_PATH)
        for suf in ("f64", "f32"):
            getattr(L, f"oracle_create_{suf}").restype = ctypes.c_void_p
            getattr(L, f"oracle_create_{suf}").argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p]
            getattr(L, f"oracle_destroy_{suf}").argtypes = [ctypes.c_void_p]
            getattr(L, f"oracle_step_{suf}").argtypes = [ctypes.c_void_p, ctypes.c_int]
            for fn in ("get_macros", "get_forces", "get_populations"):
                n = {"get_macros": 4, "get_forces": 3, "get_populations": 1}[fn]
                getattr(L, f"oracle_{fn}_{suf}").argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * n
            getattr(L, f"oracle_set_populations_{suf}").argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            getattr(L, f"oracle_set_closure_{suf}").argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                                               ctypes.c_double, ctypes.c_double]
        L.mlp_forward_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.mlp_loss_grad_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_void_p]
        L.mlp_loss_grad_f64.restype = ctypes.c_double
        L.mlp_adam_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double]
        L.mlp_infer_bf16.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.mlp_round_bf16.argtypes = [ctypes.c_float]
        L.mlp_round_bf16.restype = ctypes.c_float
        L.mlp_nparam.restype = ctypes.c_int
        L.oracle_get_weights_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


class Oracle:
    """One lattice run on the CPU: the reference's single-rank time step."""

    def __init__(self, lattice: int, nx: int, ny: int, nz: int, nd_type, *, omega: float,
                 u_lbm: float = 0.0, rho_lbm: float = 1.0, cs: float = 0.0,
                 dynamics: int = 1, precision: str = "f64", exact_weights: bool = False):
        if precision not in ("f64", "f32"):
            raise ValueError(precision)
        self.suf = precision
        self.dtype = np.float64 if precision == "f64" else np.float32
        self.Q = lattice
        self.N = nx * ny * nz
        p = _Params(lattice, nx, ny, nz, dynamics, cs, rho_lbm, u_lbm, omega, int(exact_weights))
        nd = np.ascontiguousarray(np.asarray(nd_type, dtype=np.int32).ravel())
        if nd.size != self.N:
            raise ValueError(f"ndType has {nd.size} entries, lattice has {self.N}")
        self._h = getattr(lib(), f"oracle_create_{self.suf}")(ctypes.byref(p), nd.ctypes.data)
        if not self._h:
            raise ValueError("oracle rejected the lattice, sizes or node types")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), f"oracle_destroy_{self.suf}")(h)
            self._h = None

    def step(self, n: int = 1) -> None:
        getattr(lib(), f"oracle_step_{self.suf}")(self._h, int(n))

    def macros(self):
        out = [np.empty(self.N, self.dtype) for _ in range(4)]
        getattr(lib(), f"oracle_get_macros_{self.suf}")(self._h, *[a.ctypes.data for a in out])
        return dict(zip(("ux", "uy", "uz", "rho"), out))

    def forces(self):
        out = [np.empty(self.N, self.dtype) for _ in range(3)]
        getattr(lib(), f"oracle_get_forces_{self.suf}")(self._h, *[a.ctypes.data for a in out])
        return dict(zip(("Fx", "Fy", "Fz"), out))

    def populations(self):
        f = np.empty(self.N * self.Q, self.dtype)
        getattr(lib(), f"oracle_get_populations_{self.suf}")(self._h, f.ctypes.data)
        return f.reshape(self.N, self.Q)

    def set_populations(self, f) -> None:
        a = np.ascontiguousarray(f, dtype=self.dtype).ravel()
        if a.size != self.N * self.Q:
            raise ValueError(f"expected {self.N * self.Q} populations, got {a.size}")
        getattr(lib(), f"oracle_set_populations_{self.suf}")(self._h, a.ctypes.data)

    def set_closure(self, params, cl_in: float, cl_out: float) -> None:
        if params is None:
            getattr(lib(), f"oracle_set_closure_{self.suf}")(self._h, None, 0.0, 0.0)
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        getattr(lib(), f"oracle_set_closure_{self.suf}")(self._h, p.ctypes.data, cl_in, cl_out)

    def weights(self):
        if self.suf != "f64":
            raise ValueError("weights() reads the f64 table")
        w = np.empty(self.Q, np.float64)
        lib().oracle_get_weights_f64(self._h, w.ctypes.data)
        return w


# ---- closure network (FP64 reference of the HIP training / inference) ----

MLP_NPARAM = 1345


def mlp_forward(params, x):
    p = np.ascontiguousarray(params, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 7)
    y = np.empty(len(xx))
    lib().mlp_forward_f64(p.ctypes.data, xx.ctypes.data, len(xx), y.ctypes.data)
    return y


def mlp_loss_grad(params, x, t):
    p = np.ascontiguousarray(params, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 7)
    tt = np.ascontiguousarray(t, dtype=np.float64)
    g = np.empty(MLP_NPARAM)
    loss = lib().mlp_loss_grad_f64(p.ctypes.data, xx.ctypes.data, tt.ctypes.data, len(xx), g.ctypes.data)
    return loss, g


def mlp_adam(params, grad, mom, vel, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    for a in (params, mom, vel):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    g = np.ascontiguousarray(grad, dtype=np.float64)
    lib().mlp_adam_f64(params.ctypes.data, g.ctypes.data, mom.ctypes.data, vel.ctypes.data, step, lr,
                       beta1, beta2, eps)


def mlp_infer_bf16(params, x):
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 7)
    z = np.empty(len(xx))
    lib().mlp_infer_bf16(p.ctypes.data, xx.ctypes.data, len(xx), z.ctypes.data)
    return z


def round_bf16(v):
    return np.array([lib().mlp_round_bf16(float(a)) for a in np.ravel(v)], dtype=np.float32)
