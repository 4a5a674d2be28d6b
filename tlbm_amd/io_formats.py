"""Readers and writers of tLBM's text inputs and binary outputs.

* params.lbm: 23 whitespace-separated tokens in the order
  src/Problem.cpp:10-32 reads them (written one per line as
  python/tLBM_write_params.py:138-166 does, but floats with %.17g so a
  double parse recovers the value exactly).
* ndType.lbm / parts.lbm: one int per line in global node order
  (tLBM_write_params.py:103-108; TLBM_Partition.cpp:120-141, 210-252).
* <field><k>.b_dat: raw native-endian `real` values of the local nodes,
  rank-concatenated (TLBM_Partition.cpp:608-683); ordering.b_dat holds the
  global node index of each entry as int32 (TLBM_Partition.cpp:542-562).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

PARAM_ORDER = (
    "latticeType", "dynamics", "numTs", "tsRepFreq", "warmupTs", "plotFreq",
    "cs", "rhoLBM", "uLBM", "omega", "nx", "ny", "nz", "restartFlag",
    "timeAvgFlag", "lx_p", "ly_p", "lz_p", "tConvFact", "lConvFact",
    "pConvFact", "pRefIdx", "ssDataFlag",
)
_INT_FIELDS = {"dynamics", "numTs", "tsRepFreq", "warmupTs", "plotFreq", "nx", "ny",
               "nz", "restartFlag", "timeAvgFlag", "pRefIdx", "ssDataFlag"}


@dataclass
class Params:
    """The Problem record (include/Problem.h:10-45)."""
    latticeType: str = "D3Q19"
    dynamics: int = 1
    numTs: int = 1
    tsRepFreq: int = 1
    warmupTs: int = 0
    plotFreq: int = 1
    cs: float = 0.0
    rhoLBM: float = 1.0
    uLBM: float = 0.0
    omega: float = 1.0
    nx: int = 1
    ny: int = 1
    nz: int = 1
    restartFlag: int = 0
    timeAvgFlag: int = 0
    lx_p: float = 1.0
    ly_p: float = 1.0
    lz_p: float = 1.0
    tConvFact: float = 1.0
    lConvFact: float = 1.0
    pConvFact: float = 1.0
    pRefIdx: int = 0
    ssDataFlag: int = 0

    @property
    def lattice_q(self) -> int:
        return {"D3Q15": 15, "D3Q19": 19, "D3Q27": 27}[self.latticeType]

    @property
    def num_nodes(self) -> int:
        return self.nx * self.ny * self.nz


def write_params(path: str, p: Params) -> None:
    with open(path, "w") as fh:
        for name in PARAM_ORDER:
            v = getattr(p, name)
            if name == "latticeType":
                fh.write(f"{v} \n")
            elif name in _INT_FIELDS:
                fh.write("%d \n" % int(v))
            else:
                fh.write("%.17g \n" % float(v))


def read_params(path: str) -> Params:
    with open(path) as fh:
        tok = fh.read().split()
    if len(tok) < len(PARAM_ORDER):
        raise ValueError(f"{path}: {len(tok)} tokens, expected {len(PARAM_ORDER)}")
    kw = {}
    for name, t in zip(PARAM_ORDER, tok):
        if name == "latticeType":
            kw[name] = t
        elif name in _INT_FIELDS:
            kw[name] = int(t)
        else:
            kw[name] = float(t)
    return Params(**kw)


def write_int_lines(path: str, values) -> None:
    arr = np.asarray(values, dtype=np.int64).ravel()
    with open(path, "w") as fh:
        fh.write("".join("%d \n" % v for v in arr.tolist()))


def read_int_lines(path: str) -> np.ndarray:
    return np.loadtxt(path, dtype=np.int64, ndmin=1)


def slab_parts(nx: int, ny: int, nz: int, nparts: int) -> np.ndarray:
    """1-D z-slab partition ids (contiguous global ranges per part)."""
    z = np.repeat(np.arange(nz), nx * ny)
    return (z * nparts // nz).astype(np.int64)


def write_case(directory: str, p: Params, nd_type, parts=None) -> None:
    """params.lbm, ndType.lbm and parts.lbm for one run directory."""
    os.makedirs(directory, exist_ok=True)
    write_params(os.path.join(directory, "params.lbm"), p)
    write_int_lines(os.path.join(directory, "ndType.lbm"), nd_type)
    if parts is None:
        parts = np.zeros(p.num_nodes, dtype=np.int64)
    write_int_lines(os.path.join(directory, "parts.lbm"), parts)


def read_bdat(directory: str, field: str, k: int, dtype) -> np.ndarray:
    """One dump field reordered to global order through ordering.b_dat."""
    raw = np.fromfile(os.path.join(directory, f"{field}{k}.b_dat"), dtype=dtype)
    order = np.fromfile(os.path.join(directory, "ordering.b_dat"), dtype=np.int32)
    if raw.size != order.size:
        raise ValueError(f"{field}{k}.b_dat has {raw.size} entries, ordering has {order.size}")
    out = np.empty_like(raw)
    out[order] = raw
    return out


DUMP_FIELDS = ("ux", "uy", "uz", "density", "Fx", "Fy", "Fz")
