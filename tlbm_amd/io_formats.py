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
import zlib
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


def int_lines_bytes(values) -> bytes:
    """The text "%d \n" per value (non-negative or negative ints), built
    with array arithmetic: 268M node types take seconds, not minutes."""
    raw = np.asarray(values).ravel()
    if raw.size == 0:
        return b""
    if raw.min() >= 0 and raw.max() <= 9:        # node types: 3 bytes a line
        out = np.empty((raw.size, 3), np.uint8)
        out[:, 0] = raw.astype(np.uint8) + ord("0")
        out[:, 1] = ord(" ")
        out[:, 2] = ord("\n")
        return out.tobytes()
    v = raw.astype(np.int64)
    neg = v < 0
    a = np.abs(v)
    ndig = np.ones(v.size, np.int64)
    t = a // 10
    while np.any(t > 0):
        ndig += t > 0
        t //= 10
    width = ndig + neg + 2                      # digits, sign, " \n"
    end = np.cumsum(width)
    start = end - width
    out = np.empty(int(end[-1]), np.uint8)
    out[end - 1] = ord("\n")
    out[end - 2] = ord(" ")
    out[start[neg]] = ord("-")
    first = start + neg                         # position of the leading digit
    rem = a.copy()
    for k in range(int(ndig.max()) - 1, -1, -1):
        has = ndig > k                          # values with a digit at place 10^k
        pos = first[has] + (ndig[has] - 1 - k)
        out[pos] = ord("0") + (rem[has] // 10 ** k) % 10
    return out.tobytes()


def write_int_lines(path: str, values) -> None:
    with open(path, "wb") as fh:
        fh.write(int_lines_bytes(values))


# Binary sidecar of ndType.lbm: magic, node count (little-endian uint64),
# then one byte per node in global order.  The drop-in driver reads it in
# place of the text file when it is present and not older than ndType.lbm.
NDTYPE_BIN_MAGIC = b"TLBMNDT2"


def _text_signature(text_path):
    """(size in bytes, CRC-32) of a text file, or (0, 0) when absent."""
    if text_path is None or not os.path.exists(text_path):
        return 0, 0
    crc = 0
    with open(text_path, "rb") as fh:
        while True:
            block = fh.read(1 << 22)
            if not block:
                break
            crc = zlib.crc32(block, crc)
    return os.path.getsize(text_path), crc


def write_ndtype_bin(path: str, nd_type, text_path=None) -> None:
    """The binary node-type sidecar: "TLBMNDT2", the node count, the size
    and CRC-32 of the ndType.lbm it was made from (zeros when none; each
    a little-endian uint64), then one byte per node in global order.  The
    driver reads it in place of ndType.lbm only when that text file is
    absent or is the one recorded."""
    nd = np.asarray(nd_type).ravel()
    if nd.size and (nd.min() < 0 or nd.max() > 255):
        raise ValueError("node types must fit in a byte")
    size, crc = _text_signature(text_path)
    with open(path, "wb") as fh:
        fh.write(NDTYPE_BIN_MAGIC)
        fh.write(np.array([nd.size, size, crc], np.uint64).astype("<u8").tobytes())
        fh.write(nd.astype(np.uint8).tobytes())


def read_ndtype_bin(path: str) -> np.ndarray:
    with open(path, "rb") as fh:
        head = fh.read(32)
        if len(head) != 32 or head[:8] != NDTYPE_BIN_MAGIC:
            raise ValueError(f"{path}: not an ndType sidecar")
        n = int(np.frombuffer(head[8:16], "<u8")[0])
        nd = np.frombuffer(fh.read(), np.uint8)
    if nd.size != n:
        raise ValueError(f"{path}: {nd.size} node types, header says {n}")
    return nd.copy()


def read_int_lines(path: str) -> np.ndarray:
    return np.loadtxt(path, dtype=np.int64, ndmin=1)


def slab_parts(nx: int, ny: int, nz: int, nparts: int) -> np.ndarray:
    """1-D z-slab partition ids (contiguous global ranges per part)."""
    z = np.repeat(np.arange(nz), nx * ny)
    return (z * nparts // nz).astype(np.int64)


def write_case(directory: str, p: Params, nd_type, parts=None, binary: bool = False) -> None:
    """params.lbm, ndType.lbm and parts.lbm for one run directory (and,
    with binary, the ndType.bin sidecar, written after the text file)."""
    os.makedirs(directory, exist_ok=True)
    write_params(os.path.join(directory, "params.lbm"), p)
    write_int_lines(os.path.join(directory, "ndType.lbm"), nd_type)
    if parts is None:
        parts = np.zeros(p.num_nodes, dtype=np.int64)
    write_int_lines(os.path.join(directory, "parts.lbm"), parts)
    if binary:
        write_ndtype_bin(os.path.join(directory, "ndType.bin"), nd_type,
                         text_path=os.path.join(directory, "ndType.lbm"))


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
