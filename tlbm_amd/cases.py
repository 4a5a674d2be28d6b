"""Run directories for the measured configurations, in tLBM's input formats.

Each case writes params.lbm, ndType.lbm and parts.lbm (the files
python/tLBM_write_params.py:103-166 writes and the reference's tLBMexec and
this build's tlbm_exec both read), optionally with the ndType.bin sidecar
that tlbm_exec reads instead of parsing a large text file.  Geometry is in
integer lattice units with the FluidChannel classification rules
(tlbm_amd/geometry.py).

    python -m tlbm_amd.cases channel   <dir>            # 24x24x96 D3Q15, sphere r 4
    python -m tlbm_amd.cases cylinders <dir>            # 128x128x512 D3Q19, cylinder array
    python -m tlbm_amd.cases porous    <dir>            # 384^3 D3Q27 Smagorinsky, sphere pack
    python -m tlbm_amd.cases closure   <dir>            # 256x128x128 D3Q19 fine run of the closure
    python -m tlbm_amd.cases spheroid  <dir> --binary   # 1024x512x512 D3Q19, prolate spheroid

Options: --parts N (z-slab partition ids for N ranks), --steps, --plot-freq,
--binary.  The closure case's seeded inlet perturbation and the porous
case's start-up ramp are not expressible in params.lbm (uniform u_LBM); the
Python runners apply them through the C-ABI.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import geometry, io_formats


def _params(lattice, dims, omega, u, steps, cs=0.0, plot_freq=None, rep=None):
    nx, ny, nz = dims
    return io_formats.Params(latticeType=lattice, dynamics=1, numTs=steps,
                             tsRepFreq=rep or max(1, steps // 10), warmupTs=0,
                             plotFreq=plot_freq or steps, cs=cs, rhoLBM=1.0, uLBM=u,
                             omega=omega, nx=nx, ny=ny, nz=nz)


def channel():
    """24x24x96 D3Q15 BGK FP64 channel: walls, sphere r 4 at (12, 12, 24),
    inlet u_z 0.05, tau 0.6, 2000 steps."""
    dims = (24, 24, 96)
    nd = geometry.channel(*dims, obstacles=[geometry.sphere(12, 12, 24, 4)])
    return _params("D3Q15", dims, 1.0 / 0.6, 0.05, 2000), nd


def cylinders():
    """128x128x512 D3Q19 BGK, periodic x and y, cylinder along x, r 12 at
    (y, z) = (64, 128), inlet u_z 0.04, tau 0.55, 20000 steps."""
    dims = (128, 128, 512)
    nd = geometry.channel(*dims, walls=(), obstacles=[geometry.cylinder_x(64, 128, 12)])
    return _params("D3Q19", dims, 1.0 / 0.55, 0.04, 20000), nd


def porous():
    """384^3 D3Q27 Smagorinsky (cs 0.1) porous channel, sphere pack to
    porosity 0.65, inlet u_z 0.05, tau_0 0.505, 5000 steps."""
    from .workloads import porous_case
    nd, _ = porous_case(384)
    return _params("D3Q27", (384, 384, 384), 1.0 / 0.505, 0.05, 5000, cs=0.1), nd


def closure():
    """The closure's fine run: 256x128x128 D3Q19, tau 0.502, Smagorinsky cs
    0.1, 4x2 staggered spheres r 8, periodic x and y, inlet u_z 0.06."""
    from .closure import ClosureCase, fine_geometry
    case = ClosureCase()
    nd = fine_geometry(case)
    return _params("D3Q19", (case.nx, case.ny, case.nz), 1.0 / case.tau_fine, case.u_in,
                   case.spinup + case.snapshots * case.every, cs=case.cs), nd


def spheroid():
    """1024x512x512 D3Q19, prolate spheroid (semi-axes 96 along z, 32 across)
    centred at (512, 256, 256), walls, inlet u_z 0.05, tau 0.51, 10000 steps."""
    from .workloads import spheroid_case
    dims, nd, _ = spheroid_case(1)
    return _params("D3Q19", dims, 1.0 / 0.51, 0.05, 10000), nd


CASES = {"channel": channel, "cylinders": cylinders, "porous": porous, "closure": closure,
         "spheroid": spheroid}


def write(name: str, directory: str, parts: int = 1, steps: int | None = None,
          plot_freq: int | None = None, binary: bool = False):
    p, nd = CASES[name]()
    if steps is not None:
        p.numTs = steps
        p.tsRepFreq = max(1, steps // 10)
        p.plotFreq = steps
    if plot_freq is not None:
        p.plotFreq = plot_freq
    part_ids = io_formats.slab_parts(p.nx, p.ny, p.nz, parts) if parts > 1 else None
    io_formats.write_case(directory, p, nd, part_ids, binary=binary)
    return p, nd


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("case", choices=sorted(CASES))
    ap.add_argument("directory")
    ap.add_argument("--parts", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--plot-freq", type=int, default=None)
    ap.add_argument("--binary", action="store_true", help="also write the ndType.bin sidecar")
    a = ap.parse_args(argv)
    p, nd = write(a.case, a.directory, a.parts, a.steps, a.plot_freq, a.binary)
    counts = np.bincount(nd, minlength=6)
    print(f"{a.case}: {p.latticeType} {p.nx}x{p.ny}x{p.nz}, {p.numTs} steps, node types "
          f"fluid {counts[0]} solid {counts[1]} inlet {counts[2]} outlet {counts[3]} -> {a.directory}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
