"""Generate the golden fixtures from the compiled reference.

Runs the reference executables built by oracle/Makefile (oracle/_ref/
tLBMexec: real=float as shipped; tLBMexec_fp64: real=double) on small
lattices written by tlbm_amd.geometry / tlbm_amd.io_formats, and stores the
inputs with the dumped ux/uy/uz/density/Fx/Fy/Fz fields as compressed .npz
files next to this script.  Needs /root/reference (to build the
executables) and the image's MPICH in /opt/conda; it is run by hand in the
development container, never by the tests.

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tlbm_amd import geometry, io_formats  # noqa: E402

REF_DIR = os.path.join(REPO, "oracle", "_ref")
MPIEXEC = "/opt/conda/bin/mpiexec"
LATT = {15: "D3Q15", 19: "D3Q19", 27: "D3Q27"}

# dyadic parameters: float and double parses of the text agree exactly
SMALL = dict(u=0.0625, omega=1.6875, rho=1.0)
CS_ON = 0.125


def run_reference(exe: str, params: io_formats.Params, nd, nranks: int = 1) -> dict:
    dtype = np.float64 if exe.endswith("_fp64") else np.float32
    with tempfile.TemporaryDirectory(prefix="tlbm_golden_") as d:
        parts = io_formats.slab_parts(params.nx, params.ny, params.nz, nranks)
        io_formats.write_case(d, params, nd, parts)
        cmd = [MPIEXEC, "-n", str(nranks), os.path.join(REF_DIR, exe)]
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"{cmd} failed:\n{r.stdout}\n{r.stderr}")
        return {f: io_formats.read_bdat(d, f, 0, dtype) for f in io_formats.DUMP_FIELDS}


def make_params(q, nx, ny, nz, nsteps, *, u, omega, rho, cs=0.0, dynamics=1):
    return io_formats.Params(latticeType=LATT[q], dynamics=dynamics, numTs=nsteps,
                             tsRepFreq=nsteps, warmupTs=0, plotFreq=nsteps, cs=cs,
                             rhoLBM=rho, uLBM=u, omega=omega, nx=nx, ny=ny, nz=nz)


def save(name: str, params_list, nd, outputs, meta=None):
    blob = {"ndType": np.asarray(nd, dtype=np.uint8)}
    plist = []
    for i, (p, out) in enumerate(zip(params_list, outputs)):
        plist.append({k: getattr(p, k) for k in io_formats.PARAM_ORDER})
        for f, v in out.items():
            blob[f"{f}_{i}"] = v
    blob["meta"] = np.frombuffer(json.dumps({"params": plist, **(meta or {})}).encode(), np.uint8)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **blob)
    print(f"wrote {path} ({os.path.getsize(path)} B)")


def small_channel():
    nx, ny, nz = 8, 7, 12
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 3, 4, 2)])
    return nx, ny, nz, nd


def main():
    if not os.path.exists(os.path.join(REF_DIR, "tLBMexec_fp64")):
        raise SystemExit("build the reference first: make -C oracle ref")
    nx, ny, nz, nd = small_channel()
    steps = (1, 2, 50)
    # channel matrix: lattice x Smagorinsky on/off x precision, three step counts each
    for q in (15, 19, 27):
        for cs in (0.0, CS_ON):
            for exe, tag in (("tLBMexec_fp64", "f64"), ("tLBMexec", "f32")):
                plist = [make_params(q, nx, ny, nz, n, cs=cs, **SMALL) for n in steps]
                outs = [run_reference(exe, p, nd) for p in plist]
                save(f"chan_q{q}_cs{int(cs > 0)}_{tag}", plist, nd, outs)
    # rank-count invariance of the moments (and not of the forces)
    p = make_params(19, nx, ny, nz, 50, **SMALL)
    one = run_reference("tLBMexec_fp64", p, nd, 1)
    four = run_reference("tLBMexec_fp64", p, nd, 4)
    for f in ("ux", "uy", "uz", "density"):
        assert np.array_equal(one[f], four[f]), f
    print("1 rank == 4 ranks for ux/uy/uz/density")
    # lid-driven cavity: prescribed-u_z nodes (type 5)
    lnx, lny, lnz = 7, 8, 9
    lid = geometry.lid_cavity(lnx, lny, lnz)
    for exe, tag in (("tLBMexec_fp64", "f64"), ("tLBMexec", "f32")):
        plist = [make_params(19, lnx, lny, lnz, n, **SMALL) for n in (1, 50)]
        save(f"lid_q19_{tag}", plist, lid, [run_reference(exe, pp, lid) for pp in plist])
    # dynamics = 2: strain computed, no relaxation
    p2 = make_params(19, nx, ny, nz, 20, dynamics=2, **SMALL)
    save("dyn2_q19_f64", [p2], nd, [run_reference("tLBMexec_fp64", p2, nd)])
    # periodic box with a sphere at rest: mass grows with the float weights
    bnx, bny, bnz = 6, 5, 7
    box = geometry.periodic_box(bnx, bny, bnz, obstacles=[geometry.sphere(2, 2, 3, 1.5)])
    pb = make_params(19, bnx, bny, bnz, 100, u=0.0, omega=1.5, rho=1.0)
    save("rest_q19_f64", [pb], box, [run_reference("tLBMexec_fp64", pb, box)])
    # 24x24x96 D3Q15 channel, sphere r=4 at (12,12,24), tau=0.6, u=0.05, 2000 steps, 4 ranks
    c1 = geometry.channel(24, 24, 96, obstacles=[geometry.sphere(12, 12, 24, 4)])
    p1 = make_params(15, 24, 24, 96, 2000, u=0.05, omega=1.0 / 0.6, rho=1.0)
    out1 = run_reference("tLBMexec_fp64", p1, c1, 4)
    out1 = {f: out1[f] for f in ("ux", "uy", "uz", "density")}
    save("channel_d3q15_24x24x96_f64", [p1], c1, [out1], meta={"ranks": 4})


if __name__ == "__main__":
    main()
