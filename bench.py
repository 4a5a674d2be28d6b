#!/usr/bin/env python3
"""Benchmark of the lattice-Boltzmann time step on one MI355X.

Workload: D3Q19 BGK flow through an infinite array of cylinders, lattice
128x128x512 (x, y, z; flow along z), periodic in x and y, cylinder axis x,
radius 12, centred at (y, z) = (64, 128), bounce-back cylinder, Zou-He
velocity inlet u_z = 0.04 at z = 0, pressure outlet rho = 1 at z = 511,
tau = 0.55, starting from rest (f = w).

A step is one fused stream-collide launch over the whole lattice with the
populations resident in HBM.  The headline `value` is MLUPS with FP32
storage (nx*ny*nz*steps / wall time of the timed steps); FP64, FP16 (f - w)
and FP32X storage run the same timed steps, reported under `variants`.
After the timed region every storage runs the configuration's full 20000
steps from rest on the full lattice (--accuracy-steps), and each reduced
storage reports its velocity L2 over fluid nodes, its mass drift (FP64
on-device sums) and the reference's own validation metric, the relative
1-norm difference of the speed |u| over all nodes (python/validate.py:33-43),
against the FP64 run.  `roofline` prices the FP32 kernel at 2*Q*4 = 152
algorithmic bytes per lattice update against 8 TB/s, with its average
duration from HIP events on the stream the kernel runs on.
`cpu_baseline` times the reference's own MPI executable (built from its
sources into oracle/_ref) on the host's physical cores (those this process
may run on, at most its CPU share) for 200 steps of the same lattice, or the
C oracle when that executable cannot run.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 100]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tlbm_amd import geometry, io_formats  # noqa: E402
from tlbm_amd.gpu import BYTES_PER_VALUE, Lattice  # noqa: E402

METRIC = "D3Q19 MLUPS & %HBM roofline, FP32/FP16 vs FP64 at stated L2 error"
PEAK_HBM_GBS = 8000.0
Q = 19
NX, NY, NZ = 128, 128, 512
TAU = 0.55
U_IN = 0.04
WORKLOAD = "cylinder_array_d3q19_128x128x512"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_stream_collide_d3q19_128x128x512.json")


def node_types():
    return geometry.channel(NX, NY, NZ, walls=(), obstacles=[geometry.cylinder_x(64, 128, 12)])


def run_gpu(storage, nd, steps, warmup):
    kw = dict(omega=1.0 / TAU, u_lbm=U_IN, rho_lbm=1.0)
    with Lattice(Q, NX, NY, NZ, nd, storage=storage, **kw) as lat:
        lat.step(warmup)
        lat.sync()
        t0 = time.perf_counter()
        ms_events = lat.step_timed(steps)
        wall = time.perf_counter() - t0
    n = NX * NY * NZ
    return {
        "wall_s": wall,
        "kernel_ms": ms_events / steps,
        "mlups": n * steps / wall / 1e6,
    }


def accuracy_runs(nd, storages, steps):
    """Each storage from rest over `steps` steps on the full lattice (one
    lattice at a time), the reduced storages compared with FP64: velocity
    L2 over the nodes that are not solid (inlet and outlet included, as the
    workload runners count them), mass drift of the FP64 on-device sums, and
    the relative 1-norm of the speed difference over all nodes."""
    fluid = nd != geometry.SOLID
    kw = dict(omega=1.0 / TAU, u_lbm=U_IN, rho_lbm=1.0)
    ref = None
    out = {}
    for st in ["f64"] + [s for s in storages if s != "f64"]:
        print(f"[bench] accuracy: {st}, {steps} steps", file=sys.stderr, flush=True)
        with Lattice(Q, NX, NY, NZ, nd, storage=st, **kw) as lat:
            lat.step(steps)
            m = lat.macros()
            mass = lat.reduce()["mass"]
        if st == "f64":
            ref = (m, mass)
            continue
        out[st] = {"accuracy_steps": steps,
                   "l2_velocity_vs_f64": l2_velocity(m, ref[0], fluid),
                   "mass_drift_vs_f64": abs(mass - ref[1]) / ref[1],
                   "rel_l1_speed_vs_f64": rel_l1_speed(m, ref[0])}
    return out


def l2_velocity(m, ref, fluid):
    a = np.stack([m[k][fluid] for k in ("ux", "uy", "uz")])
    b = np.stack([ref[k][fluid] for k in ("ux", "uy", "uz")])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_l1_speed(m, ref):
    """The reference's validation metric (python/validate.py:33-43):
    ||(|u| - |u_ref|)||_1 / || |u_ref| ||_1 over every node."""
    s = np.sqrt(m["ux"] ** 2 + m["uy"] ** 2 + m["uz"] ** 2)
    r = np.sqrt(ref["ux"] ** 2 + ref["uy"] ** 2 + ref["uz"] ** 2)
    return float(np.sum(np.abs(s - r)) / np.sum(np.abs(r)))


def traffic_bytes(storage):
    """HBM bytes per launch measured with rocprofv3 PMC counters (committed
    summary), or None when no measurement exists for this storage."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
        return t[storage]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_reference(nd, steps, cores, timeout):
    """The reference executable (real=float) on `cores` MPI ranks with 1-D
    z-slab partitions; returns its own `Estimated LPU/s` in MLUPS."""
    exe = os.path.join(ROOT, "oracle", "_ref", "tLBMexec")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        raise RuntimeError("reference executable or mpiexec missing")
    p = io_formats.Params(latticeType="D3Q19", dynamics=1, numTs=steps, tsRepFreq=steps,
                          warmupTs=0, plotFreq=steps + 1, cs=0.0, rhoLBM=1.0, uLBM=U_IN,
                          omega=1.0 / TAU, nx=NX, ny=NY, nz=NZ)
    with tempfile.TemporaryDirectory(prefix="tlbm_cpu_") as d:
        io_formats.write_params(os.path.join(d, "params.lbm"), p)
        io_formats.write_int_lines(os.path.join(d, "ndType.lbm"), nd)
        io_formats.write_int_lines(os.path.join(d, "parts.lbm"), io_formats.slab_parts(NX, NY, NZ, cores))
        r = subprocess.run([mpiexec, "-launcher", "fork", "-n", str(cores), exe], cwd=d,
                           capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}): {r.stderr[-400:]}")
        mt = re.search(r"Estimated LPU/s = ([0-9.eE+-]+)", r.stdout)
        if not mt:
            raise RuntimeError("no LPU/s line in the reference output")
        return float(mt.group(1)) / 1e6


def cpu_oracle(nd, steps):
    """The C oracle (one thread, FP32 restatement) on the same lattice."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    o = Oracle(Q, NX, NY, NZ, nd.astype(np.int32), omega=1.0 / TAU, u_lbm=U_IN, rho_lbm=1.0,
               precision="f32")
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return NX * NY * NZ * steps / dt / 1e6


def host_cores():
    """(ranks, physical cores, logical CPUs, CPU model, share): the physical
    cores among the CPUs this process may run on (lscpu's CORE/SOCKET of
    each CPU in the affinity set), capped by the CPU share the host grants
    one GPU job (OMP_NUM_THREADS when set)."""
    aff = sorted(os.sched_getaffinity(0))
    phys = set()
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        allowed = set(aff)
        for ln in out.splitlines():
            if ln.startswith("#"):
                continue
            cpu, core, sock = (int(v) for v in ln.split(",")[:3])
            if cpu in allowed:
                phys.add((sock, core))
    except (OSError, ValueError, subprocess.SubprocessError):
        phys = set(aff)
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(aff)
    ncores = max(1, len(phys))
    return min(ncores, share), ncores, len(aff), model, share


def cpu_baseline(nd, args):
    ranks, ncores, nlogical, model, share = host_cores()
    if args.cpu_cores > 0:
        ranks = args.cpu_cores
    where = (f"{model}; {ncores} physical cores / {nlogical} logical CPUs in this process's "
             f"affinity set, CPU share {share}")
    try:
        v = cpu_reference(nd, args.cpu_steps, ranks, timeout=args.cpu_timeout)
        return {"value": round(v, 3), "unit": "MLUPS", "cores": ranks, "kind": "reference",
                "sample": f"reference tLBMexec (real=float, MPICH, {ranks} ranks, one per physical "
                          f"core, 1-D z-slabs) on the full {NX}x{NY}x{NZ} D3Q19 lattice for "
                          f"{args.cpu_steps} steps; its own Estimated LPU/s (loop time, no dumps); "
                          f"host: {where}"}
    except Exception as exc:  # the reference cannot run here: time the oracle instead
        steps = max(1, args.cpu_steps // 20)
        v = cpu_oracle(nd, steps)
        return {"value": round(v, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
                "sample": f"C oracle (FP32 restatement, 1 thread) on the full {NX}x{NY}x{NZ} "
                          f"lattice for {steps} steps (reference run unavailable: {exc}); "
                          f"host: {where}"}


def copy_calibration(mib, reps=5):
    """Device-to-device copies of `mib` MiB through the HIP runtime (a plain
    copy kernel of known size): the reference point for scaling the
    FETCH_SIZE / WRITE_SIZE counters of the stream-collide kernel."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mib) << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(reps):
            if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=20000,
                    help="after the timed region, run every storage this many steps from rest "
                         "on the full lattice for its errors against FP64 (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="MPI ranks of the CPU baseline (0: one per physical core, at most the "
                         "host's CPU share)")
    ap.add_argument("--cpu-timeout", type=float, default=300.0)
    ap.add_argument("--storages", default="f32,f64,f16,f32x",
                    help="storages to run; the first is the headline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the lattice lives whole on one MI355X and this build has "
              f"no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    nd = node_types()
    storages = args.storages.split(",")
    res = {}
    for st in storages:
        print(f"[bench] {st}: {args.warmup} warm-up + {args.steps} timed steps", file=sys.stderr, flush=True)
        res[st] = run_gpu(st, nd, args.steps, args.warmup)
    head = storages[0]
    h = res[head]
    n = NX * NY * NZ
    b_alg = 2 * Q * BYTES_PER_VALUE[head]
    achieved = b_alg * n / (h["kernel_ms"] * 1e-3) / 1e9
    variants = {}
    for st, r in res.items():
        variants[st] = {"mlups": round(r["mlups"], 1), "kernel_ms": round(r["kernel_ms"], 5),
                        "bytes_per_update": 2 * Q * BYTES_PER_VALUE[st],
                        "hbm_frac": round(2 * Q * BYTES_PER_VALUE[st] * n / (r["kernel_ms"] * 1e-3)
                                          / 1e9 / PEAK_HBM_GBS, 4)}
    if args.accuracy_steps > 0:
        for st, e in accuracy_runs(nd, storages, args.accuracy_steps).items():
            variants[st].update(e)
    line = {
        "metric": METRIC,
        "value": round(h["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["wall_s"] / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head in ("f32", "f16") else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "lattice": "D3Q19", "nx": NX, "ny": NY, "nz": NZ,
                   "storage": head, "tau": TAU, "u_inlet": U_IN, "collision": "BGK"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic_bytes(head)},
        "variants": variants,
    }
    if not args.no_cpu_baseline:
        print(f"[bench] cpu baseline: {args.cpu_steps} steps", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(nd, args)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
