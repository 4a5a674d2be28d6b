"""Runners of the measured workloads besides the stream-collide benchmark.

* porous: D3Q27 Smagorinsky (cs 0.1, the reference's linear-cs form,
  include/LatticeStructure.hpp:140-153) porous channel, 384^3, seeded
  random sequential addition of non-overlapping spheres (radius uniform in
  [6, 14], kept 4 nodes clear of the inlet and outlet planes) to porosity
  0.65, bounce-back side walls, Zou-He inlet u_z 0.05 and outlet rho 1,
  tau_0 0.505, 5000 steps; FP32 (deviation) storage against the FP64 run of
  the same lattice.  Started impulsively this flow leaves the lattice
  Boltzmann regime within ~100 steps and diverges (FP64 and FP32 alike; the
  oracle, which follows the reference's rules, does the same on smaller
  packs); the inlet velocity is therefore raised over the first 3000 steps
  (--ramp 0 runs the impulsive start).
* spheroid: D3Q19 1024x512x512 flow past a prolate spheroid (semi-axes 96
  along z, 32 along x and y, centred at (512, 256, 256)), bounce-back side
  walls, inlet u_z 0.05, outlet rho 1, tau 0.51, 10000 steps; FP16
  deviation storage with the closure network in the collide kernel against
  an FP32 Smagorinsky run; and on the 2x-downsampled copy (512x256x256,
  semi-axes 48/16, tau 0.505 -- acoustic scaling keeps the lattice velocity
  and halves the lattice viscosity -- 5000 steps) FP64, FP32 and FP16 runs,
  the FP64 Smagorinsky run as the reference.  These runs start from
  uniform flow (u_z 0.05 on every non-solid node): from rest, the
  inlet-outlet channel's acoustic mode loads the body with a force that
  swings between +130 and -35 over the run.

* cylinders: the stream-collide benchmark's lattice (D3Q19 BGK 128x128x512,
  cylinder array, inlet u_z 0.04, tau 0.55) over the configuration's full
  20000 steps from rest, FP32 and FP16 against FP64 at checkpoints.
* slab: one GPU's share of the spheroid lattice cut into z-slabs (1024 x
  512 x 64 for 8 GPUs, the slab below the body's centre plane: walls and the
  body's mid-section, no inlet or outlet), periodic in z -- a ring of one,
  whose own z wrap stands in for the ghost planes -- timed for FP16 +
  network and FP32 Smagorinsky beside the face exchange a sharded step would
  overlap, modelled (not measured) as the bytes of the 5 crossing D3Q19
  directions of one face over one xGMI link's 153 GB/s.

Every run is timed with HIP events around batches of steps (kernel time on
the lattice's stream), so MLUPS = nodes x steps / time, counting every node
as the reference does (tLBM.cpp:71-72), and HBM GB/s = MLUPS x 2 Q bytes per
value (each population read once and written once).  Errors: relative L2
of u over fluid nodes; mass drift |M - M_ref| / M_ref with M the FP64
on-device sum of the populations (it also moves with the flow state: the
Zou-He inlet and outlet do not conserve mass, so runs with different
closures settle at different totals); drag = the z force of the momentum
exchange (TLBM_Partition.cpp:143-208) summed over the obstacle's nodes,
averaged over the second half of the run.

    python -m tlbm_amd.workloads porous  [--n 384] [--steps 5000]
    python -m tlbm_amd.workloads spheroid [--scale 1] [--steps 10000] [--network FILE]
    python -m tlbm_amd.workloads cylinders [--steps 20000]
    python -m tlbm_amd.workloads slab [--gpus 8] [--steps 1000]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import geometry
from .closure import uniform_start
from .gpu import Lattice

STORAGE_BYTES = {"f64": 8, "f32": 4, "f16": 2, "f32x": 4}
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_NETWORK = os.path.join(HERE, "closure_d3q19_256x128x128.json")


def _log(msg: str) -> None:
    print(f"[workload] {msg}", flush=True)


def channel_types(obstacle: np.ndarray, walls: bool = True) -> np.ndarray:
    """ndType (flat, global order) of a channel along z from an obstacle
    mask shaped (nz, ny, nx): side walls x = 0, nx-1, y = 0, ny-1 and the
    obstacle solid, inlet z = 0 and outlet z = nz-1 elsewhere
    (FluidChannel.py:1358-1397, 1517-1538).  Built plane by plane."""
    nz, ny, nx = obstacle.shape
    nd = np.zeros((nz, ny, nx), np.uint8)
    wall = np.zeros((ny, nx), bool)
    if walls:
        wall[:, 0] = wall[:, nx - 1] = True
        wall[0, :] = wall[ny - 1, :] = True
    for z in range(nz):
        solid = obstacle[z] | wall
        plane = nd[z]
        plane[solid] = geometry.SOLID
        if z == 0:
            plane[~solid] = geometry.INLET
        elif z == nz - 1:
            plane[~solid] = geometry.OUTLET
    return nd.ravel()


def spheroid_mask(nx, ny, nz, xc, yc, zc, ab, c):
    """Prolate spheroid on or inside (g <= 0, FluidChannel.py:378-393),
    evaluated by broadcasting."""
    x = (np.arange(nx, dtype=np.float64) - xc)[None, None, :]
    y = (np.arange(ny, dtype=np.float64) - yc)[None, :, None]
    z = (np.arange(nz, dtype=np.float64) - zc)[:, None, None]
    out = np.empty((nz, ny, nx), bool)
    for k in range(nz):
        out[k] = (x[0] ** 2 + y[0] ** 2) / float(ab * ab) + z[k] ** 2 / float(c * c) - 1.0 <= 0.0
    return out


def timed_run(lat: Lattice, steps: int, chunk: int, tag: str) -> float:
    """Runs `steps` steps in batches, returns the kernel milliseconds."""
    ms, done = 0.0, 0
    while done < steps:
        k = min(chunk, steps - done)
        ms += lat.step_timed(k)
        done += k
        _log(f"{tag}: {done}/{steps} steps, {ms / done:.3f} ms/step")
    return ms


def _rate(N, Q, storage, steps, ms):
    mlups = N * steps / (ms * 1e3)
    return {"mlups": mlups, "ms_per_step": ms / steps,
            "hbm_gbs": mlups * 1e6 * 2 * Q * STORAGE_BYTES[storage] / 1e9,
            "bytes_per_update": 2 * Q * STORAGE_BYTES[storage]}


def _l2(a: dict, b: dict, mask) -> float:
    ua = np.stack([a[k][mask] for k in ("ux", "uy", "uz")])
    ub = np.stack([b[k][mask] for k in ("ux", "uy", "uz")])
    return float(np.linalg.norm(ua - ub) / np.linalg.norm(ub))


# ---------------------------------------------------------------------------


def porous_case(n: int = 384, seed: int = 20251202, porosity: float = 0.65, z_margin: int = 4):
    """Spheres added until the lattice (walls included) has the porosity.

    The spheres stay z_margin nodes clear of the inlet and outlet planes:
    an inlet node whose downstream neighbour is solid keeps receiving the
    Zou-He inflow with nowhere to pass it on, and its density grows without
    bound (to 6 in 60 steps on a 32^3 pack in the oracle, which follows the
    reference's rules, D3Q27LatticeStructure.hpp:112-141)."""
    solid = np.zeros((n, n, n), bool)
    solid[:, :, 0] = solid[:, :, n - 1] = solid[:, 0, :] = solid[:, n - 1, :] = True
    centres, radii = geometry.rsa_spheres(n, n, n, 6.0, 14.0, porosity, seed, solid_mask=solid,
                                          z_margin=z_margin)
    mask = geometry.sphere_pack_mask(n, n, n, centres, radii)
    nd = channel_types(mask)
    return nd, {"spheres": int(len(radii)), "porosity": float(np.mean(nd != geometry.SOLID))}


def run_porous(n: int = 384, steps: int = 5000, chunk: int = 500, seed: int = 20251202,
               ramp: int = 3000, checkpoints=(100, 500, 1000, 2000, 3000, 5000)):
    """FP64 and FP32 lattices advanced side by side (both fit in HBM), the
    velocity L2 and mass drift of FP32 against FP64 recorded at checkpoints;
    each lattice's kernel time accumulated from its own HIP events."""
    t0 = time.time()
    nd, geo = porous_case(n, seed)
    _log(f"porous {n}^3: {geo['spheres']} spheres, porosity {geo['porosity']:.4f}, "
         f"{time.time() - t0:.1f} s")
    N, Q = nd.size, 27
    fluid = nd != geometry.SOLID
    out = {"workload": f"porous_d3q27_{n}^3", "steps": steps, "seed": seed, "inlet_ramp": ramp,
           **geo, "runs": {}, "checkpoints": []}
    kw = dict(omega=1.0 / 0.505, u_lbm=0.05, cs=0.1)
    lats = {st: Lattice(27, n, n, n, nd, storage=st, **kw) for st in ("f64", "f32")}
    ms = {st: 0.0 for st in lats}
    done, every = 0, 50
    try:
        while done < steps:
            if done < ramp:
                prof = np.full(n * n, 0.05 * min(1.0, (done + every) / ramp))
                for lat in lats.values():
                    lat.set_inlet_profile(prof)
            k = min(every, steps - done)
            for st, lat in lats.items():
                ms[st] += lat.step_timed(k)
            done += k
            if done in checkpoints or done == steps:
                m64, m32 = lats["f64"].macros(), lats["f32"].macros()
                r64, r32 = lats["f64"].reduce()["mass"], lats["f32"].reduce()["mass"]
                cp = {"step": done, "l2_velocity_vs_f64": _l2(m32, m64, fluid),
                      "mass_drift_vs_f64": abs(r32 - r64) / r64,
                      "max_abs_uz": float(np.nanmax(np.abs(m64["uz"][fluid]))),
                      "rho_range": [float(np.nanmin(m64["rho"][fluid])),
                                    float(np.nanmax(m64["rho"][fluid]))]}
                out["checkpoints"].append(cp)
                _log(f"porous step {done}: {cp}, ms/step f64 {ms['f64'] / done:.3f} "
                     f"f32 {ms['f32'] / done:.3f}")
                del m64, m32
        for st, lat in lats.items():
            r = _rate(N, Q, st, steps, ms[st])
            r["mass"] = float(lat.reduce()["mass"])
            out["runs"][st] = r
    finally:
        for lat in lats.values():
            lat.close()
    last = out["checkpoints"][-1]
    out["runs"]["f32"]["l2_velocity_vs_f64"] = last["l2_velocity_vs_f64"]
    out["runs"]["f32"]["mass_drift_vs_f64"] = last["mass_drift_vs_f64"]
    _log(f"porous runs: {out['runs']}")
    return out


# ---------------------------------------------------------------------------


def run_cylinders(steps: int = 20000, chunk: int = 500,
                  checkpoints=(1100, 2000, 5000, 10000, 20000), q: int = 19, nx: int = 128,
                  perturb: float = 0.0):
    """The stream-collide benchmark's lattice over the configuration's full
    length: D3Q19 (or `q`) BGK 128x128x512, periodic x and y, a cylinder along x
    (r 12 at (y, z) = (64, 128)), Zou-He inlet u_z 0.04 and outlet rho 1,
    tau 0.55, from rest.  FP64, FP32, FP16 and FP32X (FP32 storage, FP64
    arithmetic) lattices advance side by side; the velocity L2 and mass drift
    of each against FP64, and the cylinder's drag (z momentum exchange) in
    each, are recorded at the checkpoints.  nx = 1 runs one x-row, which the
    full lattice repeats exactly for FP64 and FP32 (x is periodic, geometry
    and start uniform in x).  perturb > 0 adds an FP64 lattice whose inlet
    profile is multiplied by 1 + perturb U[-1, 1] (seeded): how far the flow
    itself carries a difference that small."""
    ny, nz = 128, 512
    nd = geometry.channel(nx, ny, nz, walls=(), obstacles=[geometry.cylinder_x(64, 128, 12)])
    N, Q = nd.size, q
    fluid = nd != geometry.SOLID
    body = nd == geometry.SOLID
    out = {"workload": f"cylinder_array_d3q{q}_{nx}x128x512", "steps": steps, "runs": {},
           "checkpoints": []}
    kw = dict(omega=1.0 / 0.55, u_lbm=0.04)
    lats = {st: Lattice(q, nx, ny, nz, nd, storage=st, **kw) for st in ("f64", "f32", "f16", "f32x")}
    compared = ["f32", "f16", "f32x"]
    if perturb > 0:
        lats["f64p"] = Lattice(q, nx, ny, nz, nd, storage="f64", **kw)
        rng = np.random.default_rng(5)
        lats["f64p"].set_inlet_profile(0.04 * (1.0 + perturb * rng.uniform(-1.0, 1.0, nx * ny)))
        compared.append("f64p")
        out["perturbation"] = perturb
    ms = {st: 0.0 for st in lats}
    done = 0
    try:
        while done < steps:
            nxt = min([c for c in checkpoints if c > done] + [steps, done + chunk])
            for st, lat in lats.items():
                ms[st] += lat.step_timed(nxt - done)
            done = nxt
            if done in checkpoints or done == steps:
                m = {st: lat.macros() for st, lat in lats.items()}
                # D3Q15's own growing mode can carry a reduced storage to NaN
                # on this flow: recorded, not raised
                red = {st: lat.reduce(allow_nonfinite=True) for st, lat in lats.items()}
                mass = {st: r["mass"] for st, r in red.items()}
                drag = {st: float(np.sum(lat.forces()["Fz"][body])) for st, lat in lats.items()}
                cp = {"step": done, "drag": drag}
                bad = [st for st, r in red.items() if not r["finite"]]
                if bad:
                    cp["nonfinite"] = bad
                for st in compared:
                    cp[f"l2_velocity_{st}_vs_f64"] = _l2(m[st], m["f64"], fluid)
                    cp[f"mass_drift_{st}_vs_f64"] = abs(mass[st] - mass["f64"]) / mass["f64"]
                out["checkpoints"].append(cp)
                _log(f"cylinders step {done}: {cp}")
                del m
        for st, lat in lats.items():
            out["runs"][st] = _rate(N, Q, "f64" if st == "f64p" else st, steps, ms[st])
    finally:
        for lat in lats.values():
            lat.close()
    _log(f"cylinders runs: {out['runs']}")
    return out


def spheroid_case(scale: int = 1):
    """(nd, spheroid mask flat) of the spheroid channel at 1/scale size."""
    nx, ny, nz = 1024 // scale, 512 // scale, 512 // scale
    mask = spheroid_mask(nx, ny, nz, nx / 2, ny / 2, nz / 2, 32 / scale, 96 / scale)
    return (nx, ny, nz), channel_types(mask), mask.ravel()


def _spheroid_run(dims, nd, body, storage, tau, steps, chunk, closure=None, cs=0.1, samples=10):
    """One run; drag = Fz on the body, final and averaged over `samples`
    instants evenly spread over the second half of the run (the wake is
    unsteady, so single instants of runs that drift apart differ)."""
    from .closure import split_closure
    nx, ny, nz = dims
    net, offset = split_closure(closure)
    with Lattice(19, nx, ny, nz, nd, storage=storage, omega=1.0 / (tau + offset), u_lbm=0.05,
                 cs=0.0 if net is not None else cs) as lat:
        if net is not None:
            lat.set_closure(*net)
        lat.init_from_macros(*uniform_start(nd, 0.05))
        tag = f"spheroid {nx}x{ny}x{nz} {storage}{'+network' if closure else ''}"
        ms = timed_run(lat, steps // 2, chunk, tag)
        gap = (steps - steps // 2) // samples
        drags = []
        for k in range(samples):
            n = gap if k < samples - 1 else steps - steps // 2 - gap * (samples - 1)
            ms += lat.step_timed(n)
            drags.append(float(np.sum(lat.forces()["Fz"][body])))
        _log(f"{tag}: {steps}/{steps} steps, {ms / steps:.3f} ms/step")
        m = lat.macros()
        red = lat.reduce()
    r = _rate(nd.size, 19, storage, steps, ms)
    r.update({"drag": drags[-1], "drag_mean": float(np.mean(drags)),
              "drag_samples": drags, "mass": red["mass"]})
    _log(f"{tag}: {r}")
    return r, m


def run_spheroid(scale: int = 1, steps: int = 10000, chunk: int = 1000,
                 network: str = DEFAULT_NETWORK, compare_down: bool = True):
    from .closure import load_network
    closure = load_network(network)
    out = {"workload": "spheroid_d3q19", "network": os.path.basename(network), "full": {},
           "downsampled": {}}
    dims, nd, body = spheroid_case(scale)
    fluid = nd != geometry.SOLID
    out["full"]["lattice"] = list(dims)
    out["full"]["steps"] = steps
    tau = 0.51
    r16, m16 = _spheroid_run(dims, nd, body, "f16", tau, steps, chunk, closure)
    r32, m32 = _spheroid_run(dims, nd, body, "f32", tau, steps, chunk)
    r16["l2_velocity_vs_f32_smagorinsky"] = _l2(m16, m32, fluid)
    r16["drag_mean_rel_diff_vs_f32_smagorinsky"] = (abs(r16["drag_mean"] - r32["drag_mean"]) /
                                                   abs(r32["drag_mean"]))
    out["full"]["f16_network"] = r16
    out["full"]["f32_smagorinsky"] = r32
    del m16, m32
    if compare_down:
        dims2, nd2, body2 = spheroid_case(2 * scale)
        fluid2 = nd2 != geometry.SOLID
        tau2, steps2 = 0.5 + (tau - 0.5) / 2.0, steps // 2
        out["downsampled"]["lattice"] = list(dims2)
        out["downsampled"]["steps"] = steps2
        r64, m64 = _spheroid_run(dims2, nd2, body2, "f64", tau2, steps2, chunk)
        out["downsampled"]["f64_smagorinsky"] = r64
        for storage, cl, key in (("f32", None, "f32_smagorinsky"), ("f16", closure, "f16_network"),
                                 ("f16", None, "f16_smagorinsky")):
            r, m = _spheroid_run(dims2, nd2, body2, storage, tau2, steps2, chunk, cl)
            r["l2_velocity_vs_f64"] = _l2(m, m64, fluid2)
            r["mass_drift_vs_f64"] = abs(r["mass"] - r64["mass"]) / r64["mass"]
            r["drag_mean_rel_diff_vs_f64"] = (abs(r["drag_mean"] - r64["drag_mean"]) /
                                              abs(r64["drag_mean"]))
            out["downsampled"][key] = r
    return out


def slab_case(gpus: int = 8, scale: int = 1):
    """One GPU's share of the spheroid lattice cut into `gpus` z-slabs: the
    slab just below the body's centre plane, (1024, 512, 512/gpus) at full
    scale, side walls and the body solid, no inlet or outlet (an interior
    slab).  Returns (dims, nd, z0)."""
    nx, ny, nz = 1024 // scale, 512 // scale, 512 // scale
    if nz % gpus:
        raise ValueError(f"{nz} planes do not split into {gpus} slabs")
    nzs = nz // gpus
    z0 = (gpus // 2 - 1) * nzs if gpus > 1 else 0
    body = spheroid_mask(nx, ny, nzs, nx / 2, ny / 2, nz / 2 - z0, 32 / scale, 96 / scale)
    nd = np.zeros((nzs, ny, nx), np.uint8)
    nd[body] = geometry.SOLID
    nd[:, :, 0] = nd[:, :, nx - 1] = geometry.SOLID
    nd[:, 0, :] = nd[:, ny - 1, :] = geometry.SOLID
    return (nx, ny, nzs), nd.ravel(), z0


def run_slab(gpus: int = 8, steps: int = 1000, chunk: int = 500, scale: int = 1,
             network: str = DEFAULT_NETWORK, link_gbs: float = 153.0):
    """The kernel on one GPU's share of the z-sharded spheroid lattice, at
    its real shape.  The slab is periodic in z (a ring of one: the step's
    own z wrap fills the ghost planes locally), which is the work each GPU
    does per step.  It is timed for FP16 + network and FP32 Smagorinsky and
    set beside the face exchange the sharded step would have to hide,
    modelled as the populations crossing a face (the 5 D3Q19 directions with
    e_z = +1, or -1) of nx ny nodes over one xGMI link at `link_gbs` (no
    second GPU runs here)."""
    from .closure import load_network, split_closure
    dims, nd, z0 = slab_case(gpus, scale)
    nx, ny, nzs = dims
    net, offset = split_closure(load_network(network))
    out = {"workload": "spheroid_slab_d3q19", "lattice": list(dims), "slabs": gpus,
           "z_range": [z0, z0 + nzs], "steps": steps, "runs": {}}
    zero = np.zeros(nd.size)
    uz = np.where(nd != geometry.SOLID, 0.05, 0.0)
    for storage, cl in (("f16", net), ("f32", None)):
        tau = 0.51 + (offset if cl is not None else 0.0)
        with Lattice(19, nx, ny, nzs, nd, storage=storage, omega=1.0 / tau, u_lbm=0.05,
                     cs=0.0 if cl is not None else 0.1) as lat:
            if cl is not None:
                lat.set_closure(*cl)
            lat.init_from_macros(zero, zero, uz, np.ones(nd.size))
            lat.step(20)
            tag = f"slab {nx}x{ny}x{nzs} {storage}{'+network' if cl is not None else ''}"
            ms = timed_run(lat, steps, chunk, tag)
        r = _rate(nd.size, 19, storage, steps, ms)
        face = 5 * nx * ny * STORAGE_BYTES[storage]
        r["face_bytes"] = face
        # modelled, not measured: the face's bytes over one link's bandwidth
        r["face_exchange_us_modelled"] = face / (link_gbs * 1e3)
        # the step's interior planes: the window the face exchange overlaps with
        r["interior_us"] = r["ms_per_step"] * 1e3 * (nzs - 2) / nzs
        out["runs"][f"{storage}_{'network' if cl is not None else 'smagorinsky'}"] = r
        _log(f"{tag}: {r}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="measured workloads")
    sub = ap.add_subparsers(dest="which", required=True)
    p = sub.add_parser("porous")
    p.add_argument("--n", type=int, default=384)
    p.add_argument("--steps", type=int, default=5000)
    p.add_argument("--seed", type=int, default=20251202)
    p.add_argument("--ramp", type=int, default=3000,
                   help="steps over which the inlet velocity rises to 0.05 (0: from the start)")
    s = sub.add_parser("spheroid")
    s.add_argument("--scale", type=int, default=1, help="lattice 1024/scale x 512/scale^2")
    s.add_argument("--steps", type=int, default=10000)
    s.add_argument("--network", default=DEFAULT_NETWORK)
    s.add_argument("--no-downsampled", action="store_true")
    c = sub.add_parser("cylinders", help="the benchmark lattice over its full 20000 steps")
    c.add_argument("--steps", type=int, default=20000)
    c.add_argument("--q", type=int, default=19, choices=(15, 19, 27), help="lattice (D3Qq)")
    c.add_argument("--nx", type=int, default=128, help="1: one x-row of the lattice")
    c.add_argument("--perturb", type=float, default=0.0,
                   help="also run FP64 with the inlet perturbed by this much (relative)")
    b = sub.add_parser("slab", help="one GPU's share of the spheroid lattice split into z-slabs")
    b.add_argument("--gpus", type=int, default=8)
    b.add_argument("--steps", type=int, default=1000)
    b.add_argument("--scale", type=int, default=1)
    b.add_argument("--network", default=DEFAULT_NETWORK)
    for q in (p, s, b, c):
        q.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.which == "porous":
        rep = run_porous(a.n, a.steps, seed=a.seed, ramp=a.ramp)
    elif a.which == "cylinders":
        rep = run_cylinders(a.steps, q=a.q, nx=a.nx, perturb=a.perturb)
    elif a.which == "slab":
        rep = run_slab(a.gpus, a.steps, scale=a.scale, network=a.network)
    else:
        rep = run_spheroid(a.scale, a.steps, network=a.network,
                           compare_down=not a.no_downsampled)
    text = json.dumps(rep, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
