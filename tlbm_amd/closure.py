"""The learned closure: training data, training and a-posteriori evaluation.

The reference has one closure, Smagorinsky in the form of
include/LatticeStructure.hpp:140-153 (nu_t = (sqrt(cs Qn + nu0^2) - nu0)/6,
with Qn the norm of the non-equilibrium stress).  This module builds the
learned replacement and compares it with that form:

1. A fine D3Q19 run with Smagorinsky (FP32 storage) past a staggered array
   of spheres, periodic in x and y, Zou-He inlet/outlet in z, with a seeded
   +/- 1 % perturbation of the inlet velocity, started from uniform flow
   (a start from rest rings the inlet-outlet channel's acoustic mode at
   +/- 5 % in density for thousands of steps).
2. Snapshots through the GPU 2x2x2 box filter: filtered rho, u, u_i u_j and
   viscous stress sigma = -(1 - omega/2) Pi_neq.
3. On the coarse lattice (acoustic scaling: the lattice velocity is kept,
   one coarse step is two fine steps, nu_coarse = nu_fine / 2, so
   tau0_coarse = 1/2 + (tau_fine - 1/2)/2) the stress the coarse run has to
   carry per unit density is
       T = sigma_bar / rho_bar - (bar(u_i u_j) - u_bar_i u_bar_j)
   (filtered fine viscous + Smagorinsky stress minus the sub-filter
   Reynolds stress).  Its best eddy viscosity, averaged over 3x3x3
   neighbourhoods as dynamic procedures do,
       nu* = <T : S_dev> / (2 <S_dev : S_dev>),  S from central differences
   of u_bar, S_dev its deviatoric part,
   gives the target relaxation-time increment dtau* = 3 nu* - 3 nu0, >= 0.
4. The network sees what the coarse collide kernel computes: the
   non-equilibrium stress Pi = -(2/3) rho tau S (6 components) and its
   norm Qn, each times the feature scale.  Pi depends on the relaxation
   time that produced it, so features are formed at the closure's own
   fixed point tau = tau0 + out_scale softplus(net(Pi(tau))): first at the
   Smagorinsky fixed point, then, in refit rounds, at the network's.
5. Training: the HIP trainer (Adam, mean squared error of softplus(z)
   against dtau*/out_scale).
6. Calibration, a posteriori and on the training window only: a grid over
   the output multiplier m (out_scale -> m out_scale) and a base offset c
   of the relaxation time (the coarse lattice runs at tau0 + c, the
   network adds its increment on top).  The choice minimises the sum of
   the two training-window errors relative to Smagorinsky's: short
   horizons started from training snapshots, and the mean velocity of a
   long run over the training window.
7. Evaluation on the held-out window (the second half of the snapshots):
   coarse runs with the network and with Smagorinsky (cs of the fine run)
   against the filtered fine fields -- short horizons started from
   held-out snapshots, and time statistics (mean and rms velocity) of one
   long run from uniform flow, per window -- plus the per-step cost of
   both.

Smagorinsky is calibrated the same way, on the same training window, for a
like-for-like comparison: a grid over its constant cs and the same base
offsets c.  The held-out evaluation scores the network against both the
plain (cs of the fine run, no offset) and the calibrated Smagorinsky.

A closure is None (Smagorinsky at the fine run's cs), a Smagorinsky(cs,
tau_offset), or (params, feature_scale, output_scale[, tau_offset]).

Everything runs through the C-ABI (tlbm_amd.gpu); numpy only forms the
coarse-grid targets and features.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np

from . import geometry
from .gpu import FILTER_PLANES, Lattice, MlpTrainer, mlp_infer, mlp_init

PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))  # xx, yy, zz, xy, xz, yz


@dataclasses.dataclass
class ClosureCase:
    """Fine lattice (x, y, z), flow along z; the coarse lattice is half of it."""
    nx: int = 256
    ny: int = 128
    nz: int = 128
    radius: float = 8.0
    tau_fine: float = 0.502
    cs: float = 0.1
    u_in: float = 0.06
    perturb: float = 0.01
    seed: int = 20251202
    spinup: int = 6400      # fine steps before the first snapshot
    snapshots: int = 100
    every: int = 100        # fine steps between snapshots (even)
    holdout: int = 50       # last snapshots kept out of training and calibration
    margin: int = 2         # coarse cells kept clear of solids and z ends

    @property
    def coarse(self):
        return self.nx // 2, self.ny // 2, self.nz // 2

    @property
    def tau_coarse(self) -> float:
        return 0.5 + (self.tau_fine - 0.5) / 2.0

    def sphere_centres(self):
        """4 x 2 staggered array: row A at z = 3nz/16, x = (k + 1/2) nx/4;
        row B at z = 6nz/16, x = k nx/4 (k = 0..3); all at y = ny/2."""
        za, zb = 3 * self.nz / 16.0, 6 * self.nz / 16.0
        yc = self.ny / 2.0
        a = [((k + 0.5) * self.nx / 4.0, yc, za) for k in range(4)]
        b = [(k * self.nx / 4.0, yc, zb) for k in range(4)]
        return a + b


def scaled_case(scale: float, **kw) -> ClosureCase:
    """The same flow on a lattice `scale` times the default (even sizes)."""
    base = ClosureCase()
    def ev(n):
        return max(8, int(round(n * scale / 2.0)) * 2)
    return ClosureCase(nx=ev(base.nx), ny=ev(base.ny), nz=ev(base.nz),
                       radius=base.radius * scale, **kw)


# ---------------------------------------------------------------------------
# geometry and inlet


def fine_geometry(case: ClosureCase) -> np.ndarray:
    """ndType of the fine lattice: spheres (periodic images in x and y),
    no walls, inlet z = 0 and outlet z = nz-1 outside the spheres."""
    obst = []
    for (xc, yc, zc) in case.sphere_centres():
        for ox in (-case.nx, 0, case.nx):
            for oy in (-case.ny, 0, case.ny):
                obst.append(geometry.sphere(xc + ox, yc + oy, zc, case.radius))
    return geometry.channel(case.nx, case.ny, case.nz, walls=(), obstacles=obst)


def coarse_geometry(nd_fine: np.ndarray, case: ClosureCase) -> np.ndarray:
    """Coarse cell solid when at least 4 of its 8 fine nodes are; inlet and
    outlet planes as on the fine lattice."""
    cx, cy, cz = case.coarse
    solid = (nd_fine.reshape(case.nz, case.ny, case.nx) == geometry.SOLID)
    frac = solid.reshape(cz, 2, cy, 2, cx, 2).sum(axis=(1, 3, 5))
    mask = frac >= 4

    def obstacle(x, y, z):
        return mask[z, y, x]
    return geometry.channel(cx, cy, cz, walls=(), obstacles=[obstacle])


def inlet_profiles(case: ClosureCase):
    """Fine inlet u_z(x, y) = u_in (1 + perturb U[-1, 1]) (seeded) and its
    2x2 mean for the coarse lattice."""
    rng = np.random.default_rng(case.seed)
    fine = case.u_in * (1.0 + case.perturb * rng.uniform(-1.0, 1.0, case.nx * case.ny))
    cx, cy, _ = case.coarse
    coarse = fine.reshape(cy, 2, cx, 2).mean(axis=(1, 3)).ravel()
    return fine, coarse


# ---------------------------------------------------------------------------
# coarse-grid quantities (arrays shaped (nz, ny, nx))


def strain(ux, uy, uz):
    """S_ij = (d_j u_i + d_i u_j)/2 by central differences, periodic wrap
    (z ends are excluded by the training mask).  Returns (6, ...) in PAIRS
    order."""
    u = (ux, uy, uz)
    axis = (2, 1, 0)  # d/dx, d/dy, d/dz

    def d(f, j):
        return 0.5 * (np.roll(f, -1, axis=axis[j]) - np.roll(f, 1, axis=axis[j]))
    grad = [[d(u[i], j) for j in range(3)] for i in range(3)]
    return np.stack([0.5 * (grad[i][j] + grad[j][i]) for i, j in PAIRS])


def contract(a, b):
    """A : B of symmetric tensors stored in PAIRS order."""
    return (a[0] * b[0] + a[1] * b[1] + a[2] * b[2] +
            2.0 * (a[3] * b[3] + a[4] * b[4] + a[5] * b[5]))


def box3(f):
    """Mean over the 3x3x3 neighbourhood (periodic wrap)."""
    out = f.copy()
    for ax in (0, 1, 2):
        out = (np.roll(out, 1, axis=ax) + out + np.roll(out, -1, axis=ax)) / 3.0
    return out


def stress_features(rho, S, tau):
    """Pi_neq = -(2/3) rho tau S (PAIRS order) and Qn = |Pi| as the collide
    kernel forms it: (7, ...)."""
    pi = -(2.0 / 3.0) * rho * tau * S
    qn = np.sqrt(pi[0] ** 2 + pi[1] ** 2 + pi[2] ** 2 + 2.0 * (pi[3] ** 2 + pi[4] ** 2 + pi[5] ** 2))
    return np.concatenate([pi, qn[None]], axis=0)


def smagorinsky_tau(rho, S, tau0, cs, iters: int = 30):
    """Fixed point of the reference Smagorinsky rule on a given strain:
    tau = 3 (nu0 + nu_t) + 1/2, nu_t = (sqrt(cs Qn(tau) + nu0^2) - nu0)/6
    (LatticeStructure.hpp:140-153)."""
    nu0 = (tau0 - 0.5) / 3.0
    tau = np.full_like(rho, tau0)
    for _ in range(iters):
        qn = stress_features(rho, S, tau)[6]
        nut = (np.sqrt(cs * qn + nu0 * nu0) - nu0) / 6.0
        tau = 3.0 * (nu0 + nut) + 0.5
    return tau


def snapshot_targets(filt: dict, case: ClosureCase, keep: np.ndarray):
    """(rho, S (6), dtau*, weight) on the kept coarse cells of one filtered
    snapshot.  weight = <S_dev : S_dev>: the squared stress error of a
    relaxation time is weight * (dtau - dtau*)^2 (up to a constant factor),
    so a weighted fit of dtau* fits the stress."""
    cx, cy, cz = case.coarse
    shp = (cz, cy, cx)
    g = {k: v.reshape(shp) for k, v in filt.items()}
    rho = g["rho"]
    S = strain(g["ux"], g["uy"], g["uz"])
    uu = {"xx": ("uxux", "ux", "ux"), "yy": ("uyuy", "uy", "uy"), "zz": ("uzuz", "uz", "uz"),
          "xy": ("uxuy", "ux", "uy"), "xz": ("uxuz", "ux", "uz"), "yz": ("uyuz", "uy", "uz")}
    T = []
    for comp in ("xx", "yy", "zz", "xy", "xz", "yz"):
        prod, a, b = uu[comp]
        tsgs = g[prod] - g[a] * g[b]
        T.append(g["s" + comp] / rho - tsgs)
    T = np.stack(T)
    Sd = S.copy()  # deviatoric strain: an eddy viscosity acts on it alone
    trs = (S[0] + S[1] + S[2]) / 3.0
    for k in range(3):
        Sd[k] -= trs
    num = box3(contract(T, Sd))  # T : S_dev = T_dev : S_dev
    den = box3(contract(Sd, Sd))
    nu = num / (2.0 * np.maximum(den, 1e-30))
    nu0 = (case.tau_coarse - 0.5) / 3.0
    # a total viscosity below zero is no target; outliers of tiny strain
    # carry no weight anyway
    dtau = np.clip(3.0 * (nu - nu0), -3.0 * nu0, 0.5)
    return (rho[keep].astype(np.float32), S[:, keep].astype(np.float32),
            dtau[keep].astype(np.float32), den[keep].astype(np.float32))


def training_mask(nd_fine: np.ndarray, case: ClosureCase) -> np.ndarray:
    """Coarse cells whose 8 fine nodes are all fluid, at least `margin`
    cells from any other cell and from the z ends."""
    cx, cy, cz = case.coarse
    f = (nd_fine.reshape(case.nz, case.ny, case.nx) == geometry.FLUID)
    clean = f.reshape(cz, 2, cy, 2, cx, 2).all(axis=(1, 3, 5))
    bad = ~clean
    for _ in range(case.margin):
        grown = bad.copy()
        for ax in (0, 1, 2):
            grown |= np.roll(bad, 1, axis=ax) | np.roll(bad, -1, axis=ax)
        bad = grown
    keep = ~bad
    keep[:case.margin + 1] = False
    keep[cz - case.margin - 1:] = False
    return keep


# ---------------------------------------------------------------------------
# runs


def _log(msg: str) -> None:
    print(f"[closure] {msg}", flush=True)


def fine_run(case: ClosureCase, storage: str = "f32", keep_raw: int = 0, inlet_scale=None):
    """Runs the fine lattice and returns the filtered snapshots' targets,
    the filtered velocity and density of every snapshot, and the coarse
    mask; the last `keep_raw` snapshots' filtered planes are kept whole.
    `inlet_scale` (nx*ny factors) multiplies the inlet profile."""
    nd = fine_geometry(case)
    keep = training_mask(nd, case)
    prof, _ = inlet_profiles(case)
    if inlet_scale is not None:
        prof = prof * inlet_scale
    cx, cy, cz = case.coarse
    data = {"rho": [], "S": [], "dtau": [], "weight": [], "u": [], "rho_full": [], "raw": []}
    t0 = time.time()
    with Lattice(19, case.nx, case.ny, case.nz, nd, storage=storage, omega=1.0 / case.tau_fine,
                 u_lbm=case.u_in, cs=case.cs) as lat:
        lat.set_inlet_profile(prof)
        lat.init_from_macros(*uniform_start(nd, case.u_in))
        lat.step(case.spinup)
        for k in range(case.snapshots):
            lat.step(case.every)
            filt = lat.box_filter()
            if not np.all(np.isfinite(filt["uz"])):
                raise RuntimeError(f"fine run diverged before snapshot {k}")
            rho, S, dtau, wt = snapshot_targets(filt, case, keep)
            data["rho"].append(rho)
            data["S"].append(S)
            data["dtau"].append(dtau)
            data["weight"].append(wt)
            data["u"].append(np.stack([filt["ux"], filt["uy"], filt["uz"]]).astype(np.float32))
            data["rho_full"].append(filt["rho"].astype(np.float32))
            if k >= case.snapshots - keep_raw:
                data["raw"].append(np.stack([filt[n] for n in FILTER_PLANES]).astype(np.float32))
    _log(f"fine run {case.nx}x{case.ny}x{case.nz}: {case.spinup + case.snapshots * case.every} "
         f"steps, {case.snapshots} snapshots, {time.time() - t0:.1f} s")
    return nd, keep, data


def precision_check(case: ClosureCase, perturbation: float = 1e-7) -> dict:
    """The fine run in FP32 storage (the pipeline's) against the same run in
    FP64, next to a second FP64 run whose inlet profile is perturbed by
    `perturbation` (relative, seeded), which measures how far the unsteady
    flow itself carries a difference that small.  Per snapshot: the
    relative L2 of the filtered velocity over the coarse lattice, and the
    drift of the filtered total density (a 2x2x2 mean over every node, so
    its sum is the lattice mass over 8).  Over all snapshots: the relative
    L2 of the time-mean filtered velocity."""
    _, _, d64 = fine_run(case, storage="f64")
    _, _, d32 = fine_run(case, storage="f32")
    rng = np.random.default_rng(7)
    scale = 1.0 + perturbation * rng.uniform(-1.0, 1.0, case.nx * case.ny)
    _, _, dp = fine_run(case, storage="f64", inlet_scale=scale)

    def l2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    rows = []
    for k in range(case.snapshots):
        u64 = d64["u"][k].astype(np.float64)
        m64 = float(d64["rho_full"][k].sum(dtype=np.float64))
        m32 = float(d32["rho_full"][k].sum(dtype=np.float64))
        rows.append({"snapshot": k, "step": case.spinup + (k + 1) * case.every,
                     "l2_velocity_vs_f64": l2(d32["u"][k].astype(np.float64), u64),
                     "l2_velocity_perturbed_f64": l2(dp["u"][k].astype(np.float64), u64),
                     "mass_drift_vs_f64": abs(m32 - m64) / m64})
    mean = {k: np.mean([u.astype(np.float64) for u in d[k2]], axis=0)
            for k, d, k2 in (("f64", d64, "u"), ("f32", d32, "u"), ("p", dp, "u"))}
    return {"lattice": [case.nx, case.ny, case.nz], "perturbation": perturbation,
            "mean_velocity_l2_vs_f64": l2(mean["f32"], mean["f64"]),
            "mean_velocity_l2_perturbed_f64": l2(mean["p"], mean["f64"]),
            "snapshots": rows}


def features_at(rho, S, tau, scale):
    x = stress_features(rho, S, tau) * scale
    return np.ascontiguousarray(x.T, dtype=np.float32)


def network_tau(params, rho, S, tau0, cl_in, cl_out, tau_start, iters: int = 8):
    """Fixed point tau = tau0 + cl_out softplus(net(features(tau)))."""
    tau = tau_start.astype(np.float64)
    for _ in range(iters):
        z = mlp_infer(params, features_at(rho, S, tau, cl_in)).astype(np.float64)
        tau = tau0 + cl_out * np.logaddexp(0.0, z)
    return tau


def _weighted_err(dtau_model, t, w):
    """Relative stress error sqrt(sum w (dtau - dtau*)^2 / sum w dtau*^2)."""
    return float(np.sqrt(np.sum(w * (dtau_model - t) ** 2) / np.sum(w * t * t)))


def train_closure(case: ClosureCase, data: dict, *, steps=(2000, 1500, 1500), batch=65536,
                  lr=1e-3, seed=0):
    """Trains the network on the training snapshots; returns (params, cl_in,
    cl_out, report).  Samples are drawn with probability proportional to
    their weight, so the trainer's plain mean squared error is the stress
    error."""
    ntr = case.snapshots - case.holdout
    w = np.concatenate(data["weight"][:ntr]).astype(np.float64)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(w.size, size=w.size, p=w / w.sum()))
    rho = np.concatenate(data["rho"][:ntr])[idx]
    S = np.concatenate(data["S"][:ntr], axis=1)[:, idx]
    t = np.concatenate(data["dtau"][:ntr])[idx]
    rho_v = np.concatenate(data["rho"][ntr:])
    S_v = np.concatenate(data["S"][ntr:], axis=1)
    t_v = np.concatenate(data["dtau"][ntr:]).astype(np.float64)
    w_v = np.concatenate(data["weight"][ntr:]).astype(np.float64)
    tau0 = case.tau_coarse
    tau_s = smagorinsky_tau(rho, S, tau0, case.cs)
    feats = stress_features(rho, S, tau_s)
    cl_in = float(1.0 / np.sqrt(np.mean(feats[:6] ** 2)))
    cl_out = float(max(np.mean(t), 1e-4))
    params = mlp_init(seed)
    report = {"samples": int(t.size), "distinct_samples": int(np.unique(idx).size),
              "holdout_samples": int(t_v.size), "cl_in": cl_in, "cl_out": cl_out,
              "dtau_mean_weighted": float(np.mean(t)), "dtau_std_weighted": float(np.std(t)),
              "rounds": []}
    tau = tau_s
    for r, n in enumerate(steps):
        if r > 0:
            tau = network_tau(params, rho, S, tau0, cl_in, cl_out, tau)
        x = features_at(rho, S, tau, cl_in)
        t0 = time.time()
        with MlpTrainer(x, (t / cl_out).astype(np.float32), params, seed=seed + r,
                        max_batch=batch) as tr:
            tr.step(n, batch=batch, lr=lr)
            params = tr.params()
            loss = tr.loss_grad(0, min(t.size, batch))[0]
        report["rounds"].append({"steps": n, "train_loss": float(loss),
                                 "seconds": time.time() - t0})
        _log(f"round {r}: {n} Adam steps, loss {loss:.4g}, {time.time() - t0:.1f} s")
    # a-priori stress error on the held-out snapshots, each closure at its
    # own fixed point
    tau_sv = smagorinsky_tau(rho_v, S_v, tau0, case.cs)
    tau_nv = network_tau(params, rho_v, S_v, tau0, cl_in, cl_out, tau_sv, iters=12)
    sw = w_v.sum()
    report["apriori"] = {
        "smagorinsky_stress_err": _weighted_err(tau_sv - tau0, t_v, w_v),
        "network_stress_err": _weighted_err(tau_nv - tau0, t_v, w_v),
        "constant_best_stress_err": _weighted_err(np.full_like(t_v, np.sum(w_v * t_v) / sw), t_v, w_v),
        "target_mean_dtau": float(np.sum(w_v * t_v) / sw),
        "smagorinsky_mean_dtau": float(np.sum(w_v * (tau_sv - tau0)) / sw),
        "network_mean_dtau": float(np.sum(w_v * (tau_nv - tau0)) / sw),
    }
    _log(f"a priori: {report['apriori']}")
    return params, cl_in, cl_out, report


@dataclasses.dataclass(frozen=True)
class Smagorinsky:
    """The reference's closure (include/LatticeStructure.hpp:140-153) with
    its own constant and a base relaxation-time offset: the coarse lattice
    runs at tau0 + tau_offset with Smagorinsky cs on top."""
    cs: float
    tau_offset: float = 0.0


def split_closure(closure):
    """(network (params, feature_scale, output_scale) or None, tau offset)."""
    if closure is None:
        return None, 0.0
    if isinstance(closure, Smagorinsky):
        return None, float(closure.tau_offset)
    return tuple(closure[:3]), (float(closure[3]) if len(closure) > 3 else 0.0)


def closure_cs(case: ClosureCase, closure) -> float:
    """Smagorinsky constant of a closure's run (0 under a network)."""
    if closure is None:
        return case.cs
    if isinstance(closure, Smagorinsky):
        return float(closure.cs)
    return 0.0


class CoarseRun:
    """A coarse lattice with either closure; a network closure's tau offset
    shifts the lattice's base relaxation time."""

    def __init__(self, case: ClosureCase, nd_coarse, closure, storage="f32", inlet_scale=None):
        cx, cy, cz = case.coarse
        self.case = case
        self.nd = nd_coarse
        self.fluid = nd_coarse != geometry.SOLID
        _, prof = inlet_profiles(case)
        if inlet_scale is not None:
            prof = prof * inlet_scale
        net, offset = split_closure(closure)
        self.lat = Lattice(19, cx, cy, cz, nd_coarse, storage=storage,
                           omega=1.0 / (case.tau_coarse + offset), u_lbm=case.u_in,
                           cs=closure_cs(case, closure))
        self.lat.set_inlet_profile(prof)
        if net is not None:
            self.lat.set_closure(*net)

    def start_from(self, u, rho):
        u = np.where(self.fluid, u, 0.0)
        self.lat.init_from_macros(u[0], u[1], u[2], rho)

    def velocity(self):
        m = self.lat.macros()
        return np.stack([m["ux"], m["uy"], m["uz"]])

    def close(self):
        self.lat.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _l2(a, b, mask):
    d = (a - b)[:, mask]
    return float(np.linalg.norm(d) / np.linalg.norm(b[:, mask]))


def deplane(u, fluid, case: ClosureCase):
    """u (3, Nc) minus, per component and z plane, its mean over the plane's
    fluid cells: removes plane acoustic modes of the inlet-outlet channel,
    leaves the vortical field."""
    cx, cy, cz = case.coarse
    v = u.reshape(3, cz, cy * cx)
    m = fluid.reshape(cz, cy * cx)
    cnt = np.maximum(m.sum(axis=1), 1)
    mean = (v * m).sum(axis=2) / cnt
    return (v - mean[:, :, None]).reshape(3, -1)


def uniform_start(nd: np.ndarray, u_in: float):
    """rho = 1, u = (0, 0, u_in) on every non-solid node, 0 on solids."""
    uz = np.where(nd != geometry.SOLID, u_in, 0.0)
    z = np.zeros(nd.size)
    return z, z, uz, np.ones(nd.size)


def short_horizon(case: ClosureCase, nd_c, data, closure, starts, horizons=(1, 4),
                  storage="f32", u_f=None):
    """Coarse runs started from the filtered snapshots `starts` (filtered rho
    and u), compared with the snapshot h intervals later: mean relative L2
    of u over fluid cells, per horizon."""
    fluid = nd_c == geometry.FLUID
    if u_f is None:
        u_f = np.stack(data["u"]).astype(np.float64)
    half = case.every // 2
    errs = {h: [] for h in horizons}
    with CoarseRun(case, nd_c, closure, storage) as run:
        for j in starts:
            run.start_from(u_f[j], data["rho_full"][j])
            done = 0
            for h in horizons:
                run.lat.step((h - done) * half)
                done = h
                errs[h].append(_l2(run.velocity(), u_f[j + h], fluid))
    return {str(h): float(np.mean(errs[h])) for h in horizons}


def statistics(case: ClosureCase, nd_c, data, closure, storage="f32", u_f=None,
               perturb: float = 0.0):
    """One coarse run from uniform flow over the fine run's time window
    (acoustic scaling), sampled at the snapshot times.  Per window --
    "train" (the training snapshots) and "heldout" (the rest) -- the
    relative L2 of the mean velocity and of the rms of the vortical
    fluctuation (deplane) against the filtered fine run over the same
    snapshots.  perturb > 0 multiplies the coarse inlet profile by
    1 + perturb U[-1, 1] (seeded): how far the flow itself moves these
    statistics for a change that small."""
    fluid = nd_c == geometry.FLUID
    if u_f is None:
        u_f = np.stack(data["u"]).astype(np.float64)
    ntr = case.snapshots - case.holdout
    windows = {"train": (0, ntr), "heldout": (ntr, case.snapshots)}
    half = case.every // 2
    acc = {w: [0.0, 0.0, 0.0] for w in windows}
    scale = None
    if perturb > 0:
        cx, cy, _ = case.coarse
        scale = 1.0 + perturb * np.random.default_rng(11).uniform(-1.0, 1.0, cx * cy)
    with CoarseRun(case, nd_c, closure, storage, inlet_scale=scale) as run:
        run.lat.init_from_macros(*uniform_start(nd_c, case.u_in))
        run.lat.step(case.spinup // 2)
        for k in range(case.snapshots):
            run.lat.step(half)
            u = run.velocity()
            if not np.all(np.isfinite(u)):
                bad = {"mean_l2": float("inf"), "rms_vort_l2": float("inf"),
                       "rms_vort_ratio": float("inf"), "diverged": True}
                return {w: dict(bad) for w in windows}
            v = deplane(u, fluid, case)
            w = "train" if k < ntr else "heldout"
            a = acc[w]
            a[0] = a[0] + u
            a[1] = a[1] + v
            a[2] = a[2] + v * v
    out = {}
    for w, (j0, j1) in windows.items():
        n = j1 - j0
        if n == 0:
            continue
        uw = u_f[j0:j1]
        v_rms = np.stack([deplane(u, fluid, case) for u in uw]).std(axis=0)
        s1, v1, v2 = acc[w]
        vm = v1 / n
        rms = np.sqrt(np.maximum(v2 / n - vm * vm, 0.0))
        out[w] = {"mean_l2": _l2(s1 / n, uw.mean(axis=0), fluid),
                  "rms_vort_l2": _l2(rms, v_rms, fluid),
                  "rms_vort_ratio": float(np.linalg.norm(rms[:, fluid]) /
                                          np.linalg.norm(v_rms[:, fluid]))}
    return out


def _finite(x: float) -> float:
    return x if np.isfinite(x) else float("inf")


def calibrate(case: ClosureCase, nd_fine, data, closure, multipliers=(0.25, 0.5, 1.0),
              offsets=(0.0, 0.004, 0.008, 0.012), starts=4, horizons=(1, 4, 8),
              storage="f32"):
    """A-posteriori choice of (m, c) on the TRAINING window: output_scale ->
    m output_scale and base relaxation time tau0 + c.  Score = short-horizon
    L2 (mean over horizons, runs started from training snapshots) over
    Smagorinsky's + training-window mean-velocity L2 over Smagorinsky's;
    runs that diverge are out.  Returns ((m, c), table)."""
    score, ref, js = _training_scorer(case, nd_fine, data, starts, horizons, storage)
    net, _ = split_closure(closure)
    params, cl_in, cl_out = net
    table = {"smagorinsky": {"short": ref[0], "mean": ref[1]}}
    best, best_score = None, float("inf")
    for m in multipliers:
        for c in offsets:
            sh, mean = score((params, cl_in, cl_out * m, c))
            sc = sh / ref[0] + mean / ref[1]
            table[f"{m}/{c}"] = {"short": sh, "mean": mean, "score": sc}
            _log(f"calibration m {m} c {c}: short {sh:.4g} mean {mean:.4g} score {sc:.4g}")
            if sc < best_score:
                best, best_score = (m, c), sc
    if best is None:
        raise RuntimeError("every calibrated network run diverged")
    _log(f"calibration on training window (starts {js}): smagorinsky short {ref[0]:.4g} "
         f"mean {ref[1]:.4g} -> m {best[0]} c {best[1]} (score {best_score:.4g})")
    return best, table


def _training_scorer(case: ClosureCase, nd_fine, data, starts, horizons, storage):
    """The calibration score's parts on the TRAINING window -- mean
    short-horizon L2 (runs started from `starts` training snapshots) and
    the training-window mean-velocity L2 of a long run -- as a function of
    the closure, with the plain Smagorinsky run's values and the starts."""
    nd_c = coarse_geometry(nd_fine, case)
    u_f = np.stack(data["u"]).astype(np.float64)
    ntr = case.snapshots - case.holdout
    js = [int(j) for j in np.linspace(0, ntr - 1 - max(horizons), starts).astype(int)]

    def score(cl):
        sh = short_horizon(case, nd_c, data, cl, js, horizons=horizons, storage=storage,
                           u_f=u_f)
        st = statistics(case, nd_c, data, cl, storage=storage, u_f=u_f)["train"]
        return _finite(float(np.mean(list(sh.values())))), _finite(st["mean_l2"])

    return score, score(None), js


def calibrate_smagorinsky(case: ClosureCase, nd_fine, data, cs_values=(0.025, 0.05, 0.1),
                          offsets=(0.0, 0.004, 0.008, 0.012), starts=4, horizons=(1, 4, 8),
                          storage="f32"):
    """The network's calibration applied to Smagorinsky: a grid over its
    constant cs and the same base offsets c, the same training-window score
    (each part relative to the plain Smagorinsky run's).  Returns
    (Smagorinsky(cs, c), table)."""
    score, ref, js = _training_scorer(case, nd_fine, data, starts, horizons, storage)
    table = {"smagorinsky": {"short": ref[0], "mean": ref[1]}}
    best, best_score = None, float("inf")
    for cs in cs_values:
        for c in offsets:
            sh, mean = score(Smagorinsky(cs, c))
            sc = sh / ref[0] + mean / ref[1]
            table[f"cs{cs}/{c}"] = {"short": sh, "mean": mean, "score": sc}
            _log(f"smagorinsky calibration cs {cs} c {c}: short {sh:.4g} mean {mean:.4g} "
                 f"score {sc:.4g}")
            if sc < best_score:
                best, best_score = Smagorinsky(cs, c), sc
    if best is None:
        raise RuntimeError("every calibrated Smagorinsky run diverged")
    _log(f"smagorinsky calibration (starts {js}) -> cs {best.cs} c {best.tau_offset} "
         f"(score {best_score:.4g})")
    return best, table


def evaluate(case: ClosureCase, nd_fine, data, closures: dict, *, starts=4,
             horizons=(1, 4, 8), storage="f32"):
    """Scores each closure (name -> closure) against the filtered fine run:
    short horizons from held-out snapshots, and the long-run statistics per
    window."""
    nd_c = coarse_geometry(nd_fine, case)
    fluid = nd_c == geometry.FLUID
    u_f = np.stack(data["u"]).astype(np.float64)
    first = case.snapshots - case.holdout
    js = list(range(first, case.snapshots - max(horizons)))[:starts]
    out = {"starts": [int(j) for j in js],
           "persistence": {str(h): float(np.mean([_l2(u_f[j], u_f[j + h], fluid) for j in js]))
                           for h in horizons},
           # how far the fine run's own mean moves between the two windows
           "window_mean_shift": _l2(u_f[first:].mean(axis=0), u_f[:first].mean(axis=0), fluid)}
    for name, cl in closures.items():
        out[name] = {"short": short_horizon(case, nd_c, data, cl, js, horizons=horizons,
                                            storage=storage, u_f=u_f),
                     "stats": statistics(case, nd_c, data, cl, storage=storage, u_f=u_f)}
        _log(f"{name}: {out[name]}")
    # the flow's own sensitivity: Smagorinsky with its inlet perturbed by 1e-7
    out["smagorinsky_inlet_perturbed_1e-7"] = {
        "stats": statistics(case, nd_c, data, None, storage=storage, u_f=u_f, perturb=1e-7)}
    _log(f"smagorinsky, inlet perturbed by 1e-7: {out['smagorinsky_inlet_perturbed_1e-7']}")
    return out


def step_cost(shape, closure, storage="f32", steps=200, warmup=20, repeats=3, q=19):
    """Milliseconds per step of the Smagorinsky and the network collide on a
    D3Q`q` lattice of `shape` (sphere-free channel, periodic x/y)."""
    nx, ny, nz = shape
    nd = geometry.channel(nx, ny, nz, walls=())
    res = {}
    for name in ("smagorinsky", "network"):
        cs = 0.1 if name == "smagorinsky" else 0.0
        with Lattice(q, nx, ny, nz, nd, storage=storage, omega=1.0 / 0.501, u_lbm=0.06,
                     cs=cs) as lat:
            if name == "network":
                lat.set_closure(*split_closure(closure)[0])
            lat.step(warmup)
            lat.sync()
            res[name] = min(lat.step_timed(steps) / steps for _ in range(repeats))
    res["ratio"] = res["network"] / res["smagorinsky"]
    return res


def run_pipeline(case: ClosureCase, *, train_steps=(2000, 1500, 1500), cost_shapes=(),
                 storage="f32", dump=None, dump_snapshots=8):
    t0 = time.time()
    nd, keep, data = fine_run(case, keep_raw=dump_snapshots if dump else 0)
    if dump:
        u = np.stack(data["u"])
        np.savez_compressed(dump, nd=nd, keep=keep, raw=np.stack(data["raw"]),
                            u_mean=u.mean(axis=0), u_rms=u.std(axis=0),
                            planes=np.array(FILTER_PLANES))
    params, cl_in, cl_out, rep = train_closure(case, data, steps=train_steps)
    (m, c), table = calibrate(case, nd, data, (params, cl_in, cl_out), storage=storage)
    rep["calibration"] = {"multiplier": m, "tau_offset": c, "table": table}
    smag, stable = calibrate_smagorinsky(case, nd, data, storage=storage)
    rep["smagorinsky_calibration"] = {"cs": smag.cs, "tau_offset": smag.tau_offset,
                                      "table": stable}
    closure = (params, cl_in, cl_out * m, c)
    closures = {"smagorinsky": None, "smagorinsky_calibrated": smag, "network": closure,
                "network_uncalibrated": (params, cl_in, cl_out)}
    ev = evaluate(case, nd, data, closures, storage=storage)
    cost = {}
    for shp in cost_shapes:
        for st in ("f32", "f16"):
            cost[f"{shp[0]}x{shp[1]}x{shp[2]}_{st}"] = step_cost(shp, closure, storage=st)
    return {"case": dataclasses.asdict(case), "coarse": list(case.coarse),
            "tau_coarse": case.tau_coarse, "training": rep, "evaluation": ev, "cost": cost,
            "params": params.tolist(), "feature_scale": cl_in, "output_scale": cl_out * m,
            "tau_offset": c, "seconds": time.time() - t0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--scale", type=float, default=1.0, help="lattice size relative to 256x128x128")
    ap.add_argument("--snapshots", type=int, default=100,
                    help="snapshots of the fine run; the second half is held out")
    ap.add_argument("--train-steps", type=int, nargs="+", default=[2000, 1500, 1500])
    ap.add_argument("--out", default=None, help="write the JSON report here")
    ap.add_argument("--cost", action="store_true", help="also time both closures")
    ap.add_argument("--dump", default=None,
                    help="write the last --dump-snapshots filtered snapshots (npz) here")
    ap.add_argument("--dump-snapshots", type=int, default=8)
    ap.add_argument("--save", default=None,
                    help="write the trained network (params, feature and output scale) as JSON")
    ap.add_argument("--precision", action="store_true",
                    help="only compare the fine run in FP32 with the same run in FP64")
    ap.add_argument("--cost-only", default=None, metavar="NETWORK",
                    help="only time the saved network's collide against Smagorinsky "
                         "(128x128x512, FP32 and FP16)")
    ap.add_argument("--lattices", default="19",
                    help="with --cost-only: the lattices to time, e.g. 15,19,27")
    a = ap.parse_args(argv)
    if a.precision:
        rep = precision_check(ClosureCase())
        if a.out:
            with open(a.out, "w") as fh:
                json.dump(rep, fh)
        print(json.dumps({k: v for k, v in rep.items() if k != "snapshots"}, indent=1))
        print(json.dumps(rep["snapshots"][:: max(1, len(rep["snapshots"]) // 10)], indent=1))
        return 0
    if a.cost_only:
        net = load_network(a.cost_only)
        cost = {}
        for q in (int(v) for v in a.lattices.split(",")):
            prefix = "" if q == 19 else f"d3q{q}_"
            for st in ("f32", "f16"):
                cost[f"{prefix}128x128x512_{st}"] = step_cost((128, 128, 512), net, storage=st, q=q)
        print(json.dumps(cost, indent=1))
        return 0
    kw = {"snapshots": a.snapshots, "holdout": a.snapshots // 2}
    case = scaled_case(a.scale, **kw) if a.scale != 1.0 else ClosureCase(**kw)
    shapes = [case.coarse, (128, 128, 512)] if a.cost else []
    rep = run_pipeline(case, train_steps=tuple(a.train_steps), cost_shapes=shapes,
                       dump=a.dump, dump_snapshots=a.dump_snapshots)
    text = json.dumps(rep)
    if a.save:
        net = {k: rep[k] for k in ("params", "feature_scale", "output_scale", "tau_offset",
                                   "case")}
        net["calibration"] = rep["training"]["calibration"]
        net["apriori"] = rep["training"]["apriori"]
        with open(a.save, "w") as fh:
            json.dump(net, fh)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text)
    summary = {k: rep[k] for k in ("coarse", "evaluation", "cost", "seconds")}
    summary["apriori"] = rep["training"]["apriori"]
    summary["calibration"] = rep["training"]["calibration"]
    summary["smagorinsky_calibration"] = rep["training"]["smagorinsky_calibration"]
    print(json.dumps(summary, indent=1))
    return 0


def load_network(path: str):
    """(params, feature_scale, output_scale, tau_offset) of a network saved
    by --save; the lattice runs at tau0 + tau_offset under it."""
    with open(path) as fh:
        net = json.load(fh)
    return (np.asarray(net["params"], np.float32), float(net["feature_scale"]),
            float(net["output_scale"]), float(net.get("tau_offset", 0.0)))


if __name__ == "__main__":
    sys.exit(main())
