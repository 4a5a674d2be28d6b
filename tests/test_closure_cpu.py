"""The closure pipeline's host-side pieces (tlbm_amd/closure.py) on the CPU.

* Central-difference strain of a periodic field against its closed form.
* The training target recovers a planted eddy viscosity, whether the
  stress comes from the filtered viscous planes or from the sub-filter
  Reynolds stress, and ignores stress orthogonal to the strain.
* The Smagorinsky fixed point satisfies the reference rule
  (LatticeStructure.hpp:140-153) and reduces to tau0 at cs = 0.
* The feature formula Pi_neq = -(2/3) rho tau S is what the lattice itself
  produces: a decaying shear wave run on the oracle.
* Geometry: the staggered sphere array, the coarse majority rule, the
  training mask and the inlet profiles.
"""
import numpy as np
import pytest

import lattice_tables
from pyoracle import Oracle
from tlbm_amd import closure, geometry


def _wave_field(shape, seed=0):
    """A smooth velocity periodic in all three axes, shaped (nz, ny, nx)."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        a = rng.normal(0, 0.01, 3)
        out.append(a[0] * np.sin(2 * np.pi * x / nx + a[1]) * np.cos(2 * np.pi * y / ny) +
                   a[2] * np.sin(2 * np.pi * z / nz))
    return out, (x, y, z)


def test_strain_matches_closed_form():
    shape = (16, 12, 20)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky, kz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    ux = np.sin(kz * z)
    uy = np.cos(kx * x)
    uz = np.sin(ky * y) + np.sin(kx * x)
    S = closure.strain(ux, uy, uz)
    # central difference of sin(k n): cos(k n) sin(k)
    d = lambda k, v: np.sin(k) * v  # noqa: E731
    want = {
        3: 0.5 * d(kx, -np.sin(kx * x)),                     # xy: (dy ux + dx uy)/2
        4: 0.5 * (d(kz, np.cos(kz * z)) + d(kx, np.cos(kx * x))),  # xz
        5: 0.5 * d(ky, np.cos(ky * y)),                       # yz
    }
    assert np.allclose(S[0], 0) and np.allclose(S[1], 0) and np.allclose(S[2], 0)
    for k, v in want.items():
        assert np.allclose(S[k], v, atol=1e-14), k


def _filtered(u, rho, S, nu_visc, nu_sgs, extra_iso=0.0):
    """Filtered planes whose viscous stress is rho 2 nu_visc S (+ an
    isotropic part) and whose sub-filter stress is -2 nu_sgs S."""
    ux, uy, uz = u
    f = {"rho": rho, "ux": ux, "uy": uy, "uz": uz}
    names = {(0, 0): ("uxux", ux, ux), (1, 1): ("uyuy", uy, uy), (2, 2): ("uzuz", uz, uz),
             (0, 1): ("uxuy", ux, uy), (0, 2): ("uxuz", ux, uz), (1, 2): ("uyuz", uy, uz)}
    for k, ij in enumerate(closure.PAIRS):
        name, a, b = names[ij]
        f[name] = a * b - 2.0 * nu_sgs * S[k]
        iso = extra_iso if ij[0] == ij[1] else 0.0
        f["s" + "xyz"[ij[0]] + "xyz"[ij[1]]] = rho * (2.0 * nu_visc * S[k] + iso)
    return {k: np.asarray(v, np.float64).ravel() for k, v in f.items()}


@pytest.mark.parametrize("nu_visc,nu_sgs,iso", [(0.004, 0.0, 0.0), (0.0, 0.003, 0.0),
                                                (0.002, 0.001, 0.05)])
def test_target_recovers_planted_eddy_viscosity(nu_visc, nu_sgs, iso):
    case = closure.ClosureCase(nx=20, ny=12, nz=16, margin=1)
    u, _ = _wave_field((case.nz // 2, case.ny // 2, case.nx // 2))
    rho = 1.0 + 0.01 * u[0]
    S = closure.strain(*u)
    filt = _filtered(u, rho, S, nu_visc, nu_sgs, iso)
    keep = np.ones(rho.shape, bool)
    r, s, dtau, wt = closure.snapshot_targets(filt, case, keep)
    nu0 = (case.tau_coarse - 0.5) / 3.0
    want = 3.0 * (nu_visc + nu_sgs - nu0)
    assert np.allclose(dtau, want, rtol=1e-5, atol=1e-7)
    assert np.allclose(r, rho[keep]) and s.shape == (6, keep.sum())
    # weight: the 3x3x3 mean of S_dev : S_dev
    Sd = S.copy()
    tr = (S[0] + S[1] + S[2]) / 3.0
    for k in range(3):
        Sd[k] -= tr
    sd2 = (Sd[:3] ** 2).sum(axis=0) + 2.0 * (Sd[3:] ** 2).sum(axis=0)
    assert np.allclose(wt, closure.box3(sd2)[keep], rtol=1e-5)


def test_target_is_clipped_at_zero_total_viscosity():
    case = closure.ClosureCase(nx=20, ny=12, nz=16, margin=1)
    u, _ = _wave_field((case.nz // 2, case.ny // 2, case.nx // 2), seed=3)
    rho = np.ones_like(u[0])
    S = closure.strain(*u)
    filt = _filtered(u, rho, S, 0.0, -0.002)  # backscatter only
    _, _, dtau, _ = closure.snapshot_targets(filt, case, np.ones(rho.shape, bool))
    nu0 = (case.tau_coarse - 0.5) / 3.0
    assert np.allclose(dtau, -3.0 * nu0)


def test_deplane_removes_plane_means_only():
    case = closure.ClosureCase(nx=16, ny=8, nz=12)
    cx, cy, cz = case.coarse
    rng = np.random.default_rng(2)
    fluid = rng.uniform(size=cx * cy * cz) > 0.1
    vort = rng.normal(size=(3, cx * cy * cz))
    v0 = closure.deplane(vort, fluid, case)
    plane = np.repeat(rng.normal(size=(3, cz)), cx * cy, axis=1)  # a plane acoustic mode
    v1 = closure.deplane(vort + plane, fluid, case)
    assert np.allclose(v0, v1)
    means = (v0.reshape(3, cz, -1) * fluid.reshape(cz, -1)).sum(axis=2)
    assert np.allclose(means, 0.0)


def test_smagorinsky_fixed_point_satisfies_reference_rule():
    u, _ = _wave_field((8, 6, 10), seed=1)
    rho = np.ones_like(u[0])
    S = closure.strain(*u) * 5.0
    tau0, cs = 0.501, 0.1
    tau = closure.smagorinsky_tau(rho, S, tau0, cs)
    nu0 = (tau0 - 0.5) / 3.0
    qn = closure.stress_features(rho, S, tau)[6]
    rhs = 3.0 * (nu0 + (np.sqrt(cs * qn + nu0 * nu0) - nu0) / 6.0) + 0.5
    assert np.allclose(tau, rhs, rtol=0, atol=1e-12)
    assert np.all(tau > tau0)
    assert np.array_equal(closure.smagorinsky_tau(rho, S, tau0, 0.0), np.full_like(rho, tau0))


def test_feature_formula_matches_lattice_shear_wave():
    """Pi_neq of the populations the collide reads equals -(2/3) rho tau S
    for a decaying shear wave u_x = U sin(2 pi z / nz) (periodic box)."""
    nx, ny, nz = 4, 4, 64
    tau, U = 0.8, 1e-3
    e, _, w, _ = lattice_tables.tables(19)
    nd = geometry.periodic_box(nx, ny, nz)
    o = Oracle(19, nx, ny, nz, nd, omega=1.0 / tau, u_lbm=0.0, exact_weights=True)
    z = np.repeat(np.arange(nz), nx * ny)
    ux = U * np.sin(2 * np.pi * z / nz)
    cu = 3.0 * np.outer(ux, e[:, 0])
    f0 = w[None, :] * (1.0 + cu + 0.5 * cu * cu - 1.5 * (ux * ux)[:, None])
    o.set_populations(f0)
    o.step(40)
    A = o.populations()
    rho = A.sum(axis=1)
    u = (A @ e) / rho[:, None]
    cu = 3.0 * (u @ e.T)
    feq = w[None, :] * rho[:, None] * (1.0 + cu + 0.5 * cu * cu - 1.5 * (u * u).sum(axis=1)[:, None])
    pi_xz = (A - feq) @ (e[:, 0] * e[:, 2]).astype(np.float64)
    shp = (nz, ny, nx)
    S = closure.strain(u[:, 0].reshape(shp), u[:, 1].reshape(shp), u[:, 2].reshape(shp))
    want = closure.stress_features(rho.reshape(shp), S, tau)[4].ravel()
    # agreement up to the O(k^2) of the central difference and the lattice
    rel = np.linalg.norm(pi_xz - want) / np.linalg.norm(want)
    assert rel < 0.02, rel
    # a dropped factor (tau or 2/3) would be far outside
    assert np.linalg.norm(pi_xz - want / tau) / np.linalg.norm(want) > 0.2


def test_case_geometry():
    case = closure.ClosureCase(nx=64, ny=32, nz=32, radius=2.0)
    nd = closure.fine_geometry(case)
    g = nd.reshape(case.nz, case.ny, case.nx)
    # 8 spheres, two rows; the x = 0 sphere of row B wraps round in x
    for (xc, yc, zc) in case.sphere_centres():
        assert g[int(zc), int(yc), int(xc) % case.nx] == geometry.SOLID
    assert g[int(6 * case.nz / 16), int(case.ny / 2), case.nx - 1] == geometry.SOLID
    assert np.all(g[0][g[0] != geometry.SOLID] == geometry.INLET)
    assert np.all(g[-1] == geometry.OUTLET)
    nc = closure.coarse_geometry(nd, case)
    cx, cy, cz = case.coarse
    solid = (g == geometry.SOLID).reshape(cz, 2, cy, 2, cx, 2).sum(axis=(1, 3, 5)) >= 4
    assert np.array_equal(nc.reshape(cz, cy, cx) == geometry.SOLID, solid)
    keep = closure.training_mask(nd, case)
    assert keep.any() and not keep[solid].any()
    assert not keep[:case.margin + 1].any() and not keep[cz - case.margin - 1:].any()
    fine, coarse = closure.inlet_profiles(case)
    assert np.all(np.abs(fine / case.u_in - 1.0) <= case.perturb)
    assert coarse.size == cx * cy
    assert coarse[0] == pytest.approx(fine.reshape(case.ny, case.nx)[:2, :2].mean())
