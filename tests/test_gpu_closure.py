"""The learned closure on the GPU against its oracle.

* Wavefront inference (bf16 MFMA, the collide kernel's code) against the
  oracle's bf16 inference.
* Training kernels: loss and gradient against the FP64 oracle, one Adam
  step, and convergence on a known function.
* The closure inside the collide kernel: a constant network reproduces BGK
  with the matching relaxation time; a random network follows the oracle's
  closure run.
* Box filter against host means; the inlet profile.
"""
import numpy as np
import pytest

import lattice_tables
import pyoracle
from pyoracle import Oracle
from tlbm_amd import geometry
from tlbm_amd.gpu import FILTER_PLANES, MLP_NPARAM, Lattice, MlpTrainer, mlp_infer, mlp_init

pytestmark = pytest.mark.gpu


def _params(seed, scale=0.4):
    return np.random.default_rng(seed).normal(0, scale, MLP_NPARAM).astype(np.float32)


def test_wavefront_inference_matches_oracle(gpu_lib):
    rng = np.random.default_rng(7)
    for seed in range(3):
        p = _params(seed)
        n = 1000 + 37 * seed  # not a multiple of the wavefront
        x = rng.normal(0, 1, (n, 7)).astype(np.float32)
        z = mlp_infer(p, x).astype(np.float64)
        ref = pyoracle.mlp_infer_bf16(p, x)
        scale = np.max(np.abs(ref))
        err = np.abs(z - ref) / scale
        # FP32 accumulation order; a pre-activation within an ulp of a bf16
        # rounding tie may round h1 the other way (one bf16 step, ~0.4 %)
        assert np.mean(err < 1e-5) > 0.98, np.sort(err)[-10:]
        assert err.max() < 2e-2


def _dataset(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 7)).astype(np.float32)
    a = rng.normal(0, 0.5, 7)
    t = np.log1p(np.exp(x.astype(np.float64) @ a + 0.3)).astype(np.float32)
    return x, t


def test_training_gradient_matches_oracle(gpu_lib):
    x, t = _dataset(1, 3000)
    p = _params(2, 0.3)
    tr = MlpTrainer(x, t, p, seed=5, max_batch=4096)
    loss, g = tr.loss_grad(0, 3000)
    loss_ref, g_ref = pyoracle.mlp_loss_grad(p.astype(np.float64), x, t)
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-4
    loss2, g2 = tr.loss_grad(100, 700)
    _, g2_ref = pyoracle.mlp_loss_grad(p.astype(np.float64), x[100:800], t[100:800])
    assert np.linalg.norm(g2 - g2_ref) / np.linalg.norm(g2_ref) < 1e-4


def test_adam_step_matches_oracle(gpu_lib):
    # every sample identical: whichever indices the step draws, the
    # gradient is that of the one sample
    x = np.tile(np.array([[0.3, -1.2, 0.5, 0.9, -0.4, 0.1, 1.1]], np.float32), (512, 1))
    t = np.full(512, 0.7, np.float32)
    p = _params(9, 0.3)
    tr = MlpTrainer(x, t, p, seed=3, max_batch=1024)
    tr.step(1, batch=1024, lr=1e-3)
    q = tr.params()
    ref = p.astype(np.float64)
    _, g = pyoracle.mlp_loss_grad(ref, x[:1], t[:1])
    m, v = np.zeros(MLP_NPARAM), np.zeros(MLP_NPARAM)
    pyoracle.mlp_adam(ref, g, m, v, 1, lr=1e-3)
    moved = np.abs(g) > 1e-6  # where |g| ~ eps the float step is ill-conditioned
    assert np.allclose(q[moved], ref[moved], rtol=0, atol=2e-6)


def test_training_converges_on_known_function(gpu_lib):
    x, t = _dataset(4, 20000)
    tr = MlpTrainer(x, t, mlp_init(0), seed=1, max_batch=8192)
    first = tr.loss_grad(0, 8192)[0]
    tr.step(600, batch=8192, lr=3e-3)
    last = tr.loss_grad(0, 8192)[0]
    assert last < first / 20, (first, last)


def _channel(nx=16, ny=12, nz=24):
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(7, 6, 8, 3)])
    return nx, ny, nz, nd


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_constant_network_closure_equals_bgk(gpu_lib, storage, q):
    nx, ny, nz, nd = _channel()
    omega, c, cl_out = 1.6, 0.3, 0.05
    params = np.zeros(MLP_NPARAM, np.float32)
    params[-1] = c
    with Lattice(q, nx, ny, nz, nd, storage=storage, omega=omega, u_lbm=0.05) as a:
        a.set_closure(params, 1.0, cl_out)
        a.step(60)
        ma = a.macros()
    om2 = 1.0 / (1.0 / omega + cl_out * np.log1p(np.exp(c)))
    with Lattice(q, nx, ny, nz, nd, storage=storage, omega=om2, u_lbm=0.05) as b:
        b.step(60)
        mb = b.macros()
    fluid = nd == 0
    u_a = np.stack([ma[k][fluid] for k in ("ux", "uy", "uz")])
    u_b = np.stack([mb[k][fluid] for k in ("ux", "uy", "uz")])
    # the two runs round the relaxation time differently in the last float
    # bit; FP16 storage turns that into its own rounding noise (~1e-4)
    tol = 1e-5 if storage == "f32" else 1e-3
    assert np.linalg.norm(u_a - u_b) / np.linalg.norm(u_b) < tol


@pytest.mark.parametrize("shape", [(16, 12, 24), (13, 11, 19)])
@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_closure_run_follows_oracle(gpu_lib, storage, q, shape):
    """The network closure in the collide kernel against the oracle's FP64
    closure run (bf16 network roundings included); FP16 storage adds its own
    rounding (~1e-4 per step).  16x12x24 is nine whole 512-node workgroups
    whose interior ones take the relative gather; 13x11x19 (2717 nodes) ends
    in a partial workgroup, whose lanes past the last node run the network on
    that node and store nothing."""
    nx, ny, nz, nd = _channel(*shape)
    params = mlp_init(3)
    params[-1] = -1.0
    kw = dict(omega=1.0 / 0.51, u_lbm=0.06, rho_lbm=1.0)
    o = Oracle(q, nx, ny, nz, nd, **kw)
    o.set_closure(params, 40.0, 0.02)
    tol = 1e-4 if storage == "f32" else 2e-3
    print(f"closure run q={q} {storage} {nx}x{ny}x{nz}", flush=True)
    with Lattice(q, nx, ny, nz, nd, storage=storage, **kw) as lat:
        lat.set_closure(params, 40.0, 0.02)
        for n in (2, 30):
            lat.step(n)
            o.step(n)
            m, r = lat.macros(), o.macros()
            fluid = nd != 1  # after one step from rest only inlet/outlet nodes move
            u = np.stack([m[k][fluid] for k in ("ux", "uy", "uz")])
            ur = np.stack([r[k][fluid] for k in ("ux", "uy", "uz")])
            assert np.linalg.norm(u - ur) / np.linalg.norm(ur) < tol, n


def test_set_closure_rejects_bad_input(gpu_lib):
    """FP64 and FP32X storage, non-positive or non-finite scales and
    non-finite parameters are refused with an error; the lattice keeps running on its
    own rule and accepts a valid network afterwards."""
    from tlbm_amd.gpu import TlbmError
    nx, ny, nz, nd = _channel()
    params = mlp_init(3)
    for storage in ("f64", "f32x"):
        with Lattice(19, nx, ny, nz, nd, storage=storage, omega=1.6, u_lbm=0.05) as a:
            with pytest.raises(TlbmError, match="FP32 or FP16"):
                a.set_closure(params, 1.0, 0.01)
    with Lattice(19, nx, ny, nz, nd, storage="f32", omega=1.6, u_lbm=0.05) as b:
        for fs, os_ in ((0.0, 0.01), (1.0, -0.01), (np.inf, 0.01), (1.0, np.nan)):
            with pytest.raises(TlbmError, match="scales"):
                b.set_closure(params, fs, os_)
        bad = params.copy()
        bad[17] = np.nan
        with pytest.raises(TlbmError, match="finite"):
            b.set_closure(bad, 1.0, 0.01)
        with pytest.raises(ValueError):
            b.set_closure(params[:-1], 1.0, 0.01)
        b.step(5)
        b.set_closure(params, 1.0, 0.01)
        b.step(5)
        assert np.all(np.isfinite(b.macros()["uz"]))


def _host_stress(A, e, w, omega, cs):
    """Viscous stress -(1 - omega'/2) Pi_neq of pre-collision populations
    A (N, Q) on fluid nodes, with the Smagorinsky omega' of
    LatticeStructure.hpp:140-153 (numpy restatement)."""
    rho = A.sum(axis=1)
    u = (A @ e) / rho[:, None]
    cu = 3.0 * (u @ e.T)
    usq = (u * u).sum(axis=1)
    feq = w[None, :] * rho[:, None] * (1.0 + cu + 0.5 * cu * cu - 1.5 * usq[:, None])
    fn = A - feq
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    pi = np.stack([fn @ (e[:, i] * e[:, j]).astype(np.float64) for i, j in pairs])
    qn = np.sqrt(pi[0] ** 2 + pi[1] ** 2 + pi[2] ** 2 + 2.0 * (pi[3] ** 2 + pi[4] ** 2 + pi[5] ** 2))
    nu0 = (1.0 / omega - 0.5) / 3.0
    om = 1.0 / (3.0 * (nu0 + (np.sqrt(cs * qn + nu0 * nu0) - nu0) / 6.0) + 0.5)
    return (0.5 * om - 1.0) * pi


def test_box_filter_matches_host_means(gpu_lib):
    nx, ny, nz, nd = _channel()
    omega, cs = 1.6, 0.1
    e, _, w, _ = lattice_tables.tables(19)
    with Lattice(19, nx, ny, nz, nd, storage="f64", omega=omega, u_lbm=0.05, cs=cs) as lat:
        lat.step(24)
        A = lat.populations()  # what the next step reads
        lat.step(1)
        m = lat.macros()
        filt = lat.box_filter()

    def coarse(v):
        return v.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).mean(axis=(1, 3, 5)).ravel()

    want = {"rho": coarse(m["rho"]), "ux": coarse(m["ux"]), "uy": coarse(m["uy"]),
            "uz": coarse(m["uz"]), "uxux": coarse(m["ux"] * m["ux"]),
            "uyuy": coarse(m["uy"] * m["uy"]), "uzuz": coarse(m["uz"] * m["uz"]),
            "uxuy": coarse(m["ux"] * m["uy"]), "uxuz": coarse(m["ux"] * m["uz"]),
            "uyuz": coarse(m["uy"] * m["uz"])}
    assert set(filt) == set(FILTER_PLANES)
    for k, v in want.items():
        assert np.allclose(filt[k], v, rtol=1e-13, atol=1e-16), k
    # stresses: cells of fluid nodes only (the restatement has no boundary rules)
    sig = _host_stress(A, e, w, omega, cs)
    all_fluid = coarse((nd == 0).astype(np.float64)) == 1.0
    assert all_fluid.sum() > 50
    scale = np.abs(sig[:, nd == 0]).max()
    assert scale > 1e-6
    for k, name in enumerate(("sxx", "syy", "szz", "sxy", "sxz", "syz")):
        assert np.allclose(filt[name][all_fluid], coarse(sig[k])[all_fluid], rtol=1e-10,
                           atol=1e-12 * scale), name
    # solid nodes carry no stress
    solid_cells = coarse((nd == 1).astype(np.float64)) == 1.0
    assert solid_cells.any()
    for name in ("sxx", "syy", "szz", "sxy", "sxz", "syz"):
        assert np.all(filt[name][solid_cells] == 0.0)


def test_inlet_profile(gpu_lib):
    nx, ny, nz, nd = _channel()
    kw = dict(omega=1.6, u_lbm=0.0625)
    with Lattice(19, nx, ny, nz, nd, storage="f64", **kw) as a, \
            Lattice(19, nx, ny, nz, nd, storage="f64", **kw) as b:
        a.set_inlet_profile(np.full(nx * ny, 0.0625))  # exactly representable: same as uniform
        a.step(20)
        b.step(20)
        assert np.array_equal(a.populations(), b.populations())
    prof = 0.06 * (1.0 + 0.01 * np.random.default_rng(0).uniform(-1, 1, nx * ny))
    with Lattice(19, nx, ny, nz, nd, storage="f32", omega=1.6, u_lbm=0.06) as c:
        c.set_inlet_profile(prof)
        c.step(5)
        m = c.macros()
    inlet = np.where(nd == 2)[0]
    assert np.allclose(m["uz"][inlet], prof.astype(np.float32)[inlet % (nx * ny)], rtol=0, atol=0)
