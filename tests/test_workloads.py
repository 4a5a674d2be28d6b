"""The workload runners' geometry (tlbm_amd/workloads.py) on the CPU, and
short GPU runs of each runner."""
import numpy as np
import pytest

from tlbm_amd import geometry, workloads


def test_channel_types_match_generic_builder():
    nx, ny, nz = 12, 10, 14
    ob = geometry.sphere(5.5, 4.5, 7.0, 3.0)
    x, y, z = geometry.grid(nx, ny, nz)
    mask = ob(x, y, z)
    want = geometry.channel(nx, ny, nz, obstacles=[ob])
    assert np.array_equal(workloads.channel_types(mask), want)
    want_open = geometry.channel(nx, ny, nz, walls=(), obstacles=[ob])
    assert np.array_equal(workloads.channel_types(mask, walls=False), want_open)


def test_spheroid_mask_matches_generic_rule():
    nx, ny, nz = 32, 16, 16
    got = workloads.spheroid_mask(nx, ny, nz, nx / 2, ny / 2, nz / 2, 4.0, 6.0)
    x, y, z = geometry.grid(nx, ny, nz)
    want = geometry.prolate_spheroid(nx / 2, ny / 2, nz / 2, 4.0, 6.0)(x, y, z)
    assert np.array_equal(got, want) and got.any()


def test_porous_case_reaches_porosity_and_clears_the_ends():
    nd, geo = workloads.porous_case(48, seed=3)
    assert abs(geo["porosity"] - 0.65) < 0.01
    g = nd.reshape(48, 48, 48)
    # no sphere within 4 nodes of the inlet and outlet: those planes are
    # fluid apart from the side walls
    for zplane in (0, 1, 2, 3, 4, 43, 44, 45, 46, 47):
        inner = g[zplane, 1:-1, 1:-1]
        assert not np.any(inner == geometry.SOLID), zplane
    assert np.all(g[0, 1:-1, 1:-1] == geometry.INLET)
    assert np.all(g[-1, 1:-1, 1:-1] == geometry.OUTLET)


@pytest.mark.gpu
def test_porous_runner_small(gpu_lib):
    rep = workloads.run_porous(n=48, steps=300, chunk=100, seed=3, ramp=200,
                               checkpoints=(100, 300))
    for st in ("f64", "f32"):
        assert rep["runs"][st]["mlups"] > 0
    # D3Q27: FP32 follows FP64 this closely because its conserving relax aims
    # at the moments of the double weights; their float values alone would
    # add a mass source of about 1.5e-8 per step
    first = rep["checkpoints"][0]
    assert np.isfinite(first["l2_velocity_vs_f64"]) and first["l2_velocity_vs_f64"] < 1e-5
    assert first["mass_drift_vs_f64"] < 1e-7


@pytest.mark.gpu
def test_spheroid_runner_small(gpu_lib):
    rep = workloads.run_spheroid(scale=16, steps=200, chunk=100)
    full = rep["full"]
    assert full["lattice"] == [64, 32, 32]
    assert np.isfinite(full["f16_network"]["l2_velocity_vs_f32_smagorinsky"])
    down = rep["downsampled"]
    assert down["f32_smagorinsky"]["l2_velocity_vs_f64"] < 1e-4
    assert down["f16_smagorinsky"]["l2_velocity_vs_f64"] < 1e-2


def test_slab_case_geometry():
    """The 8-way slab of the quarter-scale spheroid lattice: 256x128x16,
    planes 48..63 of the full lattice, walls and the body's cross-sections
    solid, no inlet or outlet nodes."""
    (nx, ny, nz), nd, z0 = workloads.slab_case(8, scale=4)
    assert (nx, ny, nz) == (256, 128, 16) and z0 == 48
    v = nd.reshape(nz, ny, nx)
    assert set(np.unique(nd)) == {geometry.FLUID, geometry.SOLID}
    assert np.all(v[:, :, 0] == geometry.SOLID) and np.all(v[:, 0, :] == geometry.SOLID)
    full = workloads.spheroid_mask(nx, ny, 128, nx / 2, ny / 2, 64, 8.0, 24.0)
    inner = v[:, 1:-1, 1:-1] == geometry.SOLID
    assert np.array_equal(inner, full[z0:z0 + nz, 1:-1, 1:-1])
    assert inner.any()


@pytest.mark.gpu
def test_slab_runner_small(gpu_lib):
    rep = workloads.run_slab(gpus=8, steps=20, chunk=10, scale=16)
    assert rep["lattice"] == [64, 32, 4]
    for r in rep["runs"].values():
        assert r["mlups"] > 0 and r["face_exchange_us_modelled"] > 0



@pytest.mark.gpu
def test_cylinders_runner_short(gpu_lib):
    rep = workloads.run_cylinders(steps=20, chunk=10, checkpoints=(10, 20))
    assert [c["step"] for c in rep["checkpoints"]] == [10, 20]
    last = rep["checkpoints"][-1]
    assert last["l2_velocity_f32_vs_f64"] < 1e-5
    assert last["l2_velocity_f16_vs_f64"] < 1e-3
    assert last["l2_velocity_f32x_vs_f64"] < 1e-6
    assert set(rep["runs"]) == {"f64", "f32", "f16", "f32x"}


@pytest.mark.gpu
def test_cylinders_runner_row_with_perturbation(gpu_lib):
    """One x-row of the D3Q15 cylinder array with the perturbed FP64 run:
    after 20 steps the 1e-7-perturbed FP64 run and FP32 stay within 1e-5 of
    FP64 (the perturbation is not zero), FP16 within 1e-2."""
    rep = workloads.run_cylinders(steps=20, chunk=10, checkpoints=(10, 20), q=15, nx=1,
                                  perturb=1e-7)
    assert rep["workload"] == "cylinder_array_d3q15_1x128x512"
    assert set(rep["runs"]) == {"f64", "f32", "f16", "f32x", "f64p"}
    last = rep["checkpoints"][-1]
    assert 0 < last["l2_velocity_f64p_vs_f64"] < 1e-5
    assert last["l2_velocity_f32_vs_f64"] < 1e-5
    assert last["l2_velocity_f16_vs_f64"] < 1e-2
