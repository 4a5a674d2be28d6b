"""What the trained closure network gains over Smagorinsky, on the GPU.

The whole pipeline of tlbm_amd/closure.py at the learned-closure workload
(fine D3Q19 256x128x128, tau 0.502, Smagorinsky cs 0.1, 4x2 staggered
spheres r 8, inlet u_z 0.06 +/- 1 % seeded; 2x box filter to 128x64x64;
100 snapshots 100 fine steps apart: the first 50 train and calibrate, the
last 50 are held out).  Both closures are calibrated the same way on the
training window: the network's output multiplier and base offset, and
Smagorinsky's constant cs in {0.025, 0.05, 0.1} and the same base offsets
(include/LatticeStructure.hpp:140-153 is the Smagorinsky rule).

* a priori, on the held-out snapshots: the network's stress error is less
  than half of Smagorinsky's (each at its own fixed point);
* a posteriori, against the calibrated Smagorinsky:
  - coarse runs started from held-out filtered snapshots stay at least 3 %
    closer to the filtered fine run with the network after 50, 200 and 400
    coarse steps;
  - the mean velocity over the held-out window of a long coarse run from
    uniform flow is as close with either closure, within 5 %.  This does
    not claim a win: the network's lead (0.0506 against 0.0519) is smaller
    than the 4 % by which a rounding-level change of the FP32 kernels (the
    contraction of 3 rho u into the equilibrium's sums), carried through the
    whole pipeline (fine run, training, calibration), once moved the
    network's figure (to 0.0529).  A 1e-7 perturbation of the coarse
    inlet, run beside it, moves Smagorinsky's figure by only 2e-8: the
    coarse long run itself is not sensitive, the trained network is.

* cost: the network's FP32 step takes at most 1.3x the Smagorinsky step,
  on the coarse lattice and on the 128x128x512 benchmark lattice.

The run is deterministic (seeded inlet, seeded sampling, deterministic
kernels), so the margins are fixed numbers, not draws; the cost ratios are
timings (measured 1.01-1.14).
"""
import numpy as np
import pytest

from tlbm_amd import closure

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


@pytest.fixture(scope="module")
def pipeline():
    case = closure.ClosureCase()
    nd, _, data = closure.fine_run(case)
    params, cl_in, cl_out, rep = closure.train_closure(case, data)
    (m, c), _ = closure.calibrate(case, nd, data, (params, cl_in, cl_out))
    smag, _ = closure.calibrate_smagorinsky(case, nd, data)
    print("calibration", m, c, "smagorinsky", smag, flush=True)
    return case, nd, data, (params, cl_in, cl_out * m, c), smag, rep


def test_network_beats_smagorinsky_a_priori(gpu_lib, pipeline):
    *_, rep = pipeline
    ap = rep["apriori"]
    print(ap)
    assert ap["network_stress_err"] < 0.5 * ap["smagorinsky_stress_err"]
    assert ap["network_stress_err"] < ap["constant_best_stress_err"]


def test_network_beats_calibrated_smagorinsky_short_horizons(gpu_lib, pipeline):
    case, nd, data, cl, smag_cal, _ = pipeline
    nd_c = closure.coarse_geometry(nd, case)
    first = case.snapshots - case.holdout
    starts = list(range(first, first + 4))
    horizons = (1, 4, 8)
    smag = closure.short_horizon(case, nd_c, data, smag_cal, starts, horizons=horizons)
    net = closure.short_horizon(case, nd_c, data, cl, starts, horizons=horizons)
    print("calibrated smagorinsky", smag_cal, smag, "network", net)
    for h in horizons:
        assert np.isfinite(net[str(h)])
        assert net[str(h)] < 0.97 * smag[str(h)], h


def test_held_out_mean_flow_matches_calibrated_smagorinsky(gpu_lib, pipeline):
    case, nd, data, cl, smag_cal, _ = pipeline
    nd_c = closure.coarse_geometry(nd, case)
    smag = closure.statistics(case, nd_c, data, smag_cal)["heldout"]
    net = closure.statistics(case, nd_c, data, cl)["heldout"]
    noise = closure.statistics(case, nd_c, data, smag_cal, perturb=1e-7)["heldout"]
    print("calibrated smagorinsky", smag, "network", net, "smagorinsky, inlet perturbed", noise)
    assert "diverged" not in net
    assert abs(net["mean_l2"] / smag["mean_l2"] - 1.0) < 0.05


def test_network_step_cost_within_bound(gpu_lib, pipeline):
    """Per-step cost of the trained network's collide against the
    Smagorinsky collide, FP32 storage, HIP events, best of 3 x 200 steps."""
    _, _, _, cl, _, _ = pipeline
    for shape in ((128, 64, 64), (128, 128, 512)):
        cost = closure.step_cost(shape, cl, storage="f32")
        print(shape, cost, flush=True)
        assert cost["ratio"] <= 1.3, (shape, cost)
