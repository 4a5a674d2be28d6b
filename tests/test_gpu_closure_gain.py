"""What the trained closure network gains over Smagorinsky, on the GPU.

The whole pipeline of tlbm_amd/closure.py at the learned-closure workload
(fine D3Q19 256x128x128, tau 0.502, Smagorinsky cs 0.1, 4x2 staggered
spheres r 8, inlet u_z 0.06 +/- 1 % seeded; 2x box filter to 128x64x64;
100 snapshots 100 fine steps apart: the first 50 train and calibrate the
network, the last 50 are held out):

* a priori, on the held-out snapshots: the network's stress error is less
  than half of Smagorinsky's (each at its own fixed point);
* a posteriori, with the calibration (output multiplier m, base offset c)
  chosen on the training window alone:
  - coarse runs started from held-out filtered snapshots stay closer to the
    filtered fine run with the network than with Smagorinsky after 50, 200
    and 400 coarse steps;
  - a long coarse run from uniform flow has a mean velocity over the
    held-out window closer to the filtered fine run's with the network.

The run is deterministic (seeded inlet, seeded sampling, deterministic
kernels), so the margins are fixed numbers, not draws.
"""
import numpy as np
import pytest

from tlbm_amd import closure

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.fixture(scope="module")
def pipeline():
    case = closure.ClosureCase()
    nd, _, data = closure.fine_run(case)
    params, cl_in, cl_out, rep = closure.train_closure(case, data)
    (m, c), _ = closure.calibrate(case, nd, data, (params, cl_in, cl_out))
    print("calibration", m, c)
    return case, nd, data, (params, cl_in, cl_out * m, c), rep


def test_network_beats_smagorinsky_a_priori(gpu_lib, pipeline):
    _, _, _, _, rep = pipeline
    ap = rep["apriori"]
    print(ap)
    assert ap["network_stress_err"] < 0.5 * ap["smagorinsky_stress_err"]
    assert ap["network_stress_err"] < ap["constant_best_stress_err"]


def test_network_beats_smagorinsky_short_horizons(gpu_lib, pipeline):
    case, nd, data, cl, _ = pipeline
    nd_c = closure.coarse_geometry(nd, case)
    first = case.snapshots - case.holdout
    starts = list(range(first, first + 4))
    horizons = (1, 4, 8)
    smag = closure.short_horizon(case, nd_c, data, None, starts, horizons=horizons)
    net = closure.short_horizon(case, nd_c, data, cl, starts, horizons=horizons)
    print("smagorinsky", smag, "network", net)
    for h in horizons:
        assert np.isfinite(net[str(h)])
        assert net[str(h)] < 0.97 * smag[str(h)], h


def test_network_beats_smagorinsky_held_out_mean_flow(gpu_lib, pipeline):
    case, nd, data, cl, _ = pipeline
    nd_c = closure.coarse_geometry(nd, case)
    smag = closure.statistics(case, nd_c, data, None)["heldout"]
    net = closure.statistics(case, nd_c, data, cl)["heldout"]
    print("smagorinsky", smag, "network", net)
    assert "diverged" not in net
    assert net["mean_l2"] < smag["mean_l2"]
