"""The identity behind the FP32 collide kernels' conserving relax.

FP32 fluid nodes relax with the rounding residue of the mass and momentum of
f - feq removed, measured against the values those moments have in exact
arithmetic with the weights w (the reference's: float values for D3Q15 and
D3Q19, double values for D3Q27):

    sum_s (f_s - feq_s)      = rho (1.5 k1 |u|^2 - k0)
    sum_s e_s (f_s - feq_s)  = k2 rho u
    k0 = sum w - 1,  k1 = 1 + k0 - 3 W2,  k2 = 1 - 3 W2,  W2 = sum_s w e_x^2

where rho = sum f, u = sum e f / rho and feq_s = w_s rho (1 + 3 e.u +
4.5 (e.u)^2 - 1.5 |u|^2) (include/LatticeStructure.hpp:92-103).  Here both
sides are evaluated in exact rational arithmetic on random populations, for
the weights and for their float roundings (the FP32 kernels' equilibrium
weights), so a dropped term, a wrong sign or a wrong constant fails.  The
weights are the oracle's tables (equal to the product's, test_oracle.py).
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

import lattice_tables


def _weights(q, rounding):
    e, _, w, _ = lattice_tables.tables(q)
    if rounding == "float":
        w = [float(np.float32(x)) for x in w]
    return e, [Fr(float(x)) for x in w]


def test_d3q27_float_weights_do_not_sum_to_one():
    """The reason the targets come from the weights, not their float
    roundings: the D3Q27 weights sum to 1 to double precision, their float
    values to 1 + 7.45e-9, a mass source in exact arithmetic."""
    _, w = _weights(27, "double")
    _, wf = _weights(27, "float")
    assert abs(float(sum(w) - 1)) < 1e-15
    assert abs(float(sum(wf) - 1) - 7.450580596923828e-09) < 1e-15


@pytest.mark.parametrize("rounding", ["double", "float"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_moments_of_nonequilibrium_match_closed_form(q, rounding):
    e, wf = _weights(q, rounding)
    dw = sum(wf) - 1
    w2 = sum(wf[s] * int(e[s, 0]) ** 2 for s in range(q))
    # the weights are symmetric: every axis has the same second moment, no cross terms
    assert all(sum(wf[s] * int(e[s, a]) ** 2 for s in range(q)) == w2 for a in range(3))
    assert all(sum(wf[s] * int(e[s, a]) * int(e[s, b]) for s in range(q)) == 0
               for a in range(3) for b in range(3) if a != b)
    k1, k2 = 1 + dw - 3 * w2, 1 - 3 * w2
    rng = np.random.default_rng(q)
    for trial in range(5):
        f = [wf[s] * (1 + Fr(float(rng.uniform(-0.2, 0.2)))) for s in range(q)]
        rho = sum(f)
        u = [sum(f[s] * int(e[s, a]) for s in range(q)) / rho for a in range(3)]
        usq = sum(c * c for c in u)
        feq = []
        for s in range(q):
            eu = sum(int(e[s, a]) * u[a] for a in range(3))
            feq.append(wf[s] * rho * (1 + 3 * eu + Fr(9, 2) * eu * eu - Fr(3, 2) * usq))
        nq = [f[s] - feq[s] for s in range(q)]
        assert sum(nq) == rho * (Fr(3, 2) * k1 * usq - dw)
        for a in range(3):
            assert sum(nq[s] * int(e[s, a]) for s in range(q)) == k2 * rho * u[a]


@pytest.mark.parametrize("q", [15, 19, 27])
def test_correction_removes_a_residue(q):
    """w_s (m0 + 3 e_s.m) carries mass m0 and momentum m up to the weights'
    normalisation (sum w_f = 1 + dw, 3 W2 = 1 - k2, both within 1e-7 of exact),
    so subtracting it from a residue (m0, m) leaves a residue smaller by 1e-7."""
    e, wf = _weights(q, "float")
    rng = np.random.default_rng(100 + q)
    m0 = Fr(float(rng.standard_normal()))
    m = [Fr(float(rng.standard_normal())) for _ in range(3)]
    corr = [wf[s] * (m0 + 3 * sum(int(e[s, a]) * m[a] for a in range(3))) for s in range(q)]
    left0 = m0 - sum(corr)
    assert abs(left0) <= Fr(1, 10 ** 7) * abs(m0)
    for a in range(3):
        left = m[a] - sum(corr[s] * int(e[s, a]) for s in range(q))
        assert abs(left) <= Fr(1, 10 ** 7) * max(abs(x) for x in m)


def test_fp16_d3q27_equilibrium_correction():
    """FP16 D3Q27 builds its equilibrium from the float weights w_f and
    corrects it to the balance of the double weights w: 3 e.u becomes
    t3 e.u with t3 = 3 W2 / W2_f, and the rest population gains
    rho (1.5 dk1 u^2 - dk0) with dk0 = k0_f - k0, dk1 = k1_f - k1.  In exact
    arithmetic the momentum of f - feq is then k2 rho u exactly, and its
    mass rho (1.5 k1 u^2 - k0) up to the t3^2 change of the quadratic term
    (4.5 rho W2_f u^2 (c^2 - 1), c = W2 / W2_f: about 1e-11 here, against
    the 7.45e-9 that the correction removes)."""
    e, w = _weights(27, "double")
    _, wf = _weights(27, "float")
    q = 27
    k0, k0f = sum(w) - 1, sum(wf) - 1
    W2 = sum(w[s] * int(e[s, 0]) ** 2 for s in range(q))
    W2f = sum(wf[s] * int(e[s, 0]) ** 2 for s in range(q))
    k1, k1f, k2 = 1 + k0 - 3 * W2, 1 + k0f - 3 * W2f, 1 - 3 * W2
    c = W2 / W2f
    dk0, dk1 = k0f - k0, k1f - k1
    assert abs(float(dk0) - 7.450580596923828e-09) < 1e-15
    rng = np.random.default_rng(27)
    for trial in range(3):
        f = [wf[s] * (1 + Fr(float(rng.uniform(-0.2, 0.2)))) for s in range(q)]
        rho = sum(f)
        u = [sum(f[s] * int(e[s, a]) for s in range(q)) / rho for a in range(3)]
        usq = sum(x * x for x in u)
        feq = []
        for s in range(q):
            eu = c * sum(int(e[s, a]) * u[a] for a in range(3))
            feq.append(wf[s] * rho * (1 + 3 * eu + Fr(9, 2) * eu * eu - Fr(3, 2) * usq))
        feq[0] += rho * (Fr(3, 2) * dk1 * usq - dk0)
        nq = [f[s] - feq[s] for s in range(q)]
        for a in range(3):
            assert sum(nq[s] * int(e[s, a]) for s in range(q)) == k2 * rho * u[a]
        target = rho * (Fr(3, 2) * k1 * usq - k0)
        bound = Fr(9, 2) * rho * W2f * usq * abs(c * c - 1) * Fr(101, 100)
        assert abs(sum(nq) - target) <= bound
        # without the correction the mass balance misses by ~rho dk0
        assert abs(float(dk0)) > 100 * float(bound)
