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


def _host_fp16_constants(w, wf, e, q):
    """dt3, dk0, dk1 as the host computes them (tlbm_gpu.cpp fill_weights):
    double arithmetic on the weights w and their float values wf, each
    constant then rounded to float, as the kernel receives it."""
    wd = [float(x) for x in w]
    wfd = [float(x) for x in wf]
    sum_w, sum_f = 0.0, 0.0
    w2, w2f = 0.0, 0.0
    for s in range(q):
        sum_w += wd[s]
        sum_f += wfd[s]
        w2 += wd[s] * float(int(e[s, 0]) ** 2)
        w2f += wfd[s] * float(int(e[s, 0]) ** 2)
    k1f = 1.0 + (sum_f - 1.0) - 3.0 * w2f
    dt3 = np.float32(3.0 * ((w2 - w2f) / w2f))
    dk0 = np.float32((sum_f - 1.0) - (sum_w - 1.0))
    dk1 = np.float32(k1f - (1.0 + (sum_w - 1.0) - 3.0 * w2))
    return Fr(float(dt3)), Fr(float(dk0)), Fr(float(dk1))


def _exact_weight_tables(q):
    e, _, _, w_exact = lattice_tables.tables(q)
    return e, [Fr(float(x)) for x in w_exact], [Fr(float(np.float32(x))) for x in w_exact]


@pytest.mark.parametrize("case", ["d3q27", "d3q15_exact", "d3q19_exact"])
def test_fp16_equilibrium_correction(case):
    """FP16 builds its equilibrium from the float weights w_f and corrects it
    to the balance of the weights w (D3Q27's double weights; the exact
    D3Q15/D3Q19 fractions): 3 e.u becomes (3 + dt3) e.u with
    dt3 = 3 (W2 / W2_f - 1), and the rest population gains
    rho (1.5 dk1 u^2 - dk0) with dk0 = k0_f - k0, dk1 = k1_f - k1.  The
    constants are the float values the host sends (3 + dt3 itself would
    round to 3.0f).  In exact arithmetic the momentum of f - feq is then
    k2 rho u up to the float rounding of dt3 (W2_f |dt3 - float(dt3)| rho u,
    below 1e-15 rho u), against 7.45e-9 rho u without dt3; its mass is
    rho (1.5 k1 u^2 - k0) up to the (1 + dt3/3)^2 change of the quadratic
    term (about 1e-11 here, against the 7.45e-9 that dk0 removes)."""
    if case == "d3q27":
        q = 27
        e, w = _weights(27, "double")
        _, wf = _weights(27, "float")
    else:
        q = 15 if case == "d3q15_exact" else 19
        e, w, wf = _exact_weight_tables(q)
    k0, k0f = sum(w) - 1, sum(wf) - 1
    W2 = sum(w[s] * int(e[s, 0]) ** 2 for s in range(q))
    W2f = sum(wf[s] * int(e[s, 0]) ** 2 for s in range(q))
    k1, k2, k2f = 1 + k0 - 3 * W2, 1 - 3 * W2, 1 - 3 * W2f
    dt3, dk0, dk1 = _host_fp16_constants(w, wf, e, q)
    assert dt3 != 0 and abs(float(dt3) + 2.2e-8) < 0.2e-8
    assert float(3 + dt3) != float(np.float32(3 + dt3))  # the scale alone rounds to 3.0f
    c = (3 + dt3) / 3
    rng = np.random.default_rng(q)
    for trial in range(3):
        f = [wf[s] * (1 + Fr(float(rng.uniform(-0.2, 0.2)))) for s in range(q)]
        rho = sum(f)
        u = [sum(f[s] * int(e[s, a]) for s in range(q)) / rho for a in range(3)]
        usq = sum(x * x for x in u)

        def nonequilibrium(scale, with_rest):
            feq = []
            for s in range(q):
                eu = scale * sum(int(e[s, a]) * u[a] for a in range(3))
                feq.append(wf[s] * rho * (1 + 3 * eu + Fr(9, 2) * eu * eu - Fr(3, 2) * usq))
            if with_rest:
                feq[0] += rho * (Fr(3, 2) * dk1 * usq - dk0)
            return [f[s] - feq[s] for s in range(q)]

        nq = nonequilibrium(c, True)
        for a in range(3):
            mom = sum(nq[s] * int(e[s, a]) for s in range(q))
            assert abs(mom - k2 * rho * u[a]) <= Fr(1, 10 ** 15) * abs(rho * u[a])
        target = rho * (Fr(3, 2) * k1 * usq - k0)
        bound = Fr(9, 2) * rho * W2f * usq * abs(c * c - 1) * Fr(101, 100) + \
            Fr(1, 10 ** 15) * rho
        assert abs(sum(nq) - target) <= bound
        # without the correction both balances miss by ~7.45e-9 rho (u)
        assert abs(float(dk0)) > 100 * float(bound)
        plain = nonequilibrium(Fr(1), False)
        for a in range(3):
            mom = sum(plain[s] * int(e[s, a]) for s in range(q))
            assert mom == k2f * rho * u[a]
            assert abs(mom - k2 * rho * u[a]) > Fr(5, 10 ** 9) * abs(rho * u[a])


@pytest.mark.parametrize("q", [15, 19, 27])
def test_reconstruction_partners_share_their_weight(q):
    """FP16 steps in the scaled deviation h = (f - w)/w.  The bounce-back
    (f_s <- f_bb(s)) and the inlet/outlet reconstruction
    f_unk = feq_unk + f_par - feq_par keep their form in that scale only
    because a direction and its partner have the same weight: checked for
    every pair of the tables (reference and exact weights)."""
    e, bb, w, w_exact = lattice_tables.tables(q)
    _, _, _, _, unk, _ = lattice_tables._parse_tables(lattice_tables._tables(q))
    for weights in (w, w_exact, [float(np.float32(x)) for x in w],
                    [float(np.float32(x)) for x in w_exact]):
        for s in range(q):
            assert weights[s] == weights[bb[s]], (q, s)
        for in_unk, in_par, out_unk, out_par, _, _ in unk:
            assert weights[in_unk] == weights[in_par], (q, in_unk, in_par)
            assert weights[out_unk] == weights[out_par], (q, out_unk, out_par)
