"""The oracle is pinned before it is trusted.

* Every golden fixture (outputs of the compiled reference, tests/golden) is
  reproduced bitwise, FP64 and FP32, moments and forces.
* Closed forms the mathematics fixes: lattice-table invariants, the
  rest-state density growth caused by the float-literal weights, exact mass
  conservation with exact weights in a periodic box.
* Planted mistakes: oracle variants with one deliberate error each are
  rebuilt and must fail at least one fixture.
"""
import os
import subprocess

import numpy as np
import pytest

import goldens
from lattice_tables import _parse_tables, _run_c_probe, _tables
from pyoracle import Oracle
from tlbm_amd import geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = goldens.load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_oracle_reproduces_reference_bitwise(case):
    o = Oracle(case.q, *case.dims, case.nd, precision=case.precision, **case.kwargs())
    o.step(case.params["numTs"])
    m = o.macros()
    for field, key in (("ux", "ux"), ("uy", "uy"), ("uz", "uz"), ("density", "rho")):
        assert np.array_equal(m[key], case.out[field]), field
    if "Fx" in case.out:
        fo = o.forces()
        for field in ("Fx", "Fy", "Fz"):
            assert np.array_equal(fo[field], case.out[field]), field


STANDARD = {
    15: lambda e: (np.abs(e).sum(axis=1) != 2),
    19: lambda e: (np.abs(e).sum(axis=1) <= 2),
    27: lambda e: np.ones(len(e), bool),
}


@pytest.mark.parametrize("q", [15, 19, 27])
def test_lattice_tables_closed_forms(q):
    e, bb, w, w_exact, unk, zero = _parse_tables(_tables(q))
    # the standard velocity set, each vector once
    allv = np.array([[x, y, z] for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)])
    want = {tuple(v) for v in allv[STANDARD[q](allv)]}
    assert {tuple(v) for v in e} == want and len(e) == q
    assert tuple(e[0]) == (0, 0, 0)
    # bounce-back is the opposite direction
    assert np.array_equal(e[bb], -e)
    # exact weights: isotropy up to second order, sum 1, c_s^2 = 1/3
    assert abs(w_exact.sum() - 1.0) < 1e-15
    assert np.allclose(w_exact @ e, 0.0, atol=1e-16)
    assert np.allclose(np.einsum("s,si,sj->ij", w_exact, e, e), np.eye(3) / 3.0, atol=1e-15)
    # reference weights: the float-rounded fractions for D3Q15/D3Q19
    if q in (15, 19):
        assert np.array_equal(w, w_exact.astype(np.float32).astype(np.float64))
        excess = {15: 7.450580596923828e-09, 19: 1.4901161193847656e-08}[q]
        assert w.sum() - 1.0 == pytest.approx(excess, rel=1e-6)
    else:
        assert np.array_equal(w, w_exact)
    # boundary lists: inlet unknowns are e_z=+1, outlet unknowns e_z=-1,
    # partners are opposites (D3Q19/27) or z-mirror images (D3Q15)
    in_unk, in_par, out_unk, out_par, in_list, out_list = unk.T
    assert set(in_unk) == set(np.where(e[:, 2] == 1)[0])
    assert set(out_unk) == set(np.where(e[:, 2] == -1)[0])
    assert set(in_list) == set(out_unk) and set(out_list) == set(in_unk)
    assert set(zero) == set(np.where(e[:, 2] == 0)[0]) and zero == sorted(zero)
    mirror = e * np.array([1, 1, -1])
    for u, p in zip(in_unk, in_par):
        expect = e[u] * -1 if q != 15 else mirror[u]
        assert np.array_equal(e[p], expect)
    assert np.array_equal(out_par, in_unk) and np.array_equal(out_unk, in_par)


@pytest.mark.parametrize("q", [15, 19, 27])
def test_product_tables_equal_oracle_tables(q):
    """The HIP library's lattice.h carries the same tables as the oracle."""
    src = r"""
#include <cstdio>
#include "lattice.h"
template <class L> void dump() {
    for (int s = 0; s < L::Q; ++s)
        std::printf("%d %d %d %d %.17g %.17g\n", L::ex[s], L::ey[s], L::ez[s], L::bb[s],
                    L::w_ref[s], L::w_exact[s]);
    std::printf("UNK");
    for (int i = 0; i < L::n_unk; ++i) std::printf(" %d %d %d %d %d %d", L::in_unk[i], L::in_par[i],
        L::out_unk[i], L::out_par[i], L::in_list[i], L::out_list[i]);
    std::printf("\nZERO");
    for (int i = 0; i < L::n_zero; ++i) std::printf(" %d", L::zero_list[i]);
    std::printf("\n");
}
int main(int, char **argv) {
    int q = std::atoi(argv[1]);
    if (q == 15) dump<tlbm::D3Q15>(); else if (q == 19) dump<tlbm::D3Q19>(); else dump<tlbm::D3Q27>();
}
"""
    out = _run_c_probe(src, ["-std=c++17", "-I", os.path.join(ROOT, "tlbm_amd", "csrc")], [str(q)],
                       "#include <cstdlib>\n", compiler="g++")
    prod = _parse_tables(out)
    orc = _parse_tables(_tables(q))
    for a, b in zip(prod, orc):
        assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("exact", [False, True])
def test_rest_state_density_closed_form(q, exact):
    """Uniform rest state in a periodic box: u stays 0 and
    rho_n = S * (1 + omega*(S - 1))**(n-1), S = sum of the weights used."""
    nx, ny, nz = 3, 4, 5
    omega, n = 1.3, 40
    o = Oracle(q, nx, ny, nz, np.zeros(nx * ny * nz, int), omega=omega, exact_weights=exact)
    S = o.weights().sum()
    o.step(n)
    m = o.macros()
    expect = S * (1.0 + omega * (S - 1.0)) ** (n - 1)
    # rounding of 40 steps of sums stays below 5e-14; the weight excess
    # moves rho by 3.9e-7 (D3Q15) and 7.7e-7 (D3Q19) over the same steps
    assert np.allclose(m["rho"], expect, rtol=5e-14, atol=0)
    # u = 0 up to the rounding of the ascending momentum sums
    for k in ("ux", "uy", "uz"):
        assert np.max(np.abs(m[k])) < 1e-16
    if exact or q == 27:
        assert abs(m["rho"][0] - 1.0) < 1e-14


def test_exact_weights_conserve_mass_in_periodic_box_with_solids():
    """Bounce-back and BGK with exact weights conserve total mass."""
    nx, ny, nz, q = 10, 9, 8, 19
    nd = geometry.periodic_box(nx, ny, nz, obstacles=[geometry.sphere(4, 4, 3, 2.5)])
    rng = np.random.default_rng(5)
    o = Oracle(q, nx, ny, nz, nd, omega=1.7, exact_weights=True)
    w = o.weights()
    f0 = w[None, :] * (1.0 + 0.05 * rng.standard_normal((nx * ny * nz, q)))
    o.set_populations(f0)
    m0 = f0.sum()
    o.step(30)
    f = o.populations()
    assert abs(f.sum() - m0) / m0 < 1e-13


def test_reference_weights_grow_mass_per_closed_form():
    """With the float weights, a rest box gains mass by (1+omega(S-1)) per step."""
    o = Oracle(19, 4, 4, 4, np.zeros(64, int), omega=1.5)
    S = o.weights().sum()
    f = o.populations()
    m = [f.sum()]
    for _ in range(5):
        o.step(1)
        m.append(o.populations().sum())
    ratios = np.array(m[1:]) / np.array(m[:-1])
    assert np.allclose(ratios[1:], 1.0 + 1.5 * (S - 1.0), rtol=0, atol=1e-15)


PLANTS = {
    1: "equilibrium without the quadratic term",
    2: "y neighbour with the wrong sign",
    3: "opposite partners at the inlet (D3Q15 uses z-mirror partners)",
    4: "outlet velocity closure with the wrong sign",
    5: "Smagorinsky constant squared",
    6: "type-5 rule reading uy of the node itself",
}


@pytest.mark.parametrize("plant", sorted(PLANTS), ids=[PLANTS[k] for k in sorted(PLANTS)])
def test_fixtures_catch_planted_oracle_mistake(plant, tmp_path):
    """Each plausible mistake, planted in the oracle, fails some fixture."""
    so = tmp_path / f"liboracle_plant{plant}.so"
    subprocess.run(["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-ffp-contract=off",
                    f"-DORACLE_PLANT={plant}", "-o", str(so),
                    os.path.join(ROOT, "oracle", "lbm_oracle.c"),
                    os.path.join(ROOT, "oracle", "mlp_oracle.c"), "-lm"], check=True)
    import pyoracle
    saved_lib, saved_path = pyoracle._lib, pyoracle.LIB_PATH
    pyoracle._lib, pyoracle.LIB_PATH = None, str(so)
    try:
        caught = []
        for case in CASES:
            if case.precision != "f64" or case.params["numTs"] > 100:
                continue
            o = Oracle(case.q, *case.dims, case.nd, precision="f64", **case.kwargs())
            o.step(case.params["numTs"])
            m = o.macros()
            if not all(np.array_equal(m[k], case.out[f]) for f, k in
                       (("ux", "ux"), ("uy", "uy"), ("uz", "uz"), ("density", "rho"))):
                caught.append(case.name)
            del o
        assert caught, f"no fixture detects: {PLANTS[plant]}"
    finally:
        pyoracle._lib, pyoracle.LIB_PATH = saved_lib, saved_path
