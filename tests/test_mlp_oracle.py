"""The closure network's CPU oracle, pinned by what the mathematics fixes:
central finite differences of the loss, the first Adam step in closed form,
bf16 rounding against torch's float -> bfloat16 conversion, and a constant
network reducing the closure to BGK with a fixed relaxation time."""
import numpy as np
import pytest

import pyoracle
from pyoracle import MLP_NPARAM, Oracle
from tlbm_amd import geometry


def _random_problem(seed, n=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, MLP_NPARAM)
    x = rng.normal(0, 1, (n, 7))
    t = rng.uniform(0.01, 2.0, n)
    return p, x, t


def test_loss_gradient_matches_finite_differences():
    p, x, t = _random_problem(3)
    loss, g = pyoracle.mlp_loss_grad(p, x, t)
    assert loss > 0
    h = 1e-6
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(224, 20, replace=False),            # W1
                          224 + np.arange(32),                             # b1
                          256 + rng.choice(1024, 40, replace=False),       # W2
                          1280 + np.arange(32), 1312 + np.arange(32), [1344]])
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (pyoracle.mlp_loss_grad(pp, x, t)[0] - pyoracle.mlp_loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd), k


def test_forward_matches_the_loss_definition():
    p, x, t = _random_problem(5, n=10)
    y = pyoracle.mlp_forward(p, x)
    loss, _ = pyoracle.mlp_loss_grad(p, x, t)
    assert loss == pytest.approx(np.mean(0.5 * (y - t) ** 2), rel=1e-14)
    assert np.all(y > 0)


def test_first_adam_step_is_lr_times_sign():
    p = np.linspace(-1, 1, MLP_NPARAM)
    g = np.where(np.arange(MLP_NPARAM) % 3 == 0, 0.5, -2e-3)
    m, v = np.zeros(MLP_NPARAM), np.zeros(MLP_NPARAM)
    q = p.copy()
    pyoracle.mlp_adam(q, g, m, v, 1, lr=1e-3)
    # m/(1-b1) = g, v/(1-b2) = g^2: the step is lr * g / (|g| + eps)
    assert np.allclose(q, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-30, 50), [0.0, -0.0, 1.0, 3.0],
                        # exact ties between two bf16 values
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8)]),
                        np.array([np.inf, -np.inf])]).astype(np.float32)
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(pyoracle.round_bf16(v), want)


def test_constant_network_closure_is_bgk_with_fixed_tau():
    nx, ny, nz = 8, 7, 12
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 3, 4, 2)])
    omega, c, cl_out = 1.6, 0.3, 0.05
    params = np.zeros(MLP_NPARAM, np.float32)
    params[-1] = c
    a = Oracle(19, nx, ny, nz, nd, omega=omega, u_lbm=0.05)
    a.set_closure(params, 1.0, cl_out)
    sp = np.log1p(np.exp(np.float64(np.float32(c))))
    b = Oracle(19, nx, ny, nz, nd, omega=1.0 / (1.0 / omega + cl_out * sp), u_lbm=0.05)
    a.step(20)
    b.step(20)
    assert np.array_equal(a.populations(), b.populations())


def test_bf16_inference_matches_torch_emulation():
    """The oracle's inference as the collide kernel runs it, against an
    emulation built on torch's bfloat16 conversion: x, W1, W2, w3 and h1
    rounded to bf16 (h1 from its float value), biases float, h2 and the
    products in float64.  Dropping a rounding moves some outputs by far
    more than the 1e-12 allowed here."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.4, MLP_NPARAM).astype(np.float32)
    x = rng.normal(0, 1, (300, 7)).astype(np.float32)

    def bf(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16) \
            .to(torch.float64).numpy()

    W1 = p[:224].reshape(32, 7)
    b1 = p[224:256].astype(np.float64)
    W2 = p[256:1280].reshape(32, 32)
    b2 = p[1280:1312].astype(np.float64)
    w3 = p[1312:1344]
    b3 = np.float64(p[1344])
    pre1 = bf(x) @ bf(W1).T + b1
    h1 = bf(np.maximum(pre1, 0.0).astype(np.float32))
    want = np.maximum(h1 @ bf(W2).T + b2, 0.0) @ bf(w3) + b3
    got = pyoracle.mlp_infer_bf16(p, x)
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    # the h1 rounding matters at this scale
    h1_exact = np.maximum(pre1, 0.0)
    assert np.abs(np.maximum(h1_exact @ bf(W2).T + b2, 0.0) @ bf(w3) + b3 - want).max() > 1e-6
