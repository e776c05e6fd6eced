"""Pin the learned closure's CPU oracle (oracle/mlp_oracle.py) -- CPU only.
The closure has no reference counterpart; its oracle is pinned by finite
differences of the unrounded network, the BF16 rounding rule and the
activation's derivative."""
import numpy as np
import pytest
from scipy.special import erf

from oracle import mlp_oracle as M


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM + 32)
    p[:M.NPARAM] = rng.standard_normal(M.NPARAM) * scale
    p[M.OFF["W1"]:M.OFF["b1"]].reshape(64, 16)[:, 12:] = 0
    w4 = p[M.OFF["W4"]:M.OFF["b4"]].reshape(16, 64)
    w4[9:] = 0
    p[M.OFF["b4"] + 9:M.NPARAM] = 0
    p[M.NPARAM:M.NPARAM + 12] = rng.uniform(0.5, 2.0, 12)
    p[M.NPARAM + 16:M.NPARAM + 25] = rng.uniform(0.5, 2.0, 9)
    return p


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.00390625, 1.0 + 2 ** -8 + 2 ** -9, -3.14159265, 65504.0], np.float32)
    b = M.bf16(x)
    assert b[0] == 1.0
    assert b[1] == 1.0            # exactly halfway: ties to even (1.0)
    assert b[2] == np.float32(1.0078125)
    assert abs(b[3] + 3.140625) < 1e-7
    assert b[4] == 65536.0


def test_activation_and_derivative():
    x = np.linspace(-4, 4, 100)          # no point at the kink
    g, dg = M.activation(x)
    h = 1e-6
    fd = (M.activation(x + h)[0] - M.activation(x - h)[0]) / (2 * h)
    assert np.abs(dg - fd).max() < 1e-8
    assert np.array_equal(g, np.where(x > 0, x, 0.0))
    assert M.activation(np.array([0.0]))[0][0] == 0.0 and M.activation(np.array([0.0]))[1][0] == 0.0


def test_gelu_activation_and_derivative():
    """tanh-form GELU: derivative by finite differences; within 1e-3 of the
    exact x Phi(x) (the tanh form's known accuracy); limits 0 and x; and the
    kernel's rewrite x sigmoid(2u) = 0.5 x (1 + tanh u)"""
    x = np.linspace(-6, 6, 241)
    g, dg = M.activation(x, M.GELU)
    h = 1e-6
    fd = (M.activation(x + h, M.GELU)[0] - M.activation(x - h, M.GELU)[0]) / (2 * h)
    assert np.abs(dg - fd).max() < 1e-8
    exact = 0.5 * x * (1 + erf(x / np.sqrt(2)))
    assert np.abs(g - exact).max() < 1e-3
    assert M.activation(np.array([0.0]), M.GELU)[0][0] == 0.0
    assert abs(M.activation(np.array([0.0]), M.GELU)[1][0] - 0.5) < 1e-15
    big = np.array([-30.0, 30.0])
    assert np.allclose(M.activation(big, M.GELU)[0], [0.0, 30.0], atol=1e-12)
    u = 0.7978845608028654 * (x + 0.044715 * x ** 3)
    assert np.abs(x / (1 + np.exp(-2 * u)) - g).max() < 1e-12


@pytest.mark.parametrize("act", [M.RELU, M.GELU])
def test_gradient_finite_differences(act):
    p = random_params(1)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((64, 12))
    y = rng.standard_normal((64, 9))
    net = M.Net(p, bf=False, act=act)
    loss, g = net.loss_grad(f, y)
    idx = rng.choice(np.flatnonzero(g), 40, replace=False)
    for j in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (M.Net(pp, bf=False, act=act).loss_grad(f, y)[0] - M.Net(pm, bf=False, act=act).loss_grad(f, y)[0]) / (2 * h)
        assert abs(fd - g[j]) < 1e-6 * max(1.0, abs(g[j])) + 1e-9, (j, fd, g[j])


def test_zero_output_scale_gives_zero_target():
    """an output column with scale 0 (all zero in the pool) is trained towards
    0: the loss equals the one with that column's label set to anything"""
    p = random_params(3)
    p[M.NPARAM + 16 + 7] = 0.0
    rng = np.random.default_rng(3)
    f = rng.standard_normal((64, 12))
    y = rng.standard_normal((64, 9))
    net = M.Net(p, bf=False)
    l0, g0 = net.loss_grad(f, y)
    y[:, 7] = 0.0
    l1, g1 = net.loss_grad(f, y)
    assert np.isfinite(l0) and l0 > 0
    assert l0 == l1
    assert np.array_equal(g0, g1)
