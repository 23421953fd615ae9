"""Pins the surrogate network's CPU oracle (oracle/mlp_oracle.py) before it checks the GPU path:
the analytic gradient against central finite differences, the piecewise-quadratic activation against
the exact (erf) GELU it follows, hand-computed features, and the bf16 rounding rule."""
import numpy as np
import pytest
from scipy.special import erf

import mlp_oracle as mo


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -7 + 2 ** -9, 3.0e38])
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0 + 2 ** -7, 1.765625 * 2.0 ** 127])  # 3e38: 0x7f61b1e6 -> 0x7f62
    got = mo.bf16(x)
    np.testing.assert_array_equal(got, want)


def test_gelu_tracks_exact_gelu_and_its_derivative():
    z = np.linspace(-8, 8, 16001)
    exact = 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    assert np.max(np.abs(mo.gelu(z) - exact)) < 0.0199
    assert mo.gelu(0.0) == 0.0
    big = np.array([2.5, 3.5, 5.0, 20.0])  # beyond z* = 2.463
    np.testing.assert_array_equal(mo.gelu(big), big)  # saturated: exactly x and 0
    np.testing.assert_array_equal(mo.gelu(-big), -0.0 * big)
    # on (-z*, 0) act = z/2 + a z^2 + b z^3: its minimum is the root of 1/2 + 2 a z + 3 b z^2
    a, b = mo.A, mo.BX * mo.IXS
    zmin = (-2 * a + np.sqrt(4 * a * a - 6 * b)) / (6 * b)
    assert abs(z[np.argmin(mo.gelu(z))] - zmin) < 1e-3 and abs(zmin + 0.821) < 1e-3
    assert abs(mo.gelu(zmin) - (zmin / 2 + a * zmin ** 2 + b * zmin ** 3)) < 1e-12
    assert abs(mo.gelu(zmin) + 0.17) < 0.013  # the exact GELU's minimum is -0.170 at -0.752
    h = 1e-6
    zz = z[np.abs(np.abs(z) - mo.XS) > 1e-3]  # away from the two clamp points
    fd = (mo.gelu(zz + h) - mo.gelu(zz - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(zz), fd, rtol=0, atol=1e-7)
    # C1: the slope is continuous through the clamp points (1 at z*, 0 at -z*)
    for zc, want in ((mo.XS, 1.0), (-mo.XS, 0.0)):
        np.testing.assert_allclose(mo.gelu_grad(np.array([zc - 1e-5, zc + 1e-5])), [want, want], atol=1e-4)


def test_features_hand_case():
    pos = np.array([[0.0, 0.0, 0.0], [0.06, 0.0, 0.0], [0.0, 0.0, 0.04]])
    vel = np.array([[0.0, 0.0, 0.0], [0.0, 0.3, 0.0], [0.0, 0.0, 0.0]])
    nbr = -np.ones((3, 12), np.int64)
    rest = np.zeros((3, 12))
    nbr[0, 0], rest[0, 0] = 1, 0.05  # stretched by 20 %
    nbr[0, 5], rest[0, 5] = 2, 0.05  # compressed by 20 %
    H = 1.0 / 75.0
    f = mo.features(pos, vel, nbr, rest, H, 10.0).reshape(3, 12, 6)
    np.testing.assert_allclose(f[0, 0], [2.0, 0.0, 0.0, 0.0, 0.3 * H / 0.05, 0.0], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(f[0, 5], [0.0, 0.0, -2.0, 0.0, 0.0, 0.0], rtol=1e-12, atol=1e-15)
    assert not f[0, [1, 2, 3, 4, 6, 7, 8, 9, 10, 11]].any() and not f[1:].any()


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mo.init_params(1)
    X = rng.normal(scale=0.7, size=(6, 72))
    T = rng.normal(scale=0.5, size=(6, 3))
    loss, g, _ = mo.loss_and_grad(p, X, T)
    # every bias, and a spread of weights of every layer
    idx = np.concatenate([np.arange(mo.OFF_B1, mo.OFF_W2), np.arange(mo.OFF_B4, mo.N_PARAMS),
                          rng.choice(np.arange(mo.OFF_W1, mo.OFF_B1), 40, replace=False),
                          rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 40, replace=False),
                          rng.choice(np.arange(mo.OFF_W3, mo.OFF_B3), 40, replace=False),
                          rng.choice(np.arange(mo.OFF_W4, mo.OFF_B4), 40, replace=False)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, X, T)[0] - mo.loss_and_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


@pytest.mark.parametrize("emulate", [False, True])
def test_loss_is_mean_squared_error(emulate):
    rng = np.random.default_rng(2)
    p = mo.init_params(3)
    X = rng.normal(size=(5, 72))
    T = rng.normal(size=(5, 3))
    loss, _, Y = mo.loss_and_grad(p, X, T, emulate=emulate)
    assert loss == pytest.approx(float(np.mean((Y - T) ** 2)), rel=1e-12)
