"""Pins the FP64 MLP oracle: backward against central finite differences,
GELU derivative against finite differences, bf16 rounding against known
values."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -3.0e-3], np.float32)
    r = mo.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                   # tie -> even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0 + 2 ** -6         # tie -> even (up)
    assert abs(r[4] + 3.0e-3) / 3.0e-3 < 2 ** -8


def test_gelu_grad_matches_finite_difference():
    x = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(x), fd, rtol=1e-7, atol=1e-9)
    # closed-form special values
    assert mo.gelu(0.0) == 0.0 and mo.gelu_grad(0.0) == 0.5


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    p = mo.glorot(3)
    feat = rng.normal(size=(40, 8))
    tgt = rng.normal(size=40)
    g, loss = mo.gradient(p, feat, tgt, bf=False)
    assert loss == mo.loss_only(p, feat, tgt)
    idx = rng.choice(mo.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [0, 511, 512, 575, 8895, 8959, 8960]])   # W1, b1, W4, b4 corners
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_only(pp, feat, tgt) - mo.loss_only(pm, feat, tgt)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_parameter_count():
    assert mo.N_PARAMS == 8961


def test_asymmetric_gradient_finite_difference():
    """Gradient of the weighted loss against central differences (FP64)."""
    rng = np.random.default_rng(9)
    feat = rng.normal(size=(64, 8))
    tgt = rng.normal(size=64)
    p = mo.glorot(3)
    g, _ = mo.gradient(p, feat, tgt, bf=False, over=3.0)

    def loss(q):
        y, _ = mo.forward(q, feat, False)
        e = y - tgt
        return float(np.mean(np.where(e > 0, 3.0, 1.0) * e * e))

    idx = rng.choice(mo.N_PARAMS, 40, replace=False)
    for k in idx:
        d = np.zeros_like(p)
        d[k] = 1e-6
        fd = (loss(p + d) - loss(p - d)) / 2e-6
        assert fd == pytest.approx(g[k], rel=1e-4, abs=1e-8)
