"""HIP MLP kernels (BF16 matrix cores, FP32 accumulation) against the FP64
oracle: forward, parameter gradient, and training."""
import numpy as np
import pytest

from graphics_amd import sphgpu as sg
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    feat = rng.normal(size=(n, 8)).astype(np.float32)
    tgt = (np.sin(feat[:, 0]) + 0.5 * feat[:, 1] * feat[:, 2] - 0.3 * np.abs(feat[:, 3])).astype(np.float32)
    return feat, tgt


def test_forward_matches_oracle(gpu):
    feat, _ = _data(1000)
    m = sg.Mlp(4096, seed=5)
    p = m.get_params().astype(np.float64)
    y = m.forward(feat)
    ref_bf, _ = mo.forward(p, feat, bf=True)
    ref, _ = mo.forward(p, feat, bf=False)
    scale = np.abs(ref).max()
    # same rounding points: FP32 accumulation order only
    assert np.abs(y - ref_bf).max() / scale < 2e-3
    # against exact FP64: bounded by BF16 operand rounding
    assert np.abs(y - ref).max() / scale < 3e-2
    m.close()


def test_gradient_matches_oracle(gpu):
    feat, tgt = _data(2048, 1)
    m = sg.Mlp(4096, seed=7)
    p = m.get_params().astype(np.float64)
    g, loss = m.gradient(feat, tgt)
    gref_bf, lref = mo.gradient(p, feat, tgt, bf=True)
    gref, _ = mo.gradient(p, feat, tgt, bf=False)
    assert loss == pytest.approx(lref, rel=2e-3)
    for (a, b, tol) in ((g, gref_bf, 2e-2), (g, gref, 6e-2)):
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < tol, err
    # each parameter block is non-trivially present (no dropped term)
    k = 0
    for o, i in mo.SIZES:
        for blk in (o * i, o):
            a, b = g[k:k + blk], gref[k:k + blk]
            assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b) + 1e-7
            k += blk
    m.close()


def test_training_reduces_loss(gpu):
    feat, tgt = _data(4096, 2)
    m = sg.Mlp(4096, seed=3)
    _, l0 = m.gradient(feat, tgt)
    l1 = m.train(feat, tgt, 300, 3e-3)
    assert l1 < 0.1 * l0, (l0, l1)
    # the trained network still matches the oracle forward
    p = m.get_params().astype(np.float64)
    ref, _ = mo.forward(p, feat[:500], bf=True)
    assert np.abs(m.forward(feat[:500]) - ref).max() / np.abs(ref).max() < 5e-3
    m.close()


def test_asymmetric_loss_gradient(gpu):
    """Over-predictions weighted 4x: gradient and loss against the oracle,
    and the weight shows (it differs from the symmetric gradient)."""
    feat, tgt = _data(2048, 4)
    m = sg.Mlp(4096, seed=11)
    p = m.get_params().astype(np.float64)
    g1, _ = m.gradient(feat, tgt)
    m.set_loss(4.0)
    g4, loss = m.gradient(feat, tgt)
    gref, lref = mo.gradient(p, feat, tgt, bf=True, over=4.0)
    assert loss == pytest.approx(lref, rel=2e-3)
    assert np.linalg.norm(g4 - gref) / np.linalg.norm(gref) < 2e-2
    assert np.linalg.norm(g4 - g1) > 0.1 * np.linalg.norm(g1)
    with pytest.raises(sg.SphError):
        m.set_loss(0.0)
    m.close()
