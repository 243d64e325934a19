"""The surrogate MLP oracle (oracle/mlp.py) pinned by the mathematics:
with bf16 rounding switched off, its analytic gradients match central
finite differences of its own loss; its bf16 rounding matches torch's."""
import numpy as np
import pytest

from oracle import mlp


@pytest.fixture
def exact(monkeypatch):
    monkeypatch.setattr(mlp, "bf16", lambda x: np.asarray(x, dtype=np.float64))
    monkeypatch.setattr(mlp, "FT", np.float64)
    return mlp


def _small_net(rng):
    net = mlp.init_net(rng, F=8, width=6, layers=3, sigma=0.7)
    net["W"] = [w.astype(np.float64) for w in net["W"]]
    net["b"] = [rng.standard_normal(b.shape) * 0.1 for b in net["b"]]
    return net


def test_gradients_match_finite_differences(exact):
    rng = np.random.default_rng(0)
    net = _small_net(rng)
    inp = rng.uniform(-1, 1, (7, 4))
    tgt = rng.standard_normal((7, 2))

    def loss_fn():
        y, _ = mlp.forward(inp, net)
        return 0.5 * np.mean(np.sum((y.astype(np.float64) - tgt) ** 2, axis=1))

    loss, gW, gb = mlp.loss_and_grads(inp, tgt, net)
    assert np.isclose(loss, loss_fn(), rtol=1e-5)
    eps = 1e-5
    for l in range(len(net["W"])):
        r, c = net["W"][l].shape
        for (i, j) in [(0, 0), (min(1, r - 1), min(2, c - 1)), (r - 1, c - 1)]:
            w0 = net["W"][l][i, j]
            net["W"][l][i, j] = w0 + eps
            lp = loss_fn()
            net["W"][l][i, j] = w0 - eps
            lm = loss_fn()
            net["W"][l][i, j] = w0
            assert np.isclose(gW[l][i, j], (lp - lm) / (2 * eps), rtol=1e-5, atol=1e-9), (l, i, j)
        b0 = net["b"][l][0]
        net["b"][l][0] = b0 + eps
        lp = loss_fn()
        net["b"][l][0] = b0 - eps
        lm = loss_fn()
        net["b"][l][0] = b0
        assert np.isclose(gb[l][0], (lp - lm) / (2 * eps), rtol=1e-5, atol=1e-9)


def test_bf16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(1)
    x = (rng.standard_normal(10000) * 10.0 ** rng.uniform(-8, 8, 10000)).astype(np.float32)
    x[:4] = [0.0, -0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8]  # ties round to even
    ours = mlp.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours, ref)


def test_gelu_derivative_matches_difference(exact):
    a = np.linspace(-6, 6, 1001).astype(np.float64)
    num = (mlp.gelu(a + 1e-3).astype(np.float64) - mlp.gelu(a - 1e-3).astype(np.float64)) / 2e-3
    assert np.allclose(mlp.gelu_grad(a), num, atol=2e-6)
