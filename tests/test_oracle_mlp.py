"""Pins of the network oracle: its BF16 rounding against torch's float32 -> bfloat16
conversion, and its unrounded forward pass and gradient against torch.nn and autograd
in float64 (library routines, independent of the restatement)."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

torch = pytest.importorskip("torch")


def _params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(mo.NPARAM)
    for k, o, s in mo.LAYOUT:
        p[o:o + int(np.prod(s))] = rng.normal(0, 0.4, int(np.prod(s)))
    norm = (rng.normal(0, 0.3, 9), rng.uniform(0.5, 2.0, 9), rng.normal(0, 0.1, 6), rng.uniform(0.5, 1.5, 6))
    return p, norm, rng


def _torch_net(P):
    layers = []
    for i, (w, b) in enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3"), ("W4", "b4"))):
        lin = torch.nn.Linear(P[w].shape[1], P[w].shape[0]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(P[w]))
            lin.bias.copy_(torch.from_numpy(P[b]))
        layers.append(lin)
        if i < 3:
            layers.append(torch.nn.Tanh())
    return torch.nn.Sequential(*layers)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), [0.0, -0.0, 1.0, 3.0e38]])
    x = x.astype(np.float32)
    # ties: values exactly halfway between two bfloat16 numbers
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), ref)


def test_forward_unrounded_matches_torch_nn():
    p, norm, rng = _params()
    P = mo.unpack(p)
    x = rng.normal(0, 1, (300, 9))
    y = mo.forward(P, x, norm, rounding=False)[-1]
    xn = (x - norm[0]) / norm[1]
    ref = _torch_net(P)(torch.from_numpy(xn)).detach().numpy() * norm[3] + norm[2]
    assert np.abs(y - ref).max() < 1e-12


def test_gradient_unrounded_matches_autograd():
    p, norm, rng = _params(2)
    P = mo.unpack(p)
    n = 64
    x = rng.normal(0, 1, (n, 9))
    t = rng.normal(0, 1, (n, 6))
    loss, g = mo.grad(P, x, t, norm, rounding=False)
    net = _torch_net(P)
    xn = torch.from_numpy((x - norm[0]) / norm[1])
    tn = torch.from_numpy((t - norm[2]) / norm[3])
    L = ((net(xn) - tn) ** 2).sum() / (6 * n)
    L.backward()
    G = {}
    for k, mod in zip(("1", "2", "3", "4"), [m for m in net if isinstance(m, torch.nn.Linear)]):
        G["W" + k] = mod.weight.grad.numpy()
        G["b" + k] = mod.bias.grad.numpy()
    assert abs(loss - L.item()) < 1e-13
    assert np.abs(g - mo.pack(G)).max() < 1e-12


def test_smag_pointwise_matches_the_rhs_oracle_form():
    rng = np.random.default_rng(4)
    g = rng.normal(0, 1, (50, 9))
    c2 = 0.3
    tau = mo.smag_pointwise(g, c2)
    A = g.reshape(50, 3, 3)
    S = 0.5 * (A + A.transpose(0, 2, 1))
    Smag = np.sqrt(2 * (S * S).sum(axis=(1, 2)))
    full = -2 * c2 * Smag[:, None, None] * S
    idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    assert np.abs(tau - np.stack([full[:, i, j] for i, j in idx], 1)).max() < 1e-14


def test_backscatter_is_the_full_contraction_and_smagorinsky_never_backscatters():
    rng = np.random.default_rng(6)
    g = rng.normal(0, 1, (40, 9))
    tau = rng.normal(0, 1, (40, 6))
    A = g.reshape(40, 3, 3)
    S = 0.5 * (A + A.transpose(0, 2, 1))
    T = np.zeros((40, 3, 3))
    for c, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        T[:, i, j] = T[:, j, i] = tau[:, c]
    assert np.abs(mo.backscatter(tau, g) - (T * S).sum(axis=(1, 2))).max() < 1e-13
    assert (mo.backscatter(mo.smag_pointwise(g, 0.7), g) <= 0).all()
