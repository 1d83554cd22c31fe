"""The BF16/FP32 MFMA network kernels against a float64 numpy reference that rounds to
BF16 at the same points (operands of every matrix product), and a training check."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    from pyspectral_amd.learned import LAYOUT

    return {k: p[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for k, o, s in LAYOUT}


def ref_forward(P, x, norm):
    mx, sx, my, sy = (np.asarray(a, np.float64) for a in norm)
    a0 = bf16(((x - mx) / sx).astype(np.float32))
    z1 = a0 @ bf16(P["W1"]).T + P["b1"]
    a1 = np.tanh(z1)
    z2 = bf16(a1) @ bf16(P["W2"]).T + P["b2"]
    a2 = np.tanh(z2)
    z3 = bf16(a2) @ bf16(P["W3"]).T + P["b3"]
    a3 = np.tanh(z3)
    yn = bf16(a3) @ bf16(P["W4"]).T + P["b4"]
    return a0, a1, a2, a3, yn, yn * sy + my


def ref_grad(P, x, t, norm):
    mx, sx, my, sy = (np.asarray(a, np.float64) for a in norm)
    a0, a1, a2, a3, yn, _ = ref_forward(P, x, norm)
    n = x.shape[0]
    e = yn - (t - my) / sy
    loss = (e ** 2).sum() / (6 * n)
    d4 = bf16(2 * e / (6 * n))
    d3 = (d4 @ bf16(P["W4"])) * (1 - a3 ** 2)
    d2 = (bf16(d3) @ bf16(P["W3"])) * (1 - a2 ** 2)
    d1 = (bf16(d2) @ bf16(P["W2"])) * (1 - a1 ** 2)
    G = {"W4": d4.T @ bf16(a3), "b4": d4.sum(0), "W3": bf16(d3).T @ bf16(a2), "b3": bf16(d3).sum(0),
         "W2": bf16(d2).T @ bf16(a1), "b2": bf16(d2).sum(0), "W1": bf16(d1).T @ a0, "b1": bf16(d1).sum(0)}
    from pyspectral_amd.learned import LAYOUT

    g = np.zeros(9350)
    for k, o, s in LAYOUT:
        g[o:o + int(np.prod(s))] = G[k].ravel()
    return loss, g


def _net(seed=3):
    from pyspectral_amd.learned import MLP

    net = MLP(seed=seed)
    rng = np.random.default_rng(seed)
    norm = (rng.normal(0, 0.3, 9), rng.uniform(0.5, 2.0, 9), rng.normal(0, 0.1, 6), rng.uniform(0.5, 1.5, 6))
    net.set_norm(*norm)
    p = net.params()
    p[576:640] = rng.normal(0, 0.1, 64)    # non-zero biases
    p[4736:4800] = rng.normal(0, 0.1, 64)
    p[8896:8960] = rng.normal(0, 0.1, 64)
    p[9344:] = rng.normal(0, 0.1, 6)
    net.set_params(p)
    return net, net.params(), tuple(np.asarray(a, np.float64) for a in net.norm), rng


def test_forward_matches_bf16_reference():
    net, p, norm, rng = _net()
    n = 1000
    x = rng.normal(0, 1.0, (n, 9)).astype(np.float32)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = ref_forward(unpack(p), x.astype(np.float64), norm)[-1]
    err = np.abs(y - ref)
    print(f"\nforward max abs err {err.max():.3e}, rms {np.sqrt((err ** 2).mean()):.3e}, |y| max {np.abs(ref).max():.3f}")
    assert np.sqrt((err ** 2).mean()) < 2e-3 * np.abs(ref).max()
    assert err.max() < 2e-2 * np.abs(ref).max()
    net.close()


def test_gradient_matches_bf16_reference():
    net, p, norm, rng = _net(4)
    n = 4096
    x = rng.normal(0, 1.0, (n, 9)).astype(np.float32)
    t = rng.normal(0, 1.0, (n, 6)).astype(np.float32)
    loss, g = net.grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    lref, gref = ref_grad(unpack(p), x.astype(np.float64), t.astype(np.float64), norm)
    g = g.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"\nloss {float(loss.item()):.6f} ref {lref:.6f}; gradient relative L2 {rel:.3e}")
    assert abs(float(loss.item()) - lref) < 1e-3 * lref
    assert rel < 1e-2
    net.close()


def test_training_reduces_loss():
    from pyspectral_amd.learned import MLP

    rng = np.random.default_rng(7)
    n = 65536
    x = rng.normal(0, 1.0, (n, 9)).astype(np.float32)
    A = rng.normal(0, 0.5, (9, 6))
    t = np.tanh(x @ A) + 0.1 * (x[:, :6] ** 2)
    net = MLP(seed=1)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t.astype(np.float32)).cuda()
    first = float(net.train_step(xd, td, 1e-3).item())
    for _ in range(300):
        last = net.train_step(xd, td, 1e-3)
    last = float(last.item())
    print(f"\nloss {first:.4f} -> {last:.4f}")
    assert last < 0.2 * first
    net.close()
