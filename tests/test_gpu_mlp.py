"""The BF16/FP32 MFMA network kernels against a float64 numpy reference that rounds to
BF16 at the same points (operands of every matrix product), and a training check."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import mlp_oracle  # noqa: E402
from oracle.mlp_oracle import unpack  # noqa: E402

ref_forward = mlp_oracle.forward
ref_grad = mlp_oracle.grad


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


def test_adam_steps_match_torch_adam():
    """The Adam update kernel (beta 0.9/0.999, eps 1e-8, FP32 moments) against
    torch.optim.Adam fed the same gradients: five steps at two learning rates, each
    step's parameter update compared (float32 rounding of the two formulas only)."""
    net, p, norm, rng = _net(5)
    n = 2048
    x = torch.from_numpy(rng.normal(0, 1.0, (n, 9)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.normal(0, 1.0, (n, 6)).astype(np.float32)).cuda()
    w = torch.nn.Parameter(torch.from_numpy(net.params().copy()))
    opt = torch.optim.Adam([w], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step, lr in enumerate((1e-3, 1e-3, 1e-3, 3e-4, 3e-4)):
        before = net.params().copy()
        _, g = net.grad(x, t)  # the gradient train_step computes (deterministic reductions)
        for grp in opt.param_groups:
            grp["lr"] = lr
        w.grad = g.cpu().clone()
        opt.step()
        net.train_step(x, t, lr)
        after = net.params()
        d_ours = after.astype(np.float64) - before
        d_ref = w.detach().numpy().astype(np.float64) - before
        rel = np.linalg.norm(d_ours - d_ref) / np.linalg.norm(d_ref)
        print(f"\nAdam step {step + 1} (lr {lr}): update relative L2 {rel:.2e}, |update| {np.linalg.norm(d_ref):.3e}")
        assert rel < 1e-4
        # both continue from the kernel's weights, so float32 rounding does not accumulate
        with torch.no_grad():
            w.copy_(torch.from_numpy(after.copy()))
    net.close()
