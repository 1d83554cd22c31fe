"""CPU oracle of the learned closure's network (no counterpart in the reference).

TEST INFRASTRUCTURE ONLY: the checker for the BF16/FP32 matrix-core kernels of
``pyspectral_amd/csrc/mlp.hip``.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it.

The network is a pointwise MLP 9 -> 64 -> 64 -> 64 -> 6 with tanh hidden layers.  The
GPU rounds every matrix-product operand to BF16 (round to nearest even) and
accumulates in FP32; this restatement rounds the same operands to BF16 and
accumulates in float64, so it differs from the GPU only by accumulation order and
width and by the hardware tanh/exp.  With ``rounding=False`` it is the plain
float64 network.  Pinned in ``tests/test_oracle_mlp.py``: the BF16 rounding against
torch's float32 -> bfloat16 conversion, the unrounded forward against
``torch.nn.Linear``/``tanh``, the unrounded gradient against torch autograd.
"""
from __future__ import annotations

import numpy as np

NPARAM = 9350
# parameter blocks (offset, shape), row-major, as the library stores them
LAYOUT = (("W1", 0, (64, 9)), ("b1", 576, (64,)), ("W2", 640, (64, 64)), ("b2", 4736, (64,)),
          ("W3", 4800, (64, 64)), ("b3", 8896, (64,)), ("W4", 8960, (6, 64)), ("b4", 9344, (6,)))


def bf16(a):
    """float32 -> bfloat16 (round to nearest even), returned as float64 values."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    return {k: np.asarray(p[o:o + int(np.prod(s))], np.float64).reshape(s) for k, o, s in LAYOUT}


def pack(P) -> np.ndarray:
    out = np.zeros(NPARAM)
    for k, o, s in LAYOUT:
        out[o:o + int(np.prod(s))] = np.asarray(P[k]).ravel()
    return out


def _r(a, rounding):
    return bf16(a) if rounding else np.asarray(a, np.float64)


def normalise_inputs(x, norm):
    """(x - mean) * (1 / scale) in float32, as the kernel forms it (the reciprocal
    rounded to float32 once)."""
    mx, sx = (np.asarray(a, np.float32) for a in norm[:2])
    isx = np.float32(1.0) / sx
    return (np.asarray(x, np.float32) - mx) * isx


def forward(P, x, norm, rounding=True):
    """Activations (a0, a1, a2, a3), normalised output yn and output y of n points x (n, 9)."""
    my, sy = (np.asarray(a, np.float64) for a in norm[2:])
    xn = normalise_inputs(x, norm) if rounding else \
        (np.asarray(x, np.float64) - np.asarray(norm[0], np.float64)) / np.asarray(norm[1], np.float64)
    a0 = _r(xn, rounding)
    a1 = np.tanh(a0 @ _r(P["W1"], rounding).T + P["b1"])
    a2 = np.tanh(_r(a1, rounding) @ _r(P["W2"], rounding).T + P["b2"])
    a3 = np.tanh(_r(a2, rounding) @ _r(P["W3"], rounding).T + P["b3"])
    yn = _r(a3, rounding) @ _r(P["W4"], rounding).T + P["b4"]
    return a0, a1, a2, a3, yn, yn * sy + my


def grad(P, x, t, norm, rounding=True):
    """(loss, flat gradient) of the mean squared error of the normalised outputs,
    loss = sum (yn - (t - my)/sy)^2 / (6 n); the deltas are rounded where the kernels
    round them (the operands of the backward and weight-gradient products)."""
    my, sy = (np.asarray(a, np.float64) for a in norm[2:])
    a0, a1, a2, a3, yn, _ = forward(P, x, norm, rounding)
    n = x.shape[0]
    e = yn - (np.asarray(t, np.float64) - my) / sy
    loss = (e ** 2).sum() / (6 * n)
    R = (lambda a: _r(a, rounding))
    d4 = R(2 * e / (6 * n))
    d3 = (d4 @ R(P["W4"])) * (1 - a3 ** 2)
    d2 = (R(d3) @ R(P["W3"])) * (1 - a2 ** 2)
    d1 = (R(d2) @ R(P["W2"])) * (1 - a1 ** 2)
    G = {"W4": d4.T @ R(a3), "b4": d4.sum(0), "W3": R(d3).T @ R(a2), "b3": R(d3).sum(0),
         "W2": R(d2).T @ R(a1), "b2": R(d2).sum(0), "W1": R(d1).T @ a0, "b1": R(d1).sum(0)}
    return loss, pack(G)


def smag_pointwise(g, c2):
    """Smagorinsky stress -2 c2 |S| S_ij, |S| = sqrt(2 S_ij S_ij), of gradients g (n, 9),
    g[:, 3i+j] = du_i/dx_j; columns 11 22 33 12 13 23 (src/RHSfunctions.py:195-212)."""
    g = np.asarray(g, np.float64)
    S = np.stack([g[:, 0], g[:, 4], g[:, 8], 0.5 * (g[:, 1] + g[:, 3]), 0.5 * (g[:, 2] + g[:, 6]),
                  0.5 * (g[:, 5] + g[:, 7])], axis=1)
    Smag = np.sqrt(2.0 * (S[:, 0] ** 2 + S[:, 1] ** 2 + S[:, 2] ** 2
                          + 2.0 * (S[:, 3] ** 2 + S[:, 4] ** 2 + S[:, 5] ** 2)))
    return -2.0 * c2 * Smag[:, None] * S


def backscatter(tau, g):
    """tau_ij S_ij > 0 (negative subgrid dissipation -tau_ij S_ij) per point, for stresses
    (n, 6) in the order 11 22 33 12 13 23 and gradients (n, 9)."""
    g = np.asarray(g, np.float64)
    S = np.stack([g[:, 0], g[:, 4], g[:, 8], 0.5 * (g[:, 1] + g[:, 3]), 0.5 * (g[:, 2] + g[:, 6]),
                  0.5 * (g[:, 5] + g[:, 7])], axis=1)
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    return (np.asarray(tau, np.float64) * S * w).sum(axis=1)


def closure(P, norm, bound, c2, clip=False):
    """The kernel's closure as a function of float64 gradients (n, 9) -> stresses (n, 6):
    inputs cast to float32, the network where every |normalised input| <= bound (and,
    with ``clip``, where its stress is not backscatter), the Smagorinsky stress
    elsewhere; outputs rounded to float32 as the kernel stores them."""

    def f(g):
        gf = np.asarray(g, np.float32)
        xn = normalise_inputs(gf, norm)
        use_net = np.abs(xn).max(axis=1) <= bound
        y = smag_pointwise(gf, c2)
        net = forward(P, gf, norm)[-1].astype(np.float32).astype(np.float64)
        if clip:
            use_net &= backscatter(net, gf) <= 0.0
        y[use_net] = net[use_net]
        return y.astype(np.float32).astype(np.float64), use_net

    return f
