"""The learned subgrid closure's network: a pointwise MLP 9 -> 64 -> 64 -> 64 -> 6
(tanh) from filtered velocity-gradient components to the six SGS stresses, run with
BF16 operands and FP32 accumulation on the matrix cores (libpsx ``psx_mlp_*``).
No counterpart in the reference."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

NPARAM = _lib.MLP_NPARAM
# parameter blocks (offset, shape), row-major
LAYOUT = (("W1", 0, (64, 9)), ("b1", 576, (64,)), ("W2", 640, (64, 64)), ("b2", 4736, (64,)),
          ("W3", 4800, (64, 64)), ("b3", 8896, (64,)), ("W4", 8960, (6, 64)), ("b4", 9344, (6,)))


def _f32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    def __init__(self, seed: int = 0, device=None):
        if not torch.cuda.is_available():
            raise _lib.PsxError("MLP needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
        net = C.c_void_p()
        with torch.cuda.device(self.device):
            s = torch.cuda.current_stream(self.device).cuda_stream
            rc = self.lib.psx_mlp_create(C.byref(net), self.device.index, C.c_void_p(s), C.c_ulonglong(seed))
        if rc != 0:
            raise _lib.PsxError(f"psx_mlp_create failed ({rc})")
        self.net = net
        self.norm = (np.zeros(9, np.float32), np.ones(9, np.float32), np.zeros(6, np.float32), np.ones(6, np.float32))

    def close(self):
        if getattr(self, "net", None):
            self.lib.psx_mlp_destroy(self.net)
            self.net = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise _lib.PsxError(f"{what} failed ({rc}): {self.lib.psx_mlp_last_error(self.net).decode()}")

    def _stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        self._check(self.lib.psx_mlp_set_stream(self.net, C.c_void_p(s)), "psx_mlp_set_stream")

    @staticmethod
    def _arg(t, cols):
        if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != cols or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"expected a contiguous float32 (n, {cols}) device tensor")
        return C.c_void_p(t.data_ptr())

    # ---------------------------------------------------------------- parameters
    def params(self) -> np.ndarray:
        out = np.empty(NPARAM, np.float32)
        self._stream()
        self._check(self.lib.psx_mlp_get_params(self.net, _f32p(out)), "psx_mlp_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != NPARAM:
            raise ValueError(f"{NPARAM} parameters expected")
        self._stream()
        self._check(self.lib.psx_mlp_set_params(self.net, _f32p(p)), "psx_mlp_set_params")

    def set_norm(self, x_mean, x_scale, y_mean, y_scale) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (x_mean, x_scale, y_mean, y_scale)]
        self._check(self.lib.psx_mlp_set_norm(self.net, *[_f32p(a) for a in arrs]), "psx_mlp_set_norm")
        self.norm = tuple(arrs)

    # ---------------------------------------------------------------- compute
    def forward(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        n = x.shape[0]
        out = torch.empty((n, 6), dtype=torch.float32, device=self.device) if out is None else out
        self._stream()
        self._check(self.lib.psx_mlp_forward(self.net, self._arg(x, 9), self._arg(out, 6), n), "psx_mlp_forward")
        return out

    def closure(self, g: torch.Tensor, c2: float, bound: float = float("inf"), clip_backscatter: bool = False,
                out: torch.Tensor | None = None) -> torch.Tensor:
        """The stresses the learned LES uses for gradients g (n, 9): the network, or the
        Smagorinsky stress -2 c2 |S| S_ij outside [-bound, bound] (and at backscatter
        points with ``clip_backscatter``)."""
        n = g.shape[0]
        out = torch.empty((n, 6), dtype=torch.float32, device=self.device) if out is None else out
        self._stream()
        flags = _lib.PSX_CLOSURE_CLIP_BACKSCATTER if clip_backscatter else 0
        self._check(self.lib.psx_mlp_closure(self.net, self._arg(g, 9), self._arg(out, 6), n, float(c2), float(bound),
                                             flags), "psx_mlp_closure")
        return out

    def grad(self, x: torch.Tensor, target: torch.Tensor):
        """(loss, gradient) of the normalised MSE on one batch (n a multiple of 16)."""
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        g = torch.empty(NPARAM, dtype=torch.float32, device=self.device)
        self._stream()
        self._check(self.lib.psx_mlp_grad(self.net, self._arg(x, 9), self._arg(target, 6), x.shape[0],
                                          C.c_void_p(loss.data_ptr())), "psx_mlp_grad")
        self._check(self.lib.psx_mlp_gradient(self.net, C.c_void_p(g.data_ptr())), "psx_mlp_gradient")
        return loss, g

    def train_step(self, x: torch.Tensor, target: torch.Tensor, lr: float = 1e-3) -> torch.Tensor:
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        self._stream()
        self._check(self.lib.psx_mlp_train_step(self.net, self._arg(x, 9), self._arg(target, 6), x.shape[0],
                                                C.c_float(lr), C.c_void_p(loss.data_ptr())), "psx_mlp_train_step")
        return loss


# -------------------------------------------------------------------- training data
def training_pairs(dns, uhat_dns: torch.Tensor, les, delta: float | None = None, kind: str = "box",
                   deviatoric: bool = False):
    """Filtered-DNS training pairs on an LES grid.

    ``dns`` and ``les`` are ``SpectralSolver`` contexts of the DNS and LES grids (the
    same box).  The DNS velocity is filtered (``kind``, width ``delta``; default twice
    the LES grid spacing), its SGS stress tau_ij = G(u_i u_j) - G(u_i) G(u_j) formed
    with 3/2-rule products (``computeSGS_DNS``, src/Classes.py:861-929), and both are
    restricted to the LES modes.  Returns (x, t): float32 device tensors (n, 9) of the
    filtered velocity gradients du_i/dx_j (column 3i+j) and (n, 6) of the stresses
    (order 11 22 33 12 13 23) at the n = N1 N2 N3 LES grid points.  With ``deviatoric``
    the targets are tau_ij - delta_ij tau_kk/3: the isotropic part, which the pressure
    projection removes from the RHS, is left out of the fit."""
    if delta is None:
        delta = 2.0 * les.params.dx
    dims = (dns.N1, dns.N2, dns.N3)
    ubar = dns.filter(uhat_dns, kind, delta)
    _, tau = dns.sgs_dns(uhat_dns, kind, delta, w0=None, tau=True)
    u_les = les.restrict(ubar, dims)
    tau_les = les.restrict(tau, dims)
    x = les.gradients(u_les)
    t = les.physical(tau_les).reshape(6, -1)
    if deviatoric:
        t[:3] -= t[:3].mean(dim=0, keepdim=True)
    return x, t.t().contiguous().to(torch.float32)


def normalisation(x: torch.Tensor, t: torch.Tensor):
    """Per-column means and standard deviations of inputs and targets (host float32)."""
    xs = x.double().cpu().numpy()
    ts = t.double().cpu().numpy()
    sx = xs.std(0)
    sy = ts.std(0)
    return (xs.mean(0).astype(np.float32), np.where(sx > 0, sx, 1.0).astype(np.float32),
            ts.mean(0).astype(np.float32), np.where(sy > 0, sy, 1.0).astype(np.float32))


def input_range(net: MLP, x: torch.Tensor) -> float:
    """Largest |normalised input| over x: the closure's range bound for inputs like x."""
    mx, sx = net.norm[0], net.norm[1]
    xs = x.cpu().numpy()
    return float(np.abs((xs - mx) / sx).max())


def fit(net: MLP, x: torch.Tensor, t: torch.Tensor, *, steps: int, batch: int = 65536, lr: float = 1e-3,
        lr_final: float | None = None, seed: int = 0, set_norm: bool = True):
    """Train ``net`` on (x, t) with Adam: ``steps`` steps of ``batch`` points drawn from
    one shuffle per pass over the data; the learning rate decays geometrically from
    ``lr`` to ``lr_final`` (default lr/10).  Returns the per-step losses (host floats)."""
    if set_norm:
        net.set_norm(*normalisation(x, t))
    n = x.shape[0]
    batch = min(batch, n) // 16 * 16
    if batch <= 0:
        raise ValueError("need at least 16 training points")
    lr_final = lr / 10 if lr_final is None else lr_final
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    losses = []
    pos, perm = n, None
    for k in range(steps):
        if pos + batch > n:
            perm = torch.randperm(n, generator=g, device=x.device)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        xb = x.index_select(0, idx).contiguous()
        tb = t.index_select(0, idx).contiguous()
        rate = lr * (lr_final / lr) ** (k / max(steps - 1, 1))
        losses.append(net.train_step(xb, tb, rate))
    return [float(v) for v in torch.cat(losses).cpu()]


def _strain(g: np.ndarray) -> np.ndarray:
    return np.stack([g[:, 0], g[:, 4], g[:, 8], 0.5 * (g[:, 1] + g[:, 3]), 0.5 * (g[:, 2] + g[:, 6]),
                     0.5 * (g[:, 5] + g[:, 7])], axis=1)


def subgrid_dissipation(tau: np.ndarray, g: np.ndarray) -> float:
    """Mean subgrid dissipation -<tau_ij S_ij> of stresses (n, 6) (order 11 22 33 12 13 23)
    for gradients g (n, 9)."""
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    return float(-(np.asarray(tau, np.float64) * _strain(np.asarray(g, np.float64)) * w).sum(1).mean())


def calibrate_dissipation(net: MLP, x: torch.Tensor, t: torch.Tensor, c2: float, bound: float = float("inf"),
                          clip_backscatter: bool = True, max_points: int = 1 << 20) -> float:
    """Scale the network's outputs (its output mean and scale) by one factor so that the
    closure's mean subgrid dissipation on (x, t) equals that of the targets.  A network
    fitted by least squares predicts the conditional mean stress, whose energy drain to
    the subgrid scales need not match the targets'; this matches it.  Returns the
    factor (1 when either drain is not positive)."""
    n = x.shape[0]
    step = max(1, n // max_points)
    xs, ts = x[::step].contiguous(), t[::step].contiguous()
    g, tt = xs.cpu().numpy(), ts.cpu().numpy()
    d_ref = subgrid_dissipation(tt, g)
    total = 1.0
    # the Smagorinsky points are not scaled, hence a few fixed-point passes
    for _ in range(4):
        d_net = subgrid_dissipation(net.closure(xs, c2, bound, clip_backscatter).cpu().numpy(), g)
        if not (d_net > 0 and d_ref > 0):
            break
        a = d_ref / d_net
        if abs(a - 1.0) < 1e-3:
            break
        mx, sx, my, sy = net.norm
        net.set_norm(mx, sx, my * a, sy * a)
        total *= a
    return total
