"""GPU-resident pseudo-spectral solver over libpsx (one context per grid/model/GPU).

State tensors are ``torch.complex128`` device tensors of shape ``(3, N1, N2, N3//2+1)``
in the reference's per-field Fourier layout (``src/Classes.py:19-25``), with
``uhat = rfftn(u)/sqrt(N1 N2 N3)``.  PyTorch provides device memory and the stream;
every arithmetic operation runs in libpsx's HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import math
import threading

import torch

from . import _lib

_REGISTRY: dict[int, "SpectralSolver"] = {}
_REG_LOCK = threading.Lock()
_NEXT_ID = [1]


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class SpectralSolver:
    """A libpsx context: plans, twiddle tables and pipeline scratch for one grid.

    Parameters mirror the reference's case globals (``src/PySpec_3d.py:9-72``):
    ``nu``, ``L1..L3``, ``dx..dz`` (case grid spacings, used for Delta and the
    CFL limit), ``cfl`` (<= 0 means fixed dt = -cfl), ``Cs`` (0.16 in
    ``src/RHSfunctions.py:197``), ``Ct``, ``tau_a``/``tau_b`` (``dt0``/``dt1``),
    ``kc``, ``rotate`` and ``Om``; ``force`` is a linear forcing coefficient A (RHS += A u,
    no counterpart in the reference; every model).
    """

    def __init__(self, N1, N2, N3, model=0, *, nu, L=(2 * math.pi,) * 3, dx=None, dy=None, dz=None,
                 cfl=-0.01, Cs=0.16, Ct=0.1, tau_a=-10.0, tau_b=-10.0, kc=None, rotate=0,
                 Om=(0.0, 0.0, 0.0), force=0.0, device=None):
        if not torch.cuda.is_available():
            raise _lib.PsxError("SpectralSolver needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.N1, self.N2, self.N3 = int(N1), int(N2), int(N3)
        self.Nc = self.N3 // 2 + 1
        self.model = _lib.MODEL_IDS[model]
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
        L1, L2, L3 = (float(x) for x in L)
        self.params = _lib.psx_params(
            nu=float(nu), L1=L1, L2=L2, L3=L3,
            dx=float(L1 / N1 if dx is None else dx), dy=float(L2 / N2 if dy is None else dy),
            dz=float(L3 / N3 if dz is None else dz), cfl=float(cfl), Cs=float(Cs), Ct=float(Ct),
            tau_a=float(tau_a), tau_b=float(tau_b), kc=int(N1 // 2 if kc is None else kc),
            rotate=int(rotate), Om1=float(Om[0]), Om2=float(Om[1]), Om3=float(Om[2]), force=float(force))
        ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            rc = self.lib.psx_create(C.byref(ctx), self.N1, self.N2, self.N3, self.model,
                                     C.byref(self.params), self.device.index, C.c_void_p(stream))
        if rc != 0:
            raise _lib.PsxError(f"psx_create failed ({rc}) for {N1}x{N2}x{N3} model {model}")
        self.ctx = ctx
        self._dt = torch.zeros(1, dtype=torch.float64, device=self.device)
        with _REG_LOCK:
            self.id = _NEXT_ID[0]
            _NEXT_ID[0] += 1
            _REGISTRY[self.id] = self

    # ------------------------------------------------------------------ plumbing
    @property
    def shape(self):
        return (3, self.N1, self.N2, self.Nc)

    def empty_state(self) -> torch.Tensor:
        return torch.empty(self.shape, dtype=torch.complex128, device=self.device)

    def _sync_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.ctx, self.lib.psx_set_stream(self.ctx, C.c_void_p(s)), "psx_set_stream")

    def _state_arg(self, t: torch.Tensor, name: str) -> torch.Tensor:
        if t.dtype != torch.complex128 or tuple(t.shape) != self.shape:
            raise ValueError(f"{name} must be complex128 of shape {self.shape}, got {t.dtype} {tuple(t.shape)}")
        if t.device != self.device:
            raise ValueError(f"{name} must live on {self.device}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        return t

    def set_params(self, **kw):
        for k, v in kw.items():
            if k == "Om":
                self.params.Om1, self.params.Om2, self.params.Om3 = (float(x) for x in v)
            else:
                setattr(self.params, k, type(getattr(self.params, k))(v))
        _lib.check(self.ctx, self.lib.psx_set_params(self.ctx, C.byref(self.params)), "psx_set_params")

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.psx_destroy(self.ctx)
            self.ctx = None
            with _REG_LOCK:
                _REGISTRY.pop(self.id, None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ operations
    def rhs(self, uhat: torch.Tensor, t: float = 0.0, *, out=None, uhat_model=None, w0=None) -> torch.Tensor:
        """``computeRHS_<model>``: RHS(uhat) at step-start time ``t``."""
        self._state_arg(uhat, "uhat")
        out = self.empty_state() if out is None else self._state_arg(out, "out")
        um = 0 if uhat_model is None else _ptr(self._state_arg(uhat_model, "uhat_model"))
        w = 0 if w0 is None else _ptr(self._state_arg(w0, "w0"))
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_rhs(self.ctx, _ptr(uhat), _ptr(out), um or None, w or None, float(t)),
                   "psx_rhs")
        return out

    def rk_stage(self, q0, uin, out, dt: torch.Tensor, c: float, t: float = 0.0, zero_nyquist_out=False):
        for n, x in (("q0", q0), ("uin", uin), ("out", out)):
            self._state_arg(x, n)
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_rk_stage(self.ctx, _ptr(q0), _ptr(uin), _ptr(out), _ptr(dt), float(c),
                                                   float(t), int(bool(zero_nyquist_out))), "psx_rk_stage")
        return out

    def compute_dt(self, uhat: torch.Tensor, dt: torch.Tensor | None = None) -> torch.Tensor:
        self._state_arg(uhat, "uhat")
        dt = self._dt if dt is None else dt
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_compute_dt(self.ctx, _ptr(uhat), _ptr(dt)), "psx_compute_dt")
        return dt

    def step(self, q: torch.Tensor, qs: torch.Tensor, t: float = 0.0, dt: torch.Tensor | None = None) -> torch.Tensor:
        """``advanceQ_RK4``: q <- new state, qs <- stage-4 input (stale main.uhat).
        Returns the device dt tensor used for the step."""
        self._state_arg(q, "q")
        self._state_arg(qs, "qs")
        dt = self._dt if dt is None else dt
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_step(self.ctx, _ptr(q), _ptr(qs), _ptr(dt), float(t)), "psx_step")
        return dt

    def step_graph(self, q: torch.Tensor, qs: torch.Tensor, dt: torch.Tensor, t: torch.Tensor) -> None:
        """One step replayed from a captured hipGraph; ``t`` (device float64, the step-start
        time) is advanced by dt on the device.  Needs a non-default current stream."""
        self._state_arg(q, "q")
        self._state_arg(qs, "qs")
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_step_graph(self.ctx, _ptr(q), _ptr(qs), _ptr(dt), _ptr(t)),
                   "psx_step_graph")

    TUNING_KEYS = ("x_inverse_tile", "y_inverse_tile", "y_forward_tile", "z_variant", "yz_chunks", "yz_streams",
                   "closure_separate", "x_inverse_order", "y_inverse_order", "y_forward_order",
                   "stress_sums_in_epilogue")

    def tuning(self) -> dict:
        t = (C.c_int * len(self.TUNING_KEYS))()
        _lib.check(self.ctx, self.lib.psx_get_tuning(self.ctx, t), "psx_get_tuning")
        return dict(zip(self.TUNING_KEYS, list(t)))

    def set_tuning(self, tuning: dict) -> None:
        t = (C.c_int * len(self.TUNING_KEYS))(*[int(tuning.get(k, 0)) for k in self.TUNING_KEYS])
        _lib.check(self.ctx, self.lib.psx_set_tuning(self.ctx, t), "psx_set_tuning")

    def autotune(self, uhat: torch.Tensor, reps: int = 5) -> dict:
        """Time every tile width / z-pass variant / chunking on ``reps`` RHS evaluations of
        ``uhat`` and keep the fastest of each (DNS, Smagorinsky, DSMAG, learned)."""
        self._sync_stream()
        t = (C.c_int * len(self.TUNING_KEYS))()
        _lib.check(self.ctx, self.lib.psx_autotune(self.ctx, _ptr(uhat), int(reps), t), "psx_autotune")
        return dict(zip(self.TUNING_KEYS, list(t)))

    _FILTERS = {"sharp": _lib.PSX_FILTER_SHARP, "box": _lib.PSX_FILTER_BOX}

    def filter(self, uhat: torch.Tensor, kind: str = "sharp", delta: float = 0.0, out=None) -> torch.Tensor:
        """G * uhat: ``kind="sharp"`` is grid.Gf / myFilter at the context's kc; ``"box"``
        the top-hat of width ``delta``."""
        self._state_arg(uhat, "uhat")
        out = self.empty_state() if out is None else self._state_arg(out, "out")
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_filter(self.ctx, _ptr(uhat), _ptr(out), self._FILTERS[kind], float(delta)),
                   "psx_filter")
        return out

    def sgs_dns(self, uhat: torch.Tensor, kind: str = "sharp", delta: float = 0.0, *, w0=True, tau=False):
        """``computeSGS_DNS``: (w0 (3, N1, N2, Nc) or None, tau_hat (6, N1, N2, Nc) or None)."""
        self._state_arg(uhat, "uhat")
        w = self.empty_state() if w0 is True else (None if w0 is None or w0 is False else self._state_arg(w0, "w0"))
        t = None
        if tau is True:
            t = torch.empty((6,) + self.shape[1:], dtype=torch.complex128, device=self.device)
        elif tau is not None and tau is not False:
            t = tau
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_sgs_dns(self.ctx, _ptr(uhat), _ptr(w) if w is not None else None,
                                                  _ptr(t) if t is not None else None, self._FILTERS[kind],
                                                  float(delta)), "psx_sgs_dns")
        return w, t

    def physical(self, fhat: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """irfftn(fhat) * sqrt(Nt) of (nf, N1, N2, N3/2+1) spectral fields: (nf, N1, N2, N3) float64."""
        nf = fhat.shape[0]
        if fhat.dtype != torch.complex128 or tuple(fhat.shape[1:]) != self.shape[1:] or not fhat.is_contiguous():
            raise ValueError(f"fhat must be a contiguous complex128 (nf,) + {self.shape[1:]} tensor")
        out = torch.empty((nf, self.N1, self.N2, self.N3), dtype=torch.float64, device=self.device) \
            if out is None else out
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_physical(self.ctx, _ptr(fhat), int(nf), _ptr(out)), "psx_physical")
        return out

    def spectral(self, phys: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """rfftn(phys) / sqrt(Nt) of (nf, N1, N2, N3) real fields: (nf, N1, N2, N3/2+1) complex128
        (the inverse of ``physical``)."""
        nf = phys.shape[0]
        if phys.dtype != torch.float64 or tuple(phys.shape[1:]) != (self.N1, self.N2, self.N3) \
                or not phys.is_contiguous():
            raise ValueError(f"phys must be a contiguous float64 (nf, {self.N1}, {self.N2}, {self.N3}) tensor")
        out = torch.empty((nf,) + self.shape[1:], dtype=torch.complex128, device=self.device) \
            if out is None else out
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_spectral(self.ctx, _ptr(phys), int(nf), _ptr(out)), "psx_spectral")
        return out

    def gradients(self, uhat: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """du_i/dx_j at the grid points: float32 (N1*N2*N3, 9), column 3i+j."""
        self._state_arg(uhat, "uhat")
        out = torch.empty((self.N1 * self.N2 * self.N3, 9), dtype=torch.float32, device=self.device) \
            if out is None else out
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_gradients(self.ctx, _ptr(uhat), _ptr(out)), "psx_gradients")
        return out

    def restrict(self, src: torch.Tensor, src_dims, out: torch.Tensor | None = None) -> torch.Tensor:
        """Spectral fields (nf, S1, S2, S3/2+1) of a finer grid restricted to this grid."""
        s1, s2, s3 = (int(x) for x in src_dims)
        nf = src.shape[0]
        if src.dtype != torch.complex128 or tuple(src.shape[1:]) != (s1, s2, s3 // 2 + 1) or not src.is_contiguous():
            raise ValueError("src must be a contiguous complex128 (nf, S1, S2, S3/2+1) tensor")
        out = torch.empty((nf,) + self.shape[1:], dtype=torch.complex128, device=self.device) if out is None else out
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_restrict(self.ctx, _ptr(src), s1, s2, s3, int(nf), _ptr(out)),
                   "psx_restrict")
        return out

    def set_closure(self, net, bound: float = float("inf"), clip_backscatter: bool = False) -> None:
        """Attach a learned-closure network (``learned.MLP``) to a ``"learned"`` solver;
        points whose normalised inputs leave [-bound, bound] (and, with
        ``clip_backscatter``, points where the network's stress is backscatter) take the
        Smagorinsky stress."""
        self._closure = net  # keep the network alive while attached
        flags = _lib.PSX_CLOSURE_CLIP_BACKSCATTER if clip_backscatter else 0
        _lib.check(self.ctx, self.lib.psx_set_closure(self.ctx, net.net, float(bound), flags), "psx_set_closure")

    def stats(self, uhat: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """(E, Omega, E_resolved) on the device."""
        self._state_arg(uhat, "uhat")
        out = torch.empty(3, dtype=torch.float64, device=self.device) if out is None else out
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_stats(self.ctx, _ptr(uhat), _ptr(out)), "psx_stats")
        return out

    def spectrum(self, uhat: torch.Tensor, resolved: bool = False) -> torch.Tensor:
        self._state_arg(uhat, "uhat")
        nb = self.lib.psx_spectrum_bins(self.ctx)
        out = torch.empty((nb, 3), dtype=torch.float64, device=self.device)
        self._sync_stream()
        _lib.check(self.ctx, self.lib.psx_spectrum(self.ctx, _ptr(uhat), _ptr(out), int(bool(resolved))),
                   "psx_spectrum")
        return out

    def tau_values(self):
        """Dynamic tau-model: (tau, tau used, LE re, im, PLQLE_filt re, im, filt_PLQLE re, im)."""
        out = (C.c_double * 8)()
        _lib.check(self.ctx, self.lib.psx_tau(self.ctx, out), "psx_tau")
        return list(out)

    def finite(self) -> bool:
        """Wait for the context's stream; False if a CFL or diagnostics reduction since
        creation (or the last ``clear_status``) met a NaN or an infinite value."""
        rc = self.lib.psx_status(self.ctx)
        return _lib.check(self.ctx, rc, "psx_status") == _lib.PSX_OK

    def clear_status(self) -> None:
        _lib.check(self.ctx, self.lib.psx_clear_status(self.ctx), "psx_clear_status")

    def scratch_bytes(self) -> int:
        return int(self.lib.psx_scratch_bytes(self.ctx))


def get(solver_id: int) -> SpectralSolver:
    return _REGISTRY[solver_id]
