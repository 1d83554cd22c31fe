"""Ensembles of independent periodic-box instances of uneven grid sizes on one GPU.

No counterpart in the reference, which runs one grid per process.  The members are
grouped by grid size; each group is one batched context whose kernel launches cover
all of its members, and the whole ensemble step (every member's CFL dt and four
low-storage stages) can be captured as one hipGraph.  Each member advances exactly as
a single ``SpectralSolver`` of its grid would (``src/PySpec_3d.py:96-105``).
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib


class Ensemble:
    """``dims``: one (N1, N2, N3) per member; ``nu``: one viscosity per member.

    The ensemble state is one flat complex128 tensor (``new_state()``); ``member(q, i)``
    is member i's (3, N1, N2, N3/2+1) view.  ``dt`` tensors hold one value per member,
    indexed by ``slot(i)``."""

    def __init__(self, dims, nu, model=1, *, L=(2 * math.pi,) * 3, cfl=0.5, Cs=0.16, force=0.0, device=None):
        if not torch.cuda.is_available():
            raise _lib.PsxError("Ensemble needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.dims = [tuple(int(x) for x in d) for d in dims]
        self.count = len(self.dims)
        if len(nu) != self.count:
            raise ValueError("one viscosity per member")
        self.nu = [float(x) for x in nu]
        self.model = _lib.MODEL_IDS[model]
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
        L1, L2, L3 = (float(x) for x in L)
        self.params = _lib.psx_params(nu=self.nu[0], L1=L1, L2=L2, L3=L3, dx=1.0, dy=1.0, dz=1.0, cfl=float(cfl),
                                      Cs=float(Cs), Ct=0.1, tau_a=-10.0, tau_b=-10.0, kc=1, rotate=0,
                                      Om1=0.0, Om2=0.0, Om3=0.0, force=float(force))
        flat = (C.c_int * (3 * self.count))(*[n for d in self.dims for n in d])
        nus = (C.c_double * self.count)(*self.nu)
        ens = C.c_void_p()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            rc = self.lib.psx_ensemble_create(C.byref(ens), self.count, flat, nus, self.model, C.byref(self.params),
                                              self.device.index, C.c_void_p(stream))
        if rc != 0:
            raise _lib.PsxError(f"psx_ensemble_create failed ({rc})")
        self.ens = ens
        self.size = int(self.lib.psx_ensemble_size(ens))
        self.offsets = [int(self.lib.psx_ensemble_offset(ens, i)) for i in range(self.count)]
        self.slots = [int(self.lib.psx_ensemble_slot(ens, i)) for i in range(self.count)]

    def close(self):
        if getattr(self, "ens", None):
            self.lib.psx_ensemble_destroy(self.ens)
            self.ens = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def new_state(self) -> torch.Tensor:
        return torch.zeros(self.size, dtype=torch.complex128, device=self.device)

    def new_dt(self) -> torch.Tensor:
        return torch.zeros(self.count, dtype=torch.float64, device=self.device)

    def member(self, q: torch.Tensor, i: int) -> torch.Tensor:
        N1, N2, N3 = self.dims[i]
        n = 3 * N1 * N2 * (N3 // 2 + 1)
        return q[self.offsets[i]:self.offsets[i] + n].view(3, N1, N2, N3 // 2 + 1)

    def slot(self, i: int) -> int:
        return self.slots[i]

    # ---------------------------------------------------------------- stepping
    def _check(self, rc, what):
        # PSX_E_NONFINITE is a status (work enqueued; a reduction met NaN or Inf)
        if rc != 0 and rc != _lib.PSX_E_NONFINITE:
            raise _lib.PsxError(f"{what} failed ({rc}): {self.lib.psx_ensemble_last_error(self.ens).decode()}")
        return rc

    def finite(self) -> bool:
        """Wait for the ensemble's stream; False if any member's CFL or diagnostics
        reduction met a NaN or an infinite value."""
        return self._check(self.lib.psx_ensemble_status(self.ens), "psx_ensemble_status") == 0

    def _args(self, q, qs, dt):
        for name, x, n, dtype in (("q", q, self.size, torch.complex128), ("qs", qs, self.size, torch.complex128),
                                  ("dt", dt, self.count, torch.float64)):
            if x.dtype != dtype or x.numel() != n or not x.is_contiguous() or x.device != self.device:
                raise ValueError(f"{name} must be a contiguous {dtype} tensor of {n} values on {self.device}")
        s = torch.cuda.current_stream(self.device).cuda_stream
        self._check(self.lib.psx_ensemble_set_stream(self.ens, C.c_void_p(s)), "psx_ensemble_set_stream")
        return C.c_void_p(q.data_ptr()), C.c_void_p(qs.data_ptr()), C.c_void_p(dt.data_ptr())

    def step(self, q, qs, dt, t: float = 0.0):
        """Every member advances one step: q <- new state, qs <- its stage-4 input,
        dt[slot(i)] <- member i's CFL dt."""
        a = self._args(q, qs, dt)
        self._check(self.lib.psx_ensemble_step(self.ens, *a, float(t)), "psx_ensemble_step")

    def step_graph(self, q, qs, dt):
        """The same step replayed from one captured hipGraph (non-default current stream)."""
        a = self._args(q, qs, dt)
        self._check(self.lib.psx_ensemble_step_graph(self.ens, *a), "psx_ensemble_step_graph")

    def autotune(self, q, reps: int = 3) -> list:
        """Choose each size class's kernel shapes by the time of the whole ensemble step
        (classes running concurrently), one knob at a time, on steps from a copy of
        ``q``; returns one tuning dict per class, classes in order of their first
        members."""
        from .solver import SpectralSolver

        n = self.lib.psx_ensemble_classes(self.ens)
        if n < 0:
            raise _lib.PsxError(f"psx_ensemble_classes failed ({n})")
        keys = SpectralSolver.TUNING_KEYS
        out = (C.c_int * (n * len(keys)))()
        a = self._args(q, q, self.new_dt())
        self._check(self.lib.psx_ensemble_autotune(self.ens, a[0], int(reps), out), "psx_ensemble_autotune")
        torch.cuda.current_stream(self.device).synchronize()
        return [dict(zip(keys, list(out)[k * len(keys):(k + 1) * len(keys)])) for k in range(n)]

    def set_closure(self, net, bound: float = float("inf"), clip_backscatter: bool = False) -> None:
        """Attach a learned-closure network to every member of a ``"learned"`` ensemble."""
        self._closure = net
        flags = _lib.PSX_CLOSURE_CLIP_BACKSCATTER if clip_backscatter else 0
        self._check(self.lib.psx_ensemble_set_closure(self.ens, net.net, float(bound), flags),
                    "psx_ensemble_set_closure")

    def stats(self, q) -> torch.Tensor:
        """(E, Omega, E_resolved) per member, rows in member order."""
        out = torch.empty(3 * self.count, dtype=torch.float64, device=self.device)
        s = torch.cuda.current_stream(self.device).cuda_stream
        self._check(self.lib.psx_ensemble_set_stream(self.ens, C.c_void_p(s)), "psx_ensemble_set_stream")
        self._check(self.lib.psx_ensemble_stats(self.ens, C.c_void_p(q.data_ptr()), C.c_void_p(out.data_ptr())),
                    "psx_ensemble_stats")
        out = out.view(self.count, 3)
        return out[torch.tensor(self.slots, device=self.device)]
