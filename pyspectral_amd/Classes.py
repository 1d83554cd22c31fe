"""Drop-in counterparts of PySpectral's ``src/Classes.py`` objects, GPU-resident.

``variables``, ``gridclass``, ``FFTclass`` and ``utilitiesClass`` keep the
reference's constructor signatures and attribute names.  The solution lives on
the GPU as two complex128 tensors of shape ``(3, N1, N2, N3//2+1)``:

* ``main.Qd`` -- the state the reference keeps interleaved in ``main.Q``
  (``src/Classes.py:30-34``);
* ``main.Ud`` -- the velocity the reference keeps in ``main.uhat/vhat/what``
  (which after a step is the stale stage-4 input, ``src/PySpec_3d.py:96-105``).

``main.Q``, ``main.uhat``, ``main.vhat`` and ``main.what`` are numpy views
materialised on access (a device-to-host copy) and uploaded on assignment, so
case scripts and savehooks that read or write them keep working.
"""
from __future__ import annotations

import math
import sys

import numpy as np
import torch

from . import RHSfunctions as _rhs
from .solver import SpectralSolver

_MODEL_RHS = {
    0: "computeRHS_NOSGS",
    1: "computeRHS_SMAG",
    2: "computeRHS_tmodel",
    "Static tau-model": "computeRHS_staumodel",
    "Dynamic tau-model": "computeRHS_Dtaumodel",
    7: "computeRHS_DSMAG",
}


class gridclass:
    """``gridclass`` (``src/Classes.py:548-588``).  Wavenumber arrays are built on first
    access (host numpy, for savehooks and post-processing); the kernels compute
    wavenumbers from indices."""

    def __init__(self, N1, N2, N3, x, y, z, kc, L1=2 * np.pi, L2=2 * np.pi, L3=2 * np.pi):
        self.N1, self.N2, self.N3 = int(N1), int(N2), int(N3)
        self.L1, self.L2, self.L3 = float(L1), float(L2), float(L3)
        self.x, self.y, self.z = np.array(x, dtype=float), np.array(y, dtype=float), np.array(z, dtype=float)
        self.dx = float(self.x[1, 0, 0] - self.x[0, 0, 0])
        self.dy = float(self.y[0, 1, 0] - self.y[0, 0, 0])
        self.dz = float(self.z[0, 0, 1] - self.z[0, 0, 0])
        self.kc = int(kc)
        self.Delta = (self.dx * self.dy * self.dz) ** (1.0 / 3.0)
        self._k = None

    def _build(self):
        N1, N2, N3 = self.N1, self.N2, self.N3
        k1 = np.fft.fftshift(np.linspace(-N1 // 2, N1 // 2 - 1, N1)) * 2.0 * np.pi / self.L1
        k2 = np.fft.fftshift(np.linspace(-N2 // 2, N2 // 2 - 1, N2)) * 2.0 * np.pi / self.L2
        k3 = np.linspace(0, N3 // 2, N3 // 2 + 1) * 2.0 * np.pi / self.L3
        K1, K2, K3 = np.meshgrid(k1, k2, k3, indexing="ij")
        ksqr = K1 * K1 + K2 * K2 + K3 * K3 + 1.0e-50
        kc = self.kc
        Gf = np.zeros(K1.shape)
        Gf[0:kc, 0:kc, 0:kc] = 1
        Gf[0:kc, N2 - kc + 1:, 0:kc] = 1
        Gf[N1 - kc + 1:, 0:kc, 0:kc] = 1
        Gf[N1 - kc + 1:, N2 - kc + 1:, 0:kc] = 1
        self._k = dict(k1=K1, k2=K2, k3=K3, ksqr=ksqr, ksqr_i=1.0 / ksqr, Gf=Gf)

    def __getattr__(self, name):
        if name in ("k1", "k2", "k3", "ksqr", "ksqr_i", "Gf"):
            if self.__dict__.get("_k") is None:
                self._build()
            return self._k[name]
        raise AttributeError(name)


class FFTclass:
    """``FFTclass`` (``src/Classes.py:590-625``).  The reference builds pyFFTW plans here;
    the GPU transforms live in the libpsx context that ``variables`` creates (it needs
    the closure), which is attached here as ``solver``."""

    def __init__(self, N1, N2, N3, nthreads=1):
        self.nthreads = nthreads
        self.N1, self.N2, self.N3 = int(N1), int(N2), int(N3)
        self.scale = np.sqrt((3.0 / 2.0) ** 3 * np.sqrt(N1 * N2 * N3))
        self.solver = None


def _upload(a) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to("cuda")


class variables:
    """``variables`` (``src/Classes.py:7-229``) for the GPU models: turb_model 0 (DNS),
    1 (Smagorinsky), 2 (t-model), 'Static tau-model', 'Dynamic tau-model'."""

    def __init__(self, weave, turb_model, rotate, Om1, Om2, Om3, grid, uhat, vhat, what, t, dt, nu, Ct, dt0,
                 dt0_subintegrations, dt1, dt1_subintegrations, cfl, dim, myFFT=None):
        if turb_model not in _MODEL_RHS:
            raise NotImplementedError(f"turb_model {turb_model!r} has no GPU implementation "
                                      "(GPU models: 0, 1, 2, 'Static tau-model', 'Dynamic tau-model', 7)")
        if dim != 3:
            raise NotImplementedError("only dim = 3 is supported")
        self.turb_model = turb_model
        self.rotate, self.Om1, self.Om2, self.Om3 = int(rotate), float(Om1), float(Om2), float(Om3)
        self.t = float(t)
        self.t_state = self.t  # the time of the state in Qd (advanced by each step)
        self.kc = grid.N1 // 2 - 1  # np.amax(grid.k1) for L1 = 2 pi (src/Classes.py:16)
        self.dt = float(dt)
        self.nu = float(nu)
        self.cfl = float(cfl)
        self.Ct = 0.1 if Ct == -10 else float(Ct)
        self.tau_a, self.tau_b = float(dt0), float(dt1)
        if turb_model == "Dynamic tau-model":
            self.kf = int(grid.kc * 2.0 / 3.0)
        elif turb_model == "Static tau-model":
            self.kf = int(grid.kc / 2.0)
        self.nvars = 3
        # weave = 1 selects computeRHS_NOSGS_WEAVE for the DNS (src/Classes.py:43-51), whose
        # fused assembly has no Coriolis terms (src/RHSfunctionsWEAVE.py:44-66)
        self.weave = int(weave)
        weave_dns = self.weave != 0 and turb_model == 0
        self.solver = SpectralSolver(
            grid.N1, grid.N2, grid.N3, turb_model, nu=self.nu, L=(grid.L1, grid.L2, grid.L3),
            dx=grid.dx, dy=grid.dy, dz=grid.dz, cfl=self.cfl, Cs=0.16, Ct=self.Ct, tau_a=self.tau_a,
            tau_b=self.tau_b, kc=grid.kc, rotate=0 if weave_dns else self.rotate, Om=(self.Om1, self.Om2, self.Om3))
        if myFFT is not None:
            myFFT.solver = self.solver
        self.Ud = _upload(np.stack([uhat, vhat, what]))
        self.Qd = self.Ud.clone()
        self.w0d = None
        self._dt_dev = torch.zeros(1, dtype=torch.float64, device=self.Ud.device)
        self.computeRHS = _rhs.computeRHS_NOSGS_WEAVE if weave_dns else getattr(_rhs, _MODEL_RHS[turb_model])
        self.iteration = 0

    # -- the reference's Q2U / U2Q (src/Classes.py:35-42)
    def U2Q(self):
        self.Qd.copy_(self.Ud)

    def Q2U(self):
        self.Ud.copy_(self.Qd)

    # -- numpy views of the device state
    def _get_comp(self, c):
        torch.cuda.synchronize()
        return self.Ud[c].cpu().numpy()

    def _set_comp(self, c, val):
        self.Ud[c].copy_(torch.from_numpy(np.ascontiguousarray(val, dtype=np.complex128)))

    uhat = property(lambda s: s._get_comp(0), lambda s, v: s._set_comp(0, v))
    vhat = property(lambda s: s._get_comp(1), lambda s, v: s._set_comp(1, v))
    what = property(lambda s: s._get_comp(2), lambda s, v: s._set_comp(2, v))

    @property
    def Q(self):
        """The interleaved ``(3N1, 3N2, 3(N3/2+1))`` array of ``src/Classes.py:30-34``."""
        torch.cuda.synchronize()
        q = self.Qd.cpu().numpy()
        _, n1, n2, nc = q.shape
        out = np.zeros((3 * n1, 3 * n2, 3 * nc), dtype=complex)
        for c in range(3):
            out[c::3, c::3, c::3] = q[c]
        return out

    @Q.setter
    def Q(self, val):
        val = np.asarray(val)
        self.Qd.copy_(torch.from_numpy(np.ascontiguousarray(np.stack([val[c::3, c::3, c::3] for c in range(3)]))))

    def _w0(self, c):
        if self.w0d is None:
            return None
        torch.cuda.synchronize()
        return self.w0d[c].cpu().numpy()[..., None]

    w0_u = property(lambda s: s._w0(0))
    w0_v = property(lambda s: s._w0(1))
    w0_w = property(lambda s: s._w0(2))


class utilitiesClass:
    """``utilitiesClass`` (``src/Classes.py:631-1385``): the per-step diagnostics on the
    device.  Values are computed from ``main.Ud`` (the reference's ``main.uhat``)."""

    def preAdvanceQ_hook(self, main, grid, myFFT):
        pass

    def postAdvanceQ_hook(self, main, grid, myFFT):
        pass

    def _stats(self, main):
        return main.solver.stats(main.Ud).cpu().numpy()

    def computeEnergy(self, main, grid):
        return float(self._stats(main)[0])

    def computeEnergy_resolved(self, main, grid):
        return float(self._stats(main)[2])

    def computeEnstrophy(self, main, grid):
        return float(self._stats(main)[1])

    def compute_dt(self, main, grid):
        dt = main.solver.compute_dt(main.Ud, main._dt_dev)
        main.dt = float(dt.item())

    def computeAllStats(self, main, grid):
        """``computeAllStats`` (``src/Classes.py:701-731``)."""
        E, Om, _ = self._stats(main)
        dissipation = 2 * Om * main.nu
        lambda_k = (main.nu ** 3 / dissipation) ** 0.25
        tau_k = (main.nu / dissipation) ** 0.5
        Re_lambda = E * np.sqrt(20.0 / (3.0 * main.nu * dissipation))
        return Om, E, dissipation, lambda_k, tau_k, Re_lambda

    def computeSpectrum(self, main, grid):
        s = main.solver.spectrum(main.Ud).cpu().numpy()
        return np.arange(s.shape[0], dtype=float), s

    def computeSpectrum_resolved(self, main, grid):
        s = main.solver.spectrum(main.Ud, resolved=True).cpu().numpy()
        return np.arange(s.shape[0], dtype=float), s

    def computeSGS_DNS(self, main, grid, myFFT):
        """``computeSGS_DNS`` (``src/Classes.py:861-929``): the subgrid term of the sharp
        filter grid.Gf (cutoff kc) from the resolved field; returns w0_u, w0_v, w0_w of
        shape (N1, N2, N3/2+1, 1) as the reference does."""
        w0, _ = main.solver.sgs_dns(main.Ud, "sharp")
        w = w0.cpu().numpy()
        return w[0][..., None], w[1][..., None], w[2][..., None]

    def myFilter(self, uhat, kc):
        """``myFilter`` (``src/Classes.py:1271-1279``): the sharp cutoff of one field at kc."""
        solver = _filter_solver(uhat.shape)
        import torch

        x = torch.zeros(solver.shape, dtype=torch.complex128, device=solver.device)
        x[0] = torch.from_numpy(np.ascontiguousarray(uhat)).to(solver.device)
        old = solver.params.kc
        solver.set_params(kc=int(kc))
        try:
            out = solver.filter(x, "sharp")
        finally:
            solver.set_params(kc=old)
        return out[0].cpu().numpy()


_FILTER_SOLVERS = {}


def _filter_solver(shape):
    """A DNS context of the field's grid, used for the filters of utilitiesClass."""
    from .solver import SpectralSolver

    N1, N2, Nc = (int(n) for n in shape[:3])
    N3 = 2 * (Nc - 1)
    key = (N1, N2, N3)
    if key not in _FILTER_SOLVERS:
        _FILTER_SOLVERS[key] = SpectralSolver(N1, N2, N3, 0, nu=0.0)
    return _FILTER_SOLVERS[key]


def _warn(msg):
    sys.stdout.write(msg + "\n")


__all__ = ["variables", "gridclass", "FFTclass", "utilitiesClass"]
del math
