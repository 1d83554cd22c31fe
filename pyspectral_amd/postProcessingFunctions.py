"""The solution-file formats of the solver and their loaders.

Files (written by the driver in ``src/PySpec_3d.py:90-92,148`` and by the case
savehooks, ``run/TGV_Validate100/inputfile.py:47-56``):

* ``grid.npz``     k1, k2, k3 (N1, N2, N3/2+1) and the case's x, y, z arrays
* ``runinfo.npz``  turb_model, dt, save_freq, nu
* ``stats.npz``    Energy, Dissipation, Energy_resolved, Dissipation_resolved, t, Enstrophy
* ``npsol<it>.npz`` uhat, vhat, what (N1, N2, N3/2+1) complex128, t, and for a closure
  run w0_u, w0_v, w0_w (N1, N2, N3/2+1, 1), which ``variablesFromFile`` reads when
  turb_model != 0 (``src/postProcessingFunctions.py:5-29``).

The loaders mirror ``src/postProcessingFunctions.py``; they only read files (numpy,
no pickles).  ``FFTFromFile``'s transform objects run on the GPU.

For an exact restart ``save_npsol`` also writes ``Qhat`` (3, N1, N2, N3/2+1), the
state the next step starts from, its time ``t_Qhat`` (a savehook's ``t`` is the start
of the step just taken) and ``iteration``: the reference's ``uhat`` is the
stale stage-4 velocity (``src/PySpec_3d.py:96-105``), from which neither the state nor
(for the CFL limit) the next dt could be rebuilt.  ``PySpec_3d.run_case(..., resume=)``
restarts from such a file bit for bit.
"""
from __future__ import annotations

import os

import numpy as np


def save_npsol(main, iteration: int, outdir: str = "3DSolution", compressed: bool = True) -> str:
    """Write ``npsol<iteration>.npz`` as the case savehooks do (uhat, vhat, what, t;
    ``run/TGV_Validate100/inputfile.py:56``), adding w0_u, w0_v, w0_w for a closure
    model so that ``variablesFromFile`` can read the file back."""
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, f"npsol{iteration}")
    data = dict(uhat=main.uhat, vhat=main.vhat, what=main.what, t=main.t)
    # the step's starting state, for an exact restart (not in the reference's files)
    qd = getattr(main, "Qd", None)
    if qd is not None:
        import torch

        torch.cuda.synchronize()
        data["Qhat"] = qd.cpu().numpy()
        data["t_Qhat"] = float(getattr(main, "t_state", main.t))  # t holds the step's start in a savehook
        data["iteration"] = int(iteration)
    if main.turb_model != 0:
        for name in ("w0_u", "w0_v", "w0_w"):
            w = getattr(main, name)
            data[name] = np.zeros(np.shape(main.uhat) + (1,), dtype=complex) if w is None else w
    (np.savez_compressed if compressed else np.savez)(path, **data)
    return path + ".npz"


def load_stats(path: str) -> dict:
    """``stats.npz`` as a dict of arrays."""
    with np.load(path) as f:
        return {k: f[k] for k in f.files}


class variablesFromFile:
    """``variablesFromFile`` (``src/postProcessingFunctions.py:5-37``)."""

    def __init__(self, gridFile, solutionFile, infoFile, kc):
        self.turb_model = infoFile["turb_model"].mean()
        self.kc = kc
        self.uhat = np.array(solutionFile["uhat"], dtype=complex)
        self.vhat = np.array(solutionFile["vhat"], dtype=complex)
        self.what = np.array(solutionFile["what"], dtype=complex)
        if self.turb_model != 0:
            self.w0_u = solutionFile["w0_u"][:, :, :, 0]
            self.w0_v = solutionFile["w0_v"][:, :, :, 0]
            self.w0_w = solutionFile["w0_w"][:, :, :, 0]
        else:
            N1, N2, N3 = np.shape(self.uhat)
            self.w0_u = np.zeros((N1, N2, N3, 1), dtype=complex)
            self.w0_v = np.zeros((N1, N2, N3, 1), dtype=complex)
            self.w0_w = np.zeros((N1, N2, N3, 1), dtype=complex)

    def setFieldsFromFile(self, solutionFile):
        self.uhat[:, :, :] = solutionFile["uhat"]
        self.vhat[:, :, :] = solutionFile["vhat"]
        self.what[:, :, :] = solutionFile["what"]
        if self.turb_model != 0:
            self.w0_u = solutionFile["w0_u"][:, :, :, 0]
            self.w0_v = solutionFile["w0_v"][:, :, :, 0]
            self.w0_w = solutionFile["w0_w"][:, :, :, 0]


class gridFromFile:
    """``gridFromFile`` (``src/postProcessingFunctions.py:40-72``): wavenumbers from
    ``grid.npz``, integer 2x-grid wavenumbers, Delta = pi/kc and the sharp filter Gf."""

    def __init__(self, gridFile, solutionFile, infoFile, kc):
        N1, N2, n3c = np.shape(solutionFile["uhat"])
        N3 = (n3c - 1) * 2
        self.N1, self.N2, self.N3 = N1, N2, N3
        self.k1 = np.array(gridFile["k1"], dtype=float)
        self.k2 = np.array(gridFile["k2"], dtype=float)
        self.k3 = np.array(gridFile["k3"], dtype=float)
        k1f = np.fft.fftshift(np.linspace(-N1, N1 - 1, 2 * N1))
        k2f = np.fft.fftshift(np.linspace(-N2, N2 - 1, 2 * N2))
        k3f = np.linspace(0, N3, N3 + 1)
        self.k1f, self.k2f, self.k3f = np.meshgrid(k1f, k2f, k3f, indexing="ij")
        self.ksqr = self.k1 * self.k1 + self.k2 * self.k2 + self.k3 * self.k3 + 1.0e-50
        self.ksqr_i = 1.0 / self.ksqr
        self.ksqrf = self.k1f * self.k1f + self.k2f * self.k2f + self.k3f * self.k3f + 1.0e-50
        self.ksqrf_i = 1.0 / self.ksqrf
        self.kc = kc
        self.Delta = np.pi / self.kc
        self.Gf = np.zeros(np.shape(self.k1))
        self.Gf[0:kc, 0:kc, 0:kc] = 1
        self.Gf[0:kc, N2 - kc + 1:, 0:kc] = 1
        self.Gf[N1 - kc + 1:, 0:kc, 0:kc] = 1
        self.Gf[N1 - kc + 1:, N2 - kc + 1:, 0:kc] = 1


_TRANSFORM_SOLVERS = {}


def _transform_solver(dims):
    """A DNS context on the grid ``dims`` whose transforms the objects below use
    (created on first use, one per grid)."""
    from .solver import SpectralSolver

    if dims not in _TRANSFORM_SOLVERS:
        _TRANSFORM_SOLVERS[dims] = SpectralSolver(*dims, 0, nu=0.0)
    return _TRANSFORM_SOLVERS[dims]


class _GridTransform:
    """One of FFTFromFile's transform objects on the grid ``dims``, called as
    ``obj(array) -> array`` like a pyFFTW plan: the inverse (c2r) is normalised as
    pyFFTW's __call__ normalises FFTW_BACKWARD (numpy's irfftn), the forward (r2c) is
    unnormalised (numpy's rfftn).  numpy input gives a new numpy array (not a buffer the
    object owns); a torch device tensor gives a device tensor."""

    def __init__(self, dims, inverse: bool):
        self.dims = tuple(int(n) for n in dims)
        self.inverse = inverse
        self.nt = float(np.prod(self.dims))

    def __call__(self, x):
        import torch

        s = _transform_solver(self.dims)
        host = not isinstance(x, torch.Tensor)
        if self.inverse:
            t = torch.as_tensor(np.ascontiguousarray(x, dtype=complex) if host else x, dtype=torch.complex128)
            shape = self.dims[:2] + (self.dims[2] // 2 + 1,)
            if tuple(t.shape) != shape:
                raise ValueError(f"expected a complex array of shape {shape}, got {tuple(t.shape)}")
            out = s.physical(t.to(s.device).contiguous()[None])[0] / np.sqrt(self.nt)
        else:
            t = torch.as_tensor(np.ascontiguousarray(x, dtype=float) if host else x, dtype=torch.float64)
            if tuple(t.shape) != self.dims:
                raise ValueError(f"expected a real array of shape {self.dims}, got {tuple(t.shape)}")
            out = s.spectral(t.to(s.device).contiguous()[None])[0] * np.sqrt(self.nt)
        if host:
            torch.cuda.synchronize()
            return out.cpu().numpy()
        return out


class FFTFromFile:
    """``FFTFromFile`` (``src/postProcessingFunctions.py:67-108``): the transform objects of
    a saved field's 3/2-padded grid (``ifftT_obj`` c2r, ``fft_obj`` r2c), its 2x grid
    (``ifftT_obj2``, ``fft_obj2``) and its 3x grid (``ifftT_obj3``, ``fft_obj3``), with the
    reference's ``scale``.  The transforms run on the GPU (``psx_physical`` /
    ``psx_spectral`` of a context on each grid); ``nthreads`` is kept for the signature."""

    def __init__(self, gridFile, nthreads):
        N1, N2, n3 = np.shape(gridFile["k1"])
        N3 = (n3 - 1) * 2
        self.nthreads = nthreads
        self.scale = np.sqrt((3.0 / 2.0) ** 3 * np.sqrt(N1 * N2 * N3))
        grids = {"": (3 * N1 // 2, 3 * N2 // 2, 3 * N3 // 2), "2": (2 * N1, 2 * N2, 2 * N3),
                 "3": (3 * N1, 3 * N2, 3 * N3)}
        for suffix, dims in grids.items():
            setattr(self, "ifftT_obj" + suffix, _GridTransform(dims, inverse=True))
            setattr(self, "fft_obj" + suffix, _GridTransform(dims, inverse=False))
