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
no pickles).  ``FFTFromFile`` built host pyFFTW plans there; this package runs its
transforms inside the fused GPU pipeline, so it has no host transform objects to
give and raises.
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


class FFTFromFile:
    """``FFTFromFile`` (``src/postProcessingFunctions.py:76-108``) held host pyFFTW plans.
    The transforms here run inside the GPU pipeline (SpectralSolver); energies,
    spectra, filters and SGS terms of loaded fields come from a SpectralSolver on the
    field's grid, so this class has nothing to provide."""

    def __init__(self, gridFile, nthreads):
        raise NotImplementedError("FFTFromFile: host FFT plans are not part of this package; use "
                                  "pyspectral_amd.solver.SpectralSolver (stats, spectrum, filter, sgs_dns) "
                                  "on the loaded fields")
