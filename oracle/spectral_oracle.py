"""CPU oracle: a numpy restatement of PySpectral's serial pseudo-spectral hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``pyspectral_amd``) never does, and has no CPU fallback.

It restates, in Python 3 / numpy, the arithmetic of the reference's serial
solver ``PySpectral/src`` (written for Python 2 + pyFFTW, which cannot run
here).  Every function cites the reference file:line it follows.  The
conventions that bind parity:

* Fourier layout per field ``(N1, N2, N3//2+1)`` complex128, axis 0 = k1,
  axis 1 = k2 in FFT order, axis 2 = k3 >= 0 (``src/Classes.py:565-567``).
* ``uhat = rfftn(u) / sqrt(N1 N2 N3)`` (``run/TGV_Validate100/inputfile.py:42-44``).
* Padded inverse = pyFFTW's *normalised* c2r times
  ``scale = sqrt((3/2)^3 sqrt(Nt))`` (``src/Classes.py:593``); padded forward =
  unnormalised r2c.  numpy's ``irfftn``/``rfftn`` have the same normalisation.
* State ``Q`` interleaves the three components at stride 3 in every axis
  (``src/Classes.py:30-34``).
* Time stepping is the low-storage 4-stage scheme
  ``Q = Q0 + dt*c_i*RHS``, ``c = (1/4, 1/3, 1/2, 1)`` (``src/PySpec_3d.py:96-105``).

Parity status of each function is listed in ``PARITY`` at the bottom of this
file, together with the test that pins it.
"""
from __future__ import annotations

import math
import os

import numpy as np

# --------------------------------------------------------------------------
# FFT backend (stands in for the reference's pyFFTW plans, src/Classes.py:590-625)
# --------------------------------------------------------------------------


class FFTBackend:
    """Normalised c2r inverse and unnormalised r2c forward over all three axes.

    pyFFTW's ``FFTW.__call__`` normalises BACKWARD transforms by 1/size, as
    ``numpy.fft.irfftn`` does, and leaves FORWARD transforms unnormalised, as
    ``numpy.fft.rfftn`` does (``src/Classes.py:596-607``).  ``workers > 1`` uses
    ``scipy.fft`` with that many threads (the CPU-baseline configuration).
    """

    def __init__(self, workers: int = 1):
        self.workers = int(workers)
        self.name = "numpy.fft" if self.workers <= 1 else f"scipy.fft(workers={self.workers})"
        if self.workers > 1:
            import scipy.fft as sfft

            self._sfft = sfft
        else:
            self._sfft = None

    def c2r(self, X: np.ndarray, shape) -> np.ndarray:
        if self._sfft is not None:
            return self._sfft.irfftn(X, s=shape, workers=self.workers)
        return np.fft.irfftn(X, s=shape, axes=(0, 1, 2))

    def r2c(self, x: np.ndarray) -> np.ndarray:
        if self._sfft is not None:
            return self._sfft.rfftn(x, workers=self.workers)
        return np.fft.rfftn(x)


# --------------------------------------------------------------------------
# Padding / truncation (src/padding.py)
# --------------------------------------------------------------------------


def _axis_map_32(n: int, m: int):
    """Kept FFT-ordered indices of an n-point axis and their slots in an m-point axis
    for the 3/2 rule: index n/2 (the -n/2 "oddball" mode) is dropped
    (``src/padding.py:8-12,20-24``)."""
    h = n // 2
    src = np.concatenate([np.arange(0, h), np.arange(h + 1, n)])
    dst = np.concatenate([np.arange(0, h), np.arange(m - h + 1, m)])
    return src, dst


def _axis_map_2x(n: int):
    """Same for the 2x rule, which keeps the -n/2 mode (``src/padding.py:34-37,45-48``)."""
    h = n // 2
    src = np.arange(0, n)
    dst = np.concatenate([np.arange(0, h), np.arange(2 * n - h, 2 * n)])
    return src, dst


def pad(uhat: np.ndarray) -> np.ndarray:
    """3/2 zero-pad ``(N1,N2,N3c) -> (3N1/2, 3N2/2, N3c+(N3c-1)//2)``; drops the
    Nyquist index on every axis (``src/padding.py:16-25``)."""
    N1, N2, N3c = uhat.shape
    M1, M2, M3c = (3 * N1) // 2, (3 * N2) // 2, N3c + (N3c - 1) // 2
    s1, d1 = _axis_map_32(N1, M1)
    s2, d2 = _axis_map_32(N2, M2)
    kz = np.arange(0, N3c - 1)
    out = np.zeros((M1, M2, M3c), dtype=complex)
    out[np.ix_(d1, d2, kz)] = uhat[np.ix_(s1, s2, kz)]
    return out


def unpad(uhat_pad: np.ndarray) -> np.ndarray:
    """Inverse of :func:`pad` (``src/padding.py:2-13``); Nyquist entries come back zero."""
    M1, M2, M3c = uhat_pad.shape
    N1, N2 = int(M1 * 2.0 / 3.0), int(M2 * 2.0 / 3.0)
    N3c = int(M3c * 2.0 / 3.0 + 1)
    s1, d1 = _axis_map_32(N1, M1)
    s2, d2 = _axis_map_32(N2, M2)
    kz = np.arange(0, N3c - 1)
    out = np.zeros((N1, N2, N3c), dtype=complex)
    out[np.ix_(s1, s2, kz)] = uhat_pad[np.ix_(d1, d2, kz)]
    return out


def pad_2x(uhat: np.ndarray) -> np.ndarray:
    """2x zero-pad ``(N1,N2,N3c) -> (2N1, 2N2, 2N3c-1)``; keeps the -N/2 rows, drops
    kz = N3/2 (``src/padding.py:41-49``)."""
    N1, N2, N3c = uhat.shape
    s1, d1 = _axis_map_2x(N1)
    s2, d2 = _axis_map_2x(N2)
    kz = np.arange(0, N3c - 1)
    out = np.zeros((2 * N1, 2 * N2, 2 * N3c - 1), dtype=complex)
    out[np.ix_(d1, d2, kz)] = uhat[np.ix_(s1, s2, kz)]
    return out


def unpad_2x(uhat_pad: np.ndarray) -> np.ndarray:
    """Inverse of :func:`pad_2x` (``src/padding.py:28-38``)."""
    P1, P2, P3 = uhat_pad.shape
    N1, N2, N3c = P1 // 2, P2 // 2, (P3 + 1) // 2
    s1, d1 = _axis_map_2x(N1)
    s2, d2 = _axis_map_2x(N2)
    kz = np.arange(0, N3c - 1)
    out = np.zeros((N1, N2, N3c), dtype=complex)
    out[np.ix_(s1, s2, kz)] = uhat_pad[np.ix_(d1, d2, kz)]
    return out


def _p_cube_mask(shape, k1_lo: int, k2_lo: int, kz_hi: int) -> np.ndarray:
    """Boolean mask of the resolved box ``0..k1_lo-1`` and ``-(k1_lo-1)..-1`` on the
    first axis (likewise with ``k2_lo`` on the second), ``0..kz_hi-1`` on the third."""
    P1, P2, P3 = shape
    if min(k1_lo, k2_lo) < 2:
        raise ValueError("mode split needs a cutoff of at least 2")

    def ax(n, lo):
        m = np.zeros(n, dtype=bool)
        m[:lo] = True
        m[n - lo + 1:] = True
        return m

    mz = np.zeros(P3, dtype=bool)
    mz[:kz_hi] = True
    return ax(P1, k1_lo)[:, None, None] & ax(P2, k2_lo)[None, :, None] & mz[None, None, :]


def separate_modes(uhat_pad: np.ndarray):
    """p/q split of a 2x-grid field: p keeps |k1| < N1/2, |k2| < N2/2 (the -N/2 modes
    go to q) and k3 < N3/2, each axis at its own N (``src/padding.py:65-79``)."""
    P1, P2, P3 = uhat_pad.shape
    N3c = (P3 + 1) // 2
    m = _p_cube_mask(uhat_pad.shape, P1 // 4, P2 // 4, P3 - N3c)
    up = np.where(m, uhat_pad, 0)
    return up, uhat_pad - up


def separate_modes_testfilter(uhat_pad: np.ndarray, kf: int):
    """p/q split at the test-filter index ``kf`` (``src/padding.py:52-64``)."""
    m = _p_cube_mask(uhat_pad.shape, kf, kf, kf)
    up = np.where(m, uhat_pad, 0)
    return up, uhat_pad - up


# --------------------------------------------------------------------------
# Grid (src/Classes.py:548-588)
# --------------------------------------------------------------------------


class Grid:
    """Wavenumber arrays and spacings, as ``gridclass`` builds them.

    ``dx, dy, dz`` are the case's grid spacings (the reference reads them off the
    case's ``x, y, z`` arrays, ``src/Classes.py:562-564``); ``Delta`` is their
    geometric mean (``src/Classes.py:582-583``).
    """

    def __init__(self, N1, N2, N3, L1=2 * np.pi, L2=2 * np.pi, L3=2 * np.pi,
                 dx=None, dy=None, dz=None, kc=None):
        self.N1, self.N2, self.N3 = int(N1), int(N2), int(N3)
        self.L1, self.L2, self.L3 = float(L1), float(L2), float(L3)
        self.dx = float(L1 / N1 if dx is None else dx)
        self.dy = float(L2 / N2 if dy is None else dy)
        self.dz = float(L3 / N3 if dz is None else dz)
        k1 = np.fft.fftshift(np.linspace(-N1 // 2, N1 // 2 - 1, N1)) * 2.0 * np.pi / L1
        k2 = np.fft.fftshift(np.linspace(-N2 // 2, N2 // 2 - 1, N2)) * 2.0 * np.pi / L2
        k3 = np.linspace(0, N3 // 2, N3 // 2 + 1) * 2.0 * np.pi / L3
        k1f = np.fft.fftshift(np.linspace(-N1, N1 - 1, 2 * N1)) * 2.0 * np.pi / L1
        k2f = np.fft.fftshift(np.linspace(-N2, N2 - 1, 2 * N2)) * 2.0 * np.pi / L2
        k3f = np.linspace(0, N3, N3 + 1) * 2.0 * np.pi / L3
        self.k1, self.k2, self.k3 = np.meshgrid(k1, k2, k3, indexing="ij")
        self.k1f, self.k2f, self.k3f = np.meshgrid(k1f, k2f, k3f, indexing="ij")
        self.ksqr = self.k1 * self.k1 + self.k2 * self.k2 + self.k3 * self.k3 + 1.0e-50
        self.ksqr_i = 1.0 / self.ksqr
        self.ksqrf = self.k1f * self.k1f + self.k2f * self.k2f + self.k3f * self.k3f + 1.0e-50
        self.ksqrf_i = 1.0 / self.ksqrf
        self.kc = int(N1 // 2 if kc is None else kc)
        self.Delta = (self.dx * self.dy * self.dz) ** (1.0 / 3.0)
        # sharp cutoff filter at kc (src/Classes.py:584-588)
        kc_ = self.kc
        Gf = np.zeros(self.k1.shape)
        Gf[0:kc_, 0:kc_, 0:kc_] = 1
        Gf[0:kc_, self.N2 - kc_ + 1:, 0:kc_] = 1
        Gf[self.N1 - kc_ + 1:, 0:kc_, 0:kc_] = 1
        Gf[self.N1 - kc_ + 1:, self.N2 - kc_ + 1:, 0:kc_] = 1
        self.Gf = Gf

    @property
    def Nt(self):
        return self.N1 * self.N2 * self.N3


# --------------------------------------------------------------------------
# State (subset of `variables`, src/Classes.py:7-229)
# --------------------------------------------------------------------------

MODEL_NAMES = {
    0: "dns",
    1: "smagorinsky",
    2: "tmodel",
    "Static tau-model": "static_tau",
    "Dynamic tau-model": "dynamic_tau",
    7: "dsmag",
}


class State:
    """Solution state with the reference's interleaved ``Q`` and its ``Q2U/U2Q``."""

    def __init__(self, grid: Grid, uhat, vhat, what, *, nu, turb_model=0, t=0.0, dt=0.01,
                 cfl=None, Ct=-10, dt0=-10, dt1=-10, rotate=0, Om=(0.0, 0.0, 0.0), force=0.0):
        self.grid = grid
        # linear forcing f = force * u (not in the reference; the forced-turbulence cases)
        self.force = float(force)
        self.turb_model = turb_model
        self.model = MODEL_NAMES[turb_model]
        self.nu = float(nu)
        self.t = float(t)
        self.dt = float(dt)
        # src/PySpec_3d.py:35-38: cfl defaults to -dt (fixed step)
        self.cfl = float(-dt if cfl is None else cfl)
        self.rotate = int(rotate)
        self.Om1, self.Om2, self.Om3 = (float(o) for o in Om)
        self.Ct = 0.1 if Ct == -10 else float(Ct)          # src/Classes.py:103-108
        self.tau_a, self.tau_b = float(dt0), float(dt1)      # src/Classes.py:194-195
        if self.model == "dynamic_tau":
            self.kf = int(grid.kc * 2.0 / 3.0)               # src/Classes.py:131
        elif self.model == "static_tau":
            self.kf = int(grid.kc / 2.0)                     # src/Classes.py:197
        else:
            self.kf = None
        N1, N2, N3c = grid.N1, grid.N2, grid.N3 // 2 + 1
        self.uhat = np.array(uhat, dtype=complex).copy()
        self.vhat = np.array(vhat, dtype=complex).copy()
        self.what = np.array(what, dtype=complex).copy()
        assert self.uhat.shape == (N1, N2, N3c)
        self.Q = np.zeros((3 * N1, 3 * N2, 3 * N3c), dtype=complex)
        self.U2Q()
        self.w0 = None
        self.iteration = 0

    def U2Q(self):
        self.Q[0::3, 0::3, 0::3] = self.uhat
        self.Q[1::3, 1::3, 1::3] = self.vhat
        self.Q[2::3, 2::3, 2::3] = self.what

    def Q2U(self):
        self.uhat = self.Q[0::3, 0::3, 0::3].copy()
        self.vhat = self.Q[1::3, 1::3, 1::3].copy()
        self.what = self.Q[2::3, 2::3, 2::3].copy()

    def set_rhs(self, ru, rv, rw):
        self.Q[0::3, 0::3, 0::3] = ru
        self.Q[1::3, 1::3, 1::3] = rv
        self.Q[2::3, 2::3, 2::3] = rw

    def compact(self) -> np.ndarray:
        """State as the compact ``(3, N1, N2, N3c)`` array."""
        return np.stack([self.Q[c::3, c::3, c::3] for c in range(3)])

    def set_compact(self, q: np.ndarray):
        for c in range(3):
            self.Q[c::3, c::3, c::3] = q[c]


# --------------------------------------------------------------------------
# Right-hand sides (src/RHSfunctions.py)
# --------------------------------------------------------------------------


def _coriolis_p(st, g, u, v, w, k1, k2, k3, ksqr_i):
    return 2.0 * ksqr_i * 1j * (k1 * (v * st.Om3 - w * st.Om2)
                                + k2 * (w * st.Om1 - u * st.Om3)
                                + k3 * (u * st.Om2 - v * st.Om1))


def _coriolis_add(st, ru, rv, rw):
    u, v, w = st.uhat, st.vhat, st.what
    return (ru + 2.0 * (v * st.Om3 - w * st.Om2),
            rv + 2.0 * (w * st.Om1 - u * st.Om3),
            rw + 2.0 * (u * st.Om2 - v * st.Om1))


def _products_32(st, fft, uhat, vhat, what):
    """Three padded inverses and the six dealiased products u_i u_j
    (``src/RHSfunctions.py:14-23``)."""
    g = st.grid
    scale = np.sqrt((3.0 / 2.0) ** 3 * np.sqrt(g.Nt))
    M = ((3 * g.N1) // 2, (3 * g.N2) // 2, (3 * g.N3) // 2)
    ur = fft.c2r(pad(uhat), M) * scale
    vr = fft.c2r(pad(vhat), M) * scale
    wr = fft.c2r(pad(what), M) * scale
    uu = unpad(fft.r2c(ur * ur))
    vv = unpad(fft.r2c(vr * vr))
    ww = unpad(fft.r2c(wr * wr))
    uv = unpad(fft.r2c(ur * vr))
    uw = unpad(fft.r2c(ur * wr))
    vw = unpad(fft.r2c(vr * wr))
    return uu, vv, ww, uv, uw, vw


def rhs_nosgs(st: State, fft: FFTBackend) -> None:
    """``computeRHS_NOSGS`` (``src/RHSfunctions.py:3-50``); also the arithmetic of the
    scipy.weave kernel ``src/RHSfunctionsWEAVE.py:44-66``.  Writes the RHS into
    ``st.Q`` and leaves the Nyquist-truncated input in ``st.uhat/vhat/what``."""
    g = st.grid
    st.Q2U()
    st.uhat = unpad(pad(st.uhat))
    st.vhat = unpad(pad(st.vhat))
    st.what = unpad(pad(st.what))
    uu, vv, ww, uv, uw, vw = _products_32(st, fft, st.uhat, st.vhat, st.what)
    k1, k2, k3 = g.k1, g.k2, g.k3
    phat = g.ksqr_i * (-k1 * k1 * uu - k2 * k2 * vv - k3 * k3 * ww - 2.0 * k1 * k2 * uv
                       - 2.0 * k1 * k3 * uw - 2.0 * k2 * k3 * vw)
    if st.rotate == 1:
        phat = phat - _coriolis_p(st, g, st.uhat, st.vhat, st.what, k1, k2, k3, g.ksqr_i)
    ru = -1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw - 1j * k1 * phat - st.nu * g.ksqr * st.uhat
    rv = -1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw - 1j * k2 * phat - st.nu * g.ksqr * st.vhat
    rw = -1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww - 1j * k3 * phat - st.nu * g.ksqr * st.what
    if st.rotate == 1:
        ru, rv, rw = _coriolis_add(st, ru, rv, rw)
    if st.force != 0.0:
        ru, rv, rw = ru + st.force * st.uhat, rv + st.force * st.vhat, rw + st.force * st.what
    st.set_rhs(ru, rv, rw)


SMAG_CS = 0.16  # hard-coded in src/RHSfunctions.py:197


def smag_stress_hat(st: State, fft: FFTBackend, uhat, vhat, what, return_real=False, cs2=None):
    """Smagorinsky stress spectra tau_ij = unpad(FFT(-2 nu_t S_ij)), order
    (11, 22, 33, 12, 13, 23) (``src/RHSfunctions.py:174-212``).  With
    ``return_real`` also the padded-grid strain fields and nu_t (in the reference's
    scaled units ``(3/2)^(-3/2) Nt^(-1/4)`` times physical).  ``cs2`` replaces the
    hard-coded Cs^2 = 0.16^2 (the dynamic model's coefficient)."""
    g = st.grid
    k1, k2, k3 = g.k1, g.k2, g.k3
    scale = np.sqrt((3.0 / 2.0) ** 3 * np.sqrt(g.Nt))
    M = ((3 * g.N1) // 2, (3 * g.N2) // 2, (3 * g.N3) // 2)
    S11 = 1j * k1 * uhat
    S22 = 1j * k2 * vhat
    S33 = 1j * k3 * what
    S12 = 0.5 * (1j * k2 * uhat + 1j * k1 * vhat)
    S13 = 0.5 * (1j * k3 * uhat + 1j * k1 * what)
    S23 = 0.5 * (1j * k3 * vhat + 1j * k2 * what)
    Sr = [fft.c2r(pad(S) * scale, M) for S in (S11, S22, S33, S12, S13, S23)]
    s11, s22, s33, s12, s13, s23 = Sr
    Smag = np.sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * s12 * s12
                          + 2.0 * s13 * s13 + 2.0 * s23 * s23))
    cs2 = SMAG_CS * SMAG_CS if cs2 is None else cs2
    nut = cs2 * g.Delta * g.Delta * np.abs(Smag)
    tau = [unpad(fft.r2c(-2.0 * nut * s)) for s in Sr]
    if return_real:
        return tau, Sr, nut
    return tau


def sgs_divergence(g: Grid, tau):
    """w0 from stress spectra, with the reference's projection term
    ``+1j*k_i*(1j*ksqr_i*k_m k_j tau_mj)`` kept as written
    (``src/RHSfunctions.py:215-222``; same form in ``src/Classes.py:911-918``)."""
    k1, k2, k3 = g.k1, g.k2, g.k3
    t11, t22, t33, t12, t13, t23 = tau
    tp = 1j * g.ksqr_i * (k1 * k1 * t11 + k2 * k2 * t22 + k3 * k3 * t33 + 2.0 * k1 * k2 * t12
                          + 2.0 * k1 * k3 * t13 + 2.0 * k2 * k3 * t23)
    wu = -1j * k1 * t11 - 1j * k2 * t12 - 1j * k3 * t13 + 1j * k1 * tp
    wv = -1j * k1 * t12 - 1j * k2 * t22 - 1j * k3 * t23 + 1j * k2 * tp
    ww = -1j * k1 * t13 - 1j * k2 * t23 - 1j * k3 * t33 + 1j * k3 * tp
    return wu, wv, ww


def rhs_smag(st: State, fft: FFTBackend) -> None:
    """``computeRHS_SMAG`` (``src/RHSfunctions.py:146-237``).  No Nyquist truncation of
    ``uhat`` (``:153-155``), so the Nyquist entries of Q decay by -nu k^2 uhat only."""
    g = st.grid
    st.Q2U()
    uu, vv, ww, uv, uw, vw = _products_32(st, fft, st.uhat, st.vhat, st.what)
    k1, k2, k3 = g.k1, g.k2, g.k3
    ki = g.ksqr_i
    phat = (-k1 * k1 * ki * uu - k2 * k2 * ki * vv - k3 * k3 * ki * ww - 2.0 * k1 * k2 * ki * uv
            - 2.0 * k1 * k3 * ki * uw - 2.0 * k2 * k3 * ki * vw)
    if st.rotate == 1:
        phat = phat - _coriolis_p(st, g, st.uhat, st.vhat, st.what, k1, k2, k3, ki)
    tau = smag_stress_hat(st, fft, st.uhat, st.vhat, st.what)
    wu, wv, ww_ = sgs_divergence(g, tau)
    st.w0 = (wu, wv, ww_)
    ru = -1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw - 1j * k1 * phat - st.nu * g.ksqr * st.uhat + wu
    rv = -1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw - 1j * k2 * phat - st.nu * g.ksqr * st.vhat + wv
    rw = -1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww - 1j * k3 * phat - st.nu * g.ksqr * st.what + ww_
    if st.rotate == 1:
        ru, rv, rw = _coriolis_add(st, ru, rv, rw)
    if st.force != 0.0:
        ru, rv, rw = ru + st.force * st.uhat, rv + st.force * st.vhat, rw + st.force * st.what
    st.set_rhs(ru, rv, rw)


DSMAG_ALPHA = 2.0  # src/RHSfunctions.py:1665


def dsmag_coefficient(st: State, fft: FFTBackend, uhat, vhat, what):
    """The dynamic Smagorinsky coefficient of ``computeRHS_DSMAG``
    (``src/RHSfunctions.py:1663-1820``): test filter myFilter at kL = kc/alpha (the
    reference slices with the float kc/2, which truncates to int(kc/2)); Leonard
    stress L_ij = myFilter(u_i u_j) - uF_i uF_j; M_ij = -2 Delta^2 (alpha^2 FFT(|S^F| S^F_ij)
    - myFilter(FFT(|S| S_ij))); Cs^2 = mean(M_ij L_ij) / mean(M_ij M_ij) over the six
    stored components on the padded grid.  Returns (Cs^2, L, M, SS) with SS the
    spectra of |S| S_ij (unfiltered)."""
    g = st.grid
    kL = int(g.kc / DSMAG_ALPHA)
    flt = Grid(g.N1, g.N2, g.N3, kc=kL).Gf       # myFilter(., kL): the same index mask
    u = (uhat, vhat, what)
    uF = tuple(flt * a for a in u)
    Nu = _products_32(st, fft, *u)
    NF = _products_32(st, fft, *uF)
    L = [flt * a - b for a, b in zip(Nu, NF)]
    _, Sr, _ = smag_stress_hat(st, fft, *u, return_real=True, cs2=1.0)
    _, SrF, _ = smag_stress_hat(st, fft, *uF, return_real=True, cs2=1.0)

    def mag(S):
        s11, s22, s33, s12, s13, s23 = S
        return np.sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * s12 * s12 + 2.0 * s13 * s13
                              + 2.0 * s23 * s23))

    Sm, SmF = mag(Sr), mag(SrF)
    SS = [unpad(fft.r2c(Sm * s)) for s in Sr]
    SSF = [unpad(fft.r2c(SmF * s)) for s in SrF]
    d2 = g.Delta ** 2
    M = [-2.0 * d2 * (DSMAG_ALPHA ** 2 * a - flt * b) for a, b in zip(SSF, SS)]
    scale = np.sqrt((3.0 / 2.0) ** 3 * np.sqrt(g.Nt))
    Mp = ((3 * g.N1) // 2, (3 * g.N2) // 2, (3 * g.N3) // 2)
    Mr = [fft.c2r(pad(a), Mp) * scale for a in M]
    Lr = [fft.c2r(pad(a), Mp) * scale for a in L]
    num = sum(m * l for m, l in zip(Mr, Lr))
    den = sum(m * m for m in Mr)
    return float(np.mean(num) / np.mean(den)), L, M, SS


def rhs_dsmag(st: State, fft: FFTBackend) -> None:
    """``computeRHS_DSMAG`` (``src/RHSfunctions.py:1663-1862``): the Smagorinsky RHS
    (no Nyquist truncation) with nu_t = Cs^2 Delta^2 |S| and the dynamic Cs^2."""
    g = st.grid
    st.Q2U()
    cs2, _, _, _ = dsmag_coefficient(st, fft, st.uhat, st.vhat, st.what)
    st.cs2 = cs2
    uu, vv, ww, uv, uw, vw = _products_32(st, fft, st.uhat, st.vhat, st.what)
    k1, k2, k3 = g.k1, g.k2, g.k3
    ki = g.ksqr_i
    phat = (-k1 * k1 * ki * uu - k2 * k2 * ki * vv - k3 * k3 * ki * ww - 2.0 * k1 * k2 * ki * uv
            - 2.0 * k1 * k3 * ki * uw - 2.0 * k2 * k3 * ki * vw)
    if st.rotate == 1:
        phat = phat - _coriolis_p(st, g, st.uhat, st.vhat, st.what, k1, k2, k3, ki)
    tau = smag_stress_hat(st, fft, st.uhat, st.vhat, st.what, cs2=cs2)
    wu, wv, ww_ = sgs_divergence(g, tau)
    st.w0 = (wu, wv, ww_)
    ru = -1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw - 1j * k1 * phat - st.nu * g.ksqr * st.uhat + wu
    rv = -1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw - 1j * k2 * phat - st.nu * g.ksqr * st.vhat + wv
    rw = -1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww - 1j * k3 * phat - st.nu * g.ksqr * st.what + ww_
    if st.rotate == 1:
        ru, rv, rw = _coriolis_add(st, ru, rv, rw)
    if st.force != 0.0:
        ru, rv, rw = ru + st.force * st.uhat, rv + st.force * st.vhat, rw + st.force * st.what
    st.set_rhs(ru, rv, rw)


def _pl_2x(st, fft, uhat, vhat, what):
    """Full 2x-grid projected RHS PL_i and the real fields (``src/RHSfunctions.py:269-315``)."""
    g = st.grid
    scale = np.sqrt(2.0 ** 3 * np.sqrt(g.Nt))
    P = (2 * g.N1, 2 * g.N2, 2 * g.N3)
    up, vp, wp = pad_2x(uhat), pad_2x(vhat), pad_2x(what)
    ur = fft.c2r(up * scale, P)
    vr = fft.c2r(vp * scale, P)
    wr = fft.c2r(wp * scale, P)
    uu = fft.r2c(ur * ur)
    vv = fft.r2c(vr * vr)
    ww = fft.r2c(wr * wr)
    uv = fft.r2c(ur * vr)
    uw = fft.r2c(ur * wr)
    vw = fft.r2c(vr * wr)
    k1, k2, k3, ki = g.k1f, g.k2f, g.k3f, g.ksqrf_i
    phat = (-k1 * k1 * ki * uu - k2 * k2 * ki * vv - k3 * k3 * ki * ww - 2.0 * k1 * k2 * ki * uv
            - 2.0 * k1 * k3 * ki * uw - 2.0 * k2 * k3 * ki * vw)
    if st.rotate == 1:
        phat = phat - _coriolis_p(st, g, up, vp, wp, k1, k2, k3, ki)
    PLu = -1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw - 1j * k1 * phat - st.nu * g.ksqrf * up
    PLv = -1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw - 1j * k2 * phat - st.nu * g.ksqrf * vp
    PLw = -1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww - 1j * k3 * phat - st.nu * g.ksqrf * wp
    prods = (uu, vv, ww, uv, uw, vw)
    return (PLu, PLv, PLw), (ur, vr, wr), prods, phat, (up, vp, wp)


def _plql(st, fft, PL, real, qsplit):
    """PLQL_i from the q-part of PL and the real velocity on the 2x grid
    (``src/RHSfunctions.py:317-348`` / ``:443-478``)."""
    g = st.grid
    scale = np.sqrt(2.0 ** 3 * np.sqrt(g.Nt))
    P = (2 * g.N1, 2 * g.N2, 2 * g.N3)
    ur, vr, wr = real
    qs = [qsplit(x)[1] for x in PL]
    qr = [fft.c2r(q * scale, P) for q in qs]
    Luq, Lvq, Lwq = qr
    up_u = unpad_2x(fft.r2c(ur * Luq))
    vp_u = unpad_2x(fft.r2c(vr * Luq))
    wp_u = unpad_2x(fft.r2c(wr * Luq))
    up_v = unpad_2x(fft.r2c(ur * Lvq))
    vp_v = unpad_2x(fft.r2c(vr * Lvq))
    wp_v = unpad_2x(fft.r2c(wr * Lvq))
    up_w = unpad_2x(fft.r2c(ur * Lwq))
    vp_w = unpad_2x(fft.r2c(vr * Lwq))
    wp_w = unpad_2x(fft.r2c(wr * Lwq))
    k1, k2, k3 = g.k1, g.k2, g.k3
    pterm = 2.0 * g.ksqr_i * (k1 * k1 * up_u + k2 * k2 * vp_v + k3 * k3 * wp_w
                              + k1 * k2 * (up_v + vp_u) + k1 * k3 * (up_w + wp_u)
                              + k2 * k3 * (vp_w + wp_v))
    Qu = (-1j * k1 * up_u - 1j * k2 * vp_u - 1j * k3 * wp_u
          - 1j * k1 * up_u - 1j * k2 * up_v - 1j * k3 * up_w + 1j * k1 * pterm)
    Qv = (-1j * k1 * up_v - 1j * k2 * vp_v - 1j * k3 * wp_v
          - 1j * k1 * vp_u - 1j * k2 * vp_v - 1j * k3 * vp_w + 1j * k2 * pterm)
    Qw = (-1j * k1 * up_w - 1j * k2 * vp_w - 1j * k3 * wp_w
          - 1j * k1 * wp_u - 1j * k2 * wp_v - 1j * k3 * wp_w + 1j * k3 * pterm)
    return Qu, Qv, Qw


def _tau_family_finish(st, PL, w0):
    wu, wv, ww = w0
    st.w0 = (wu, wv, ww)
    ru = unpad_2x(PL[0]) + wu
    rv = unpad_2x(PL[1]) + wv
    rw = unpad_2x(PL[2]) + ww
    if st.rotate == 1:
        ru, rv, rw = _coriolis_add(st, ru, rv, rw)
    if st.force != 0.0:
        ru, rv, rw = ru + st.force * st.uhat, rv + st.force * st.vhat, rw + st.force * st.what
    st.set_rhs(ru, rv, rw)


def rhs_tmodel(st: State, fft: FFTBackend) -> None:
    """``computeRHS_tmodel`` (``src/RHSfunctions.py:240-363``): 2x padding,
    ``w0 = Ct * t * PLQL`` with ``t`` the step-start time."""
    st.Q2U()
    st.uhat = unpad(pad(st.uhat))
    st.vhat = unpad(pad(st.vhat))
    st.what = unpad(pad(st.what))
    PL, real, _, _, _ = _pl_2x(st, fft, st.uhat, st.vhat, st.what)
    Qu, Qv, Qw = _plql(st, fft, PL, real, separate_modes)
    c = st.Ct * st.t
    _tau_family_finish(st, PL, (c * Qu, c * Qv, c * Qw))


def rhs_staumodel(st: State, fft: FFTBackend) -> None:
    """``computeRHS_staumodel`` (``src/RHSfunctions.py:366-497``):
    ``w0 = max(0, tau_a + tau_b t) * PLQL``."""
    st.Q2U()
    st.uhat = unpad(pad(st.uhat))
    st.vhat = unpad(pad(st.vhat))
    st.what = unpad(pad(st.what))
    PL, real, _, _, _ = _pl_2x(st, fft, st.uhat, st.vhat, st.what)
    Qu, Qv, Qw = _plql(st, fft, PL, real, separate_modes)
    st.tau = st.tau_a + st.tau_b * st.t
    tau = max(0.0, st.tau)
    _tau_family_finish(st, PL, (tau * Qu, tau * Qv, tau * Qw))


def test_filter(g: Grid, kf: int, uhat):
    """Sharp test filter |k_i| < kf on the N grid (``src/Classes.py:132-150``)."""
    f1 = (np.abs(g.k1[:, 0, 0]) < kf).astype(float)
    f2 = (np.abs(g.k2[0, :, 0]) < kf).astype(float)
    f3 = (np.abs(g.k3[0, 0, :]) < kf).astype(float)
    return f1[:, None, None] * f2[None, :, None] * f3[None, None, :] * uhat


def test_filter_2x(g: Grid, kf: int, uhat_pad):
    """Same on the 2x grid (``src/Classes.py:152-170``)."""
    f1 = (np.abs(g.k1f[:, 0, 0]) < kf).astype(float)
    f2 = (np.abs(g.k2f[0, :, 0]) < kf).astype(float)
    f3 = (np.abs(g.k3f[0, 0, :]) < kf).astype(float)
    return f1[:, None, None] * f2[None, :, None] * f3[None, None, :] * uhat_pad


def box_transfer(g: Grid, delta: float) -> np.ndarray:
    """Transfer function of the top-hat (box) filter of width ``delta`` on every axis,
    G(k) = prod_i sin(k_i delta/2) / (k_i delta/2).  Not in the reference (its filters
    are sharp cutoffs); the filtered-DNS training data use it."""
    s = lambda k: np.sinc(k * delta / (2.0 * np.pi))  # np.sinc(x) = sin(pi x)/(pi x)
    return s(g.k1) * s(g.k2) * s(g.k3)


def filter_transfer(g: Grid, kind: str = "sharp", delta: float | None = None) -> np.ndarray:
    """``grid.Gf`` (sharp cutoff at kc, ``src/Classes.py:584-588``; the same mask as
    ``myFilter``, ``src/Classes.py:1271-1279``) or the box filter."""
    if kind == "sharp":
        return g.Gf
    if kind == "box":
        return box_transfer(g, float(delta))
    raise ValueError(kind)


def sgs_dns(st: State, fft: FFTBackend, kind: str = "sharp", delta: float | None = None):
    """``utilitiesClass.computeSGS_DNS`` (``src/Classes.py:861-929``) on the state's
    ``uhat, vhat, what``: with ubar = G uhat, tau_hat_ij = G FFT(u_i u_j) - FFT(ubar_i ubar_j)
    (3/2-rule products, ``:886-909``), and w0 by the same divergence/projection form as
    the Smagorinsky closure (``:911-926``).  Returns (w0 (3 arrays), tau_hat (6 arrays,
    order 11 22 33 12 13 23))."""
    g = st.grid
    G = filter_transfer(g, kind, delta)
    u = (st.uhat, st.vhat, st.what)
    N = _products_32(st, fft, *u)
    Nf = _products_32(st, fft, *(G * a for a in u))
    tau = [G * a - b for a, b in zip(N, Nf)]
    return list(sgs_divergence(g, tau)), tau


# --------------------------------------------------------------------------
# Grid-space fields and the learned closure (no counterpart in the reference)
# --------------------------------------------------------------------------


def physical(g: Grid, fhat) -> np.ndarray:
    """Grid values irfftn(fhat) sqrt(Nt) of a spectral field (the inverse of
    uhat = rfftn(u)/sqrt(Nt), ``run/TGV_Validate100/inputfile.py:42-44``)."""
    return np.fft.irfftn(fhat, s=(g.N1, g.N2, g.N3), axes=(0, 1, 2)) * np.sqrt(g.Nt)


def gradients(g: Grid, uhat, vhat, what) -> np.ndarray:
    """du_i/dx_j at the grid points, (N1 N2 N3, 9) with column 3i + j; spectral
    derivatives i k_j on the signed FFT frequencies."""
    ks = (g.k1, g.k2, g.k3)
    cols = [physical(g, 1j * ks[j] * f).ravel() for f in (uhat, vhat, what) for j in range(3)]
    return np.stack(cols, axis=1)


def restrict(fhat, dst: Grid) -> np.ndarray:
    """Modes |k1| < n1/2, |k2| < n2/2, 0 <= k3 < n3/2 of a finer grid's spectrum on the
    grid ``dst`` (its Nyquist entries zero), times sqrt(nt/Nt) so grid values are kept."""
    S1, S2, Sc = fhat.shape
    S3 = 2 * (Sc - 1)
    n1, n2, n3 = dst.N1, dst.N2, dst.N3
    out = np.zeros((n1, n2, n3 // 2 + 1), dtype=complex)
    i1 = np.r_[0:n1 // 2, S1 - n1 // 2 + 1:S1]
    o1 = np.r_[0:n1 // 2, n1 // 2 + 1:n1]
    i2 = np.r_[0:n2 // 2, S2 - n2 // 2 + 1:S2]
    o2 = np.r_[0:n2 // 2, n2 // 2 + 1:n2]
    out[np.ix_(o1, o2, np.arange(n3 // 2))] = fhat[np.ix_(i1, i2, np.arange(n3 // 2))]
    return out * np.sqrt((n1 * n2 * n3) / (S1 * S2 * S3))


def padded_gradients(st: State, fft: FFTBackend, uhat, vhat, what) -> np.ndarray:
    """du_i/dx_j at the points of the 3/2-padded grid, (M1 M2 M3, 9), column 3i + j,
    physical units: the padded c2r of ``src/RHSfunctions.py:174-193`` without its
    scale, times M1 M2 M3 / sqrt(Nt)."""
    g = st.grid
    M = ((3 * g.N1) // 2, (3 * g.N2) // 2, (3 * g.N3) // 2)
    f = np.prod(M) / np.sqrt(g.Nt)
    ks = (g.k1, g.k2, g.k3)
    cols = [(fft.c2r(pad(1j * ks[j] * a), M) * f).ravel() for a in (uhat, vhat, what) for j in range(3)]
    return np.stack(cols, axis=1)


def stress_hat_from_physical(g: Grid, fft: FFTBackend, tau_phys) -> list:
    """Spectra of stresses given in physical units on the padded grid, (M1 M2 M3, 6),
    in the reference's units: its padded fields are (3/2)^(-3/2) Nt^(-1/4) times
    physical (``src/Classes.py:593``), so a stress, quadratic in them, carries
    (3/2)^(-3) Nt^(-1/2) before the unnormalised r2c and unpad
    (``src/RHSfunctions.py:208-212``)."""
    M = ((3 * g.N1) // 2, (3 * g.N2) // 2, (3 * g.N3) // 2)
    conv = (2.0 / 3.0) ** 3 / np.sqrt(g.Nt)
    return [unpad(fft.r2c(np.asarray(tau_phys[:, c], np.float64).reshape(M) * conv)) for c in range(6)]


def rhs_learned(st: State, fft: FFTBackend, closure) -> np.ndarray:
    """``computeRHS_SMAG`` (``src/RHSfunctions.py:146-237``) with the Smagorinsky stress
    replaced by ``closure(gradients)`` evaluated pointwise on the padded grid:
    ``closure`` maps (n, 9) gradients to (n, 6) stresses (order 11 22 33 12 13 23,
    physical units) and may return (stresses, mask).  Returns what the closure
    returned besides the stresses (or None)."""
    g = st.grid
    st.Q2U()
    uu, vv, ww, uv, uw, vw = _products_32(st, fft, st.uhat, st.vhat, st.what)
    k1, k2, k3 = g.k1, g.k2, g.k3
    ki = g.ksqr_i
    phat = (-k1 * k1 * ki * uu - k2 * k2 * ki * vv - k3 * k3 * ki * ww - 2.0 * k1 * k2 * ki * uv
            - 2.0 * k1 * k3 * ki * uw - 2.0 * k2 * k3 * ki * vw)
    if st.rotate == 1:
        phat = phat - _coriolis_p(st, g, st.uhat, st.vhat, st.what, k1, k2, k3, ki)
    res = closure(padded_gradients(st, fft, st.uhat, st.vhat, st.what))
    extra = None
    if isinstance(res, tuple):
        res, extra = res
    tau = stress_hat_from_physical(g, fft, res)
    wu, wv, ww_ = sgs_divergence(g, tau)
    st.w0 = (wu, wv, ww_)
    ru = -1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw - 1j * k1 * phat - st.nu * g.ksqr * st.uhat + wu
    rv = -1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw - 1j * k2 * phat - st.nu * g.ksqr * st.vhat + wv
    rw = -1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww - 1j * k3 * phat - st.nu * g.ksqr * st.what + ww_
    if st.rotate == 1:
        ru, rv, rw = _coriolis_add(st, ru, rv, rw)
    if st.force != 0.0:
        ru, rv, rw = ru + st.force * st.uhat, rv + st.force * st.vhat, rw + st.force * st.what
    st.set_rhs(ru, rv, rw)
    return extra


def _hermitian_inner(a, b, n3h, Nt):
    """sum over the half spectrum with weight 2 for 0 < k3 < N3/2 and 1 for k3 = 0
    (the weighting of ``src/RHSfunctions.py:822-852``)."""
    return (np.sum(a[:, :, 1:n3h] * np.conj(b[:, :, 1:n3h] * 2))
            + np.sum(a[:, :, 0] * np.conj(b[:, :, 0]))) / Nt


def rhs_dtaumodel(st: State, fft: FFTBackend) -> None:
    """``computeRHS_Dtaumodel`` (``src/RHSfunctions.py:592-878``): the t-model pipeline
    on uhat and on the test-filtered uhat, a Leonard term and three reductions that
    set ``tau = |LE| / (|PLQLE_filt| - |filt_PLQLE| + 1e-5)``, clamped at 0."""
    g = st.grid
    st.Q2U()
    st.uhat = unpad(pad(st.uhat))
    st.vhat = unpad(pad(st.vhat))
    st.what = unpad(pad(st.what))
    kf = st.kf
    PL, real, prods, phat, _ = _pl_2x(st, fft, st.uhat, st.vhat, st.what)
    Qu, Qv, Qw = _plql(st, fft, PL, real, separate_modes)
    uf = test_filter(g, kf, st.uhat)
    vf = test_filter(g, kf, st.vhat)
    wf = test_filter(g, kf, st.what)
    PLf, realf, prodsf, phatf, padf = _pl_2x(st, fft, uf, vf, wf)
    Quf, Qvf, Qwf = _plql(st, fft, PLf, realf,
                          lambda x: separate_modes_testfilter(x, kf))
    k1, k2, k3 = g.k1f, g.k2f, g.k3f
    uu, vv, ww, uv, uw, vw = prods
    uuf, vvf, wwf, uvf, uwf, vwf = prodsf
    tf2 = lambda x: test_filter_2x(g, kf, x)  # noqa: E731
    Lu = (tf2(-1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw + 1j * k1 * phat)
          - tf2(-1j * k1 * uuf - 1j * k2 * uvf - 1j * k3 * uwf + 1j * k1 * phatf))
    Lv = (tf2(-1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw + 1j * k2 * phat)
          - tf2(-1j * k1 * uvf - 1j * k2 * vvf - 1j * k3 * vwf + 1j * k2 * phatf))
    Lw = (tf2(-1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww + 1j * k3 * phat)
          - tf2(-1j * k1 * uwf - 1j * k2 * vwf - 1j * k3 * wwf + 1j * k3 * phatf))
    n3h = g.N3 // 2
    upf, vpf, wpf = padf
    LE = (_hermitian_inner(Lu, upf, n3h, 1) + _hermitian_inner(Lv, vpf, n3h, 1)
          + _hermitian_inner(Lw, wpf, n3h, 1)) / g.Nt
    fQu, fQv, fQw = (test_filter(g, kf, x) for x in (Quf, Qvf, Qwf))
    PLQLE_filt = (_hermitian_inner(fQu, uf, n3h, 1) + _hermitian_inner(fQv, vf, n3h, 1)
                  + _hermitian_inner(fQw, wf, n3h, 1)) / g.Nt
    filt_PLQLE = (_hermitian_inner(Qu, uf, n3h, 1) + _hermitian_inner(Qv, vf, n3h, 1)
                  + _hermitian_inner(Qw, wf, n3h, 1)) / g.Nt
    tau = np.abs(LE) / (np.abs(PLQLE_filt) - np.abs(filt_PLQLE) + 1.0e-5)
    st.tau = tau
    st.tau_parts = (LE, PLQLE_filt, filt_PLQLE)
    tau = max(0.0, tau)
    _tau_family_finish(st, PL, (tau * Qu, tau * Qv, tau * Qw))


RHS = {
    "dns": rhs_nosgs,
    "smagorinsky": rhs_smag,
    "tmodel": rhs_tmodel,
    "static_tau": rhs_staumodel,
    "dynamic_tau": rhs_dtaumodel,
    "dsmag": rhs_dsmag,
}


def compute_rhs(st: State, fft: FFTBackend) -> None:
    """``main.computeRHS`` dispatch by ``turb_model`` (``src/Classes.py:43-52,95,124,186,227``)."""
    RHS[st.model](st, fft)


# --------------------------------------------------------------------------
# Time step, CFL and the low-storage RK advance
# --------------------------------------------------------------------------


def compute_dt(st: State) -> None:
    """``utilitiesClass.compute_dt`` (``src/Classes.py:687-698``), on the stale ``st.uhat``."""
    g = st.grid
    if st.cfl > 0:
        s = np.sqrt(g.Nt)
        u = np.fft.irfftn(st.uhat, s=(g.N1, g.N2, g.N3), axes=(0, 1, 2)) * s
        v = np.fft.irfftn(st.vhat, s=(g.N1, g.N2, g.N3), axes=(0, 1, 2)) * s
        w = np.fft.irfftn(st.what, s=(g.N1, g.N2, g.N3), axes=(0, 1, 2)) * s
        max_vel = np.amax(abs(u) / g.dx + abs(v) / g.dy + abs(w) / g.dz)
        st.dt = st.cfl / (max_vel + 1e-10)
        if st.nu > 0:
            st.dt = np.minimum(st.dt, 0.634 / (1.0 / g.dx ** 2 + 1.0 / g.dy ** 2 + 1.0 / g.dz ** 2)
                               / st.nu * st.cfl / 1.35)
        st.dt = float(st.dt)
    else:
        st.dt = -st.cfl


RK_COEFFS = (1.0 / 4, 1.0 / 3, 1.0 / 2, 1.0)


def advance_rk4(st: State, fft: FFTBackend) -> None:
    """``advanceQ_RK4`` (``src/PySpec_3d.py:96-105``): Q0 copy, CFL dt on the stale
    uhat, then four stages ``Q = Q0 + dt*c_i*RHS(Q)``."""
    Q0 = st.Q.copy()
    compute_dt(st)
    for c in RK_COEFFS:
        compute_rhs(st, fft)
        st.Q = Q0 + st.dt * c * st.Q


# --------------------------------------------------------------------------
# Diagnostics (src/Classes.py:652-797)
# --------------------------------------------------------------------------


def _half_energy(a, n3h):
    return np.sum(a[:, :, 1:n3h] * np.conj(a[:, :, 1:n3h] * 2)) + np.sum(a[:, :, 0] * np.conj(a[:, :, 0]))


def compute_energy(st: State) -> float:
    """``computeEnergy`` (``src/Classes.py:652-659``)."""
    g = st.grid
    n3h = g.N3 // 2
    e = _half_energy(st.uhat, n3h) + _half_energy(st.vhat, n3h) + _half_energy(st.what, n3h)
    return float(np.real(0.5 * e / g.Nt))


def compute_energy_resolved(st: State) -> float:
    """``computeEnergy_resolved`` (``src/Classes.py:666-676``), sharp cutoff ``Gf``."""
    g = st.grid
    n3h = g.N3 // 2
    e = sum(_half_energy(g.Gf * a, n3h) for a in (st.uhat, st.vhat, st.what))
    return float(np.real(0.5 * e / g.Nt))


def compute_enstrophy(st: State) -> float:
    """``computeEnstrophy`` (``src/Classes.py:733-744``)."""
    g = st.grid
    n3h = g.N3 // 2
    o1 = 1j * g.k2 * st.what - 1j * g.k3 * st.vhat
    o2 = 1j * g.k3 * st.uhat - 1j * g.k1 * st.what
    o3 = 1j * g.k1 * st.vhat - 1j * g.k2 * st.uhat
    e = _half_energy(o1, n3h) + _half_energy(o2, n3h) + _half_energy(o3, n3h)
    return float(np.real(0.5 * e / g.Nt))


def compute_all_stats(st: State):
    """``computeAllStats`` (``src/Classes.py:701-731``)."""
    enstrophy = compute_enstrophy(st)
    energy = compute_energy(st)
    dissipation = 2 * enstrophy * st.nu
    lambda_k = (st.nu ** 3 / dissipation) ** 0.25
    tau_k = (st.nu / dissipation) ** 0.5
    Re_lambda = energy * np.sqrt(20.0 / (3.0 * st.nu * dissipation))
    return enstrophy, energy, dissipation, lambda_k, tau_k, Re_lambda


def compute_spectrum(st: State, resolved: bool = False):
    """Shell spectrum ``computeSpectrum`` / ``computeSpectrum_resolved``
    (``src/Classes.py:762-797,825-844``): bins ``rint(|k|)`` in integer wavenumbers,
    weight 2 for 0 < k3 < N3/2, 1 for k3 = 0, k3 = N3/2 excluded.  Bins are plain
    integers here; the reference casts them to ``int8``, which wraps for |k| >= 128."""
    g = st.grid
    n3h = g.N3 // 2
    ksq = ((g.k1 * g.L1 / (2 * np.pi)) ** 2 + (g.k2 * g.L2 / (2 * np.pi)) ** 2
           + (g.k3 * g.L3 / (2 * np.pi)) ** 2)
    bm = np.rint(np.sqrt(ksq[:, :, 1:n3h])).astype(np.int64).ravel()
    b0 = np.rint(np.sqrt(ksq[:, :, 0])).astype(np.int64).ravel()
    kmax = int(bm.max())
    spec = np.zeros((kmax + 1, 3))
    for c, a in enumerate((st.uhat, st.vhat, st.what)):
        if resolved:
            a = g.Gf * a
        am = a[:, :, 1:n3h].ravel()
        a0 = a[:, :, 0].ravel()
        np.add.at(spec[:, c], bm, 2 * np.real(am * np.conj(am)))
        np.add.at(spec[:, c], b0, np.real(a0 * np.conj(a0)))
    return np.arange(kmax + 1, dtype=float), spec / g.Nt


# --------------------------------------------------------------------------
# Initial conditions (case setup, run/*/inputfile.py)
# --------------------------------------------------------------------------


def hermitian_symmetrize(uhat: np.ndarray, N3: int) -> np.ndarray:
    """``rfftn(irfftn(uhat))``: the Hermitian projection of a half spectrum."""
    N1, N2, _ = uhat.shape
    return np.fft.rfftn(np.fft.irfftn(uhat, s=(N1, N2, N3), axes=(0, 1, 2)))


def taylor_green(N1, N2, N3, grid: str = "exact"):
    """TGV ``u = (sin x cos y cos z, -cos x sin y cos z, 0)`` sampled on the exact
    periodic grid ``x_j = 2 pi j / N`` (``grid='exact'``) or on the reference's case grid
    ``linspace(0, 2pi - dx, N)`` with ``dx = 2pi/(N-1)`` (``grid='reference'``,
    ``run/TGV_Validate100/inputfile.py:31-44``).  On the reference grid the field is
    the case file's ``u = cos x sin y cos z, v = -sin x cos y cos z`` (arrays indexed
    (x, y, z) through its meshgrid(y, x, z) idiom); that is the same vortex shifted
    by pi/2 in x and y, with identical energy and dissipation histories on an
    exact grid.  Returns ``(uhat, vhat, what, (dx, dy, dz))``."""
    def axis(n):
        # the spacing gridclass reads off the case arrays, x[1] - x[0] (src/Classes.py:562-564)
        if grid == "exact":
            d = 2 * np.pi / n
            return np.arange(n) * d, d
        d = 2 * np.pi / (n - 1)
        x = np.linspace(0, 2 * np.pi - d, n)
        return x, x[1] - x[0]

    x, dx = axis(N1)
    y, dy = axis(N2)
    z, dz = axis(N3)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    if grid == "exact":
        u = np.sin(X) * np.cos(Y) * np.cos(Z)
        v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    else:
        u = np.cos(X) * np.sin(Y) * np.cos(Z)
        v = -np.sin(X) * np.cos(Y) * np.cos(Z)
    w = np.zeros_like(u)
    s = np.sqrt(N1 * N2 * N3)
    return np.fft.rfftn(u) / s, np.fft.rfftn(v) / s, np.fft.rfftn(w) / s, (dx, dy, dz)


def random_isotropic(N, seed=0, kp=4.0, sigma=4.0, L=2 * np.pi, urms=1.0):
    """Seeded random-phase divergence-free field with E(k) ~ k^sigma exp(-sigma/2 (k/kp)^2)
    (sigma = 4, kp = 4 gives k^4 exp(-2 (k/4)^2)), following the Mansour-Wray
    construction of ``run/HIT/MansourWray/inputfile.py:69-103``; k = 0 and the
    Nyquist planes are zeroed, the field is Hermitian-symmetrised and scaled by one
    common factor to E = 3/2 urms^2.  Phases come from ``numpy.random.default_rng(seed)``."""
    N1 = N2 = N3 = int(N)
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftshift(np.linspace(-N1 // 2, N1 // 2 - 1, N1))
    k2 = np.fft.fftshift(np.linspace(-N2 // 2, N2 // 2 - 1, N2))
    k3 = np.linspace(0, N3 // 2, N3 // 2 + 1)
    K1, K2, K3 = np.meshgrid(k1, k2, k3, indexing="ij")
    ksqr = K1 ** 2 + K2 ** 2 + K3 ** 2
    kmag = np.sqrt(ksqr) + 1e-30
    E = (kmag / kp) ** sigma * np.exp(-sigma / 2.0 * ksqr / kp ** 2)
    th1 = rng.random(K1.shape) * 2 * np.pi
    th2 = rng.random(K1.shape) * 2 * np.pi
    phi = rng.random(K1.shape) * 2 * np.pi
    amp = np.sqrt(E / (4 * np.pi * ksqr + 1e-30))
    alpha = amp * np.exp(1j * th1) * np.cos(phi)
    beta = amp * np.exp(1j * th2) * np.sin(phi)
    kp12 = np.sqrt(K1 ** 2 + K2 ** 2)
    s = 2 * np.pi / L
    uh = (alpha * kmag * K2 * s + beta * K1 * s * K3 * s) / (kmag * kp12 + 1e-30)
    vh = (-alpha * kmag * K1 * s + beta * K2 * s * K3 * s) / (kmag * kp12 + 1e-30)
    wh = -beta * kp12 / kmag
    out = []
    for a in (uh, vh, wh):
        a = a.copy()
        a[0, 0, 0] = 0
        a[N1 // 2, :, :] = 0
        a[:, N2 // 2, :] = 0
        a[:, :, N3 // 2] = 0
        a = hermitian_symmetrize(a, N3)
        out.append(a)
    # the (k1,k2)=(0,0) column of the construction is singular; remove its
    # divergence by a Leray projection so the field is exactly solenoidal
    kk = ksqr + 1e-300
    div = K1 * out[0] + K2 * out[1] + K3 * out[2]
    out = [out[0] - K1 * div / kk, out[1] - K2 * div / kk, out[2] - K3 * div / kk]
    for a in out:
        a[0, 0, 0] = 0
    # one common factor (so the field stays solenoidal) giving E = 3/2 u_rms^2
    e = sum(np.real(_half_energy(a, N3 // 2)) for a in out) / (N1 * N2 * N3)
    f = urms * np.sqrt(3.0 / e)
    return out[0] * f, out[1] * f, out[2] * f


# --------------------------------------------------------------------------
# Driver loop (src/PySpec_3d.py:107-148)
# --------------------------------------------------------------------------


def run(st: State, fft: FFTBackend, et: float, max_steps=None, callback=None):
    """The driver's ``while t <= et`` loop: advance, ``t += dt``, stats on the stale
    uhat.  Returns the history arrays written to ``stats.npz``."""
    E = [compute_energy(st)]
    Ens = [compute_enstrophy(st)]
    Er = [compute_energy_resolved(st)]
    th = [st.t]
    t = st.t
    it = 0
    while t <= et:
        if max_steps is not None and it >= max_steps:
            break
        st.t = t
        st.iteration = it
        advance_rk4(st, fft)
        t += st.dt
        it += 1
        ens, en, *_ = compute_all_stats(st)
        E.append(en)
        Ens.append(ens)
        Er.append(compute_energy_resolved(st))
        th.append(t)
        if callback is not None:
            callback(st, it, t)
    st.t = t
    th = np.array(th)
    E = np.array(E)
    Er = np.array(Er)
    D = (E[1:] - E[:-1]) / (th[1:] - th[:-1])
    Dr = (Er[1:] - Er[:-1]) / (th[1:] - th[:-1])
    return dict(t=th, Energy=E, Enstrophy=np.array(Ens), Energy_resolved=Er,
                Dissipation=D, Dissipation_resolved=Dr)


def default_workers() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


PARITY = {
    "pad/unpad/pad_2x/unpad_2x": "pinned: brute-force DFT interpolation (tests/test_oracle_fft.py)",
    "separate_modes*": "pinned: mask counts and p+q identity (tests/test_oracle_fft.py)",
    "_products_32": "pinned: brute-force truncated convolution at N=8 (tests/test_oracle_rhs.py)",
    "rhs_nosgs": "pinned: TGV closed form, k.RHS=0, energy conservation (tests/test_oracle_rhs.py)",
    "rhs_smag": "pinned: pointwise closed form for a shear mode (tests/test_oracle_rhs.py)",
    "rhs_tmodel/static_tau": "pinned: brute-force convolution restatement at N=8, k.w0 invariants (tests/test_oracle_tmodel.py)",
    "rhs_dtaumodel": "pinned: reduces to the static tau-model form with tau from the three "
                     "reductions LE, PLQLE_filt and filt_PLQLE, each rebuilt from direct convolutions "
                     "at N=8 (tests/test_oracle_tmodel.py); the Leonard pressure sign is as written",
    "advance_rk4": "pinned: linear mode amplification R(z)^n (tests/test_oracle_rhs.py)",
    "compute_dt": "pinned: closed form from max|u| (tests/test_oracle_rhs.py)",
    "rhs_dsmag": "pinned: Cs^2 as physical-space sums (Parseval), Leonard stress by brute force, "
                 "reduces to rhs_smag at the dynamic coefficient (tests/test_oracle_dsmag.py)",
    "sgs_dns/box_transfer": "pinned: filtered convolution differences by brute force, box transfer as "
                            "the top-hat average (tests/test_oracle_sgs.py)",
    "physical/gradients/restrict/padded_gradients": "pinned: analytic band-limited fields and their "
                                                    "derivatives on both grids (tests/test_oracle_learned.py)",
    "rhs_learned/stress_hat_from_physical": "pinned: with the pointwise Smagorinsky closure it equals "
                                            "rhs_smag (tests/test_oracle_learned.py)",
    "mlp_oracle.forward/grad/bf16": "pinned: torch bfloat16 conversion, torch.nn forward and autograd "
                                    "(tests/test_oracle_mlp.py)",
    "stats": "pinned: Parseval against physical-space sums (tests/test_oracle_rhs.py); "
             "E(t) against the shipped TGV fixture (tests/test_oracle_fixture.py)",
}
