"""Brute-force spectral helpers for pinning the oracle (test-only).

Fields are held as *centred full* arrays of Fourier-series coefficients
``C[k1+K1, k2+K2, k3+K3]`` with ``u(x) = sum_k C_k exp(i k.x)``.  Products are
formed by direct summation (``scipy.signal.convolve(method='direct')``), never by
an FFT, so they are independent of every padding/index convention in the oracle.
Half spectra use the reference convention ``uhat = sqrt(Nt) * C`` (Nt = N1 N2 N3
of the base grid).
"""
import numpy as np
from scipy.signal import convolve


def half_to_full(uh, N, K=None, scale=None):
    """Hermitian-extend a half spectrum ``(n1, n2, n3c)`` in FFT order into a centred
    full array keeping |k_i| <= K (default N/2 - 1 on every axis)."""
    n1, n2, n3c = uh.shape
    if K is None:
        K = N // 2 - 1
    if scale is None:
        scale = np.sqrt(float(N) ** 3)
    C = np.zeros((2 * K + 1,) * 3, dtype=complex)
    ks = np.arange(-K, K + 1)
    for a in ks:
        for b in ks:
            for c in range(0, min(K, n3c - 1) + 1):
                val = uh[a % n1, b % n2, c] / scale
                C[a + K, b + K, c + K] = val
                if c > 0:
                    C[-a + K, -b + K, -c + K] = np.conj(val)
    return C


def full_to_half(C, n1, n2, n3c, kx_keep, kz_max, scale):
    """Read a centred full array back into an FFT-ordered half spectrum of shape
    ``(n1, n2, n3c)``; ``kx_keep`` lists the signed k1 (and k2) values written,
    ``kz_max`` the largest k3 written."""
    K = (C.shape[0] - 1) // 2
    out = np.zeros((n1, n2, n3c), dtype=complex)
    for a in kx_keep:
        for b in kx_keep:
            for c in range(0, kz_max + 1):
                if abs(a) <= K and abs(b) <= K and c <= K:
                    out[a % n1, b % n2, c] = C[a + K, b + K, c + K] * scale
    return out


def conv(A, B):
    """Exact product spectrum of two centred full arrays (direct summation)."""
    return convolve(A, B, method="direct")


def synth(C, pts):
    """Evaluate ``sum_k C_k exp(i k.x)`` (real part) at points ``pts`` (..., 3)."""
    K = (C.shape[0] - 1) // 2
    ks = np.arange(-K, K + 1)
    ex = np.exp(1j * pts[..., 0, None] * ks)
    ey = np.exp(1j * pts[..., 1, None] * ks)
    ez = np.exp(1j * pts[..., 2, None] * ks)
    val = np.einsum("abc,...a,...b,...c->...", C, ex, ey, ez)
    return val.real


def kgrid(n, L=2 * np.pi):
    """Signed wavenumbers of an n-point FFT-ordered axis (as gridclass builds them)."""
    return np.fft.fftshift(np.linspace(-n // 2, n // 2 - 1, n)) * 2 * np.pi / L
