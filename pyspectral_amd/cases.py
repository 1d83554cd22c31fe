"""Case setup: initial fields in the reference's Fourier convention.

Host-side numpy, run once before the solve (the reference's case scripts,
``run/*/inputfile.py``): ``uhat = rfftn(u) / sqrt(N1 N2 N3)`` per component.
"""
from __future__ import annotations

import numpy as np


def taylor_green(N: int, grid: str = "exact"):
    """Taylor-Green vortex, stacked ``(3, N, N, N//2+1)`` complex128.

    ``grid='exact'``: u = (sin x cos y cos z, -cos x sin y cos z, 0) on x_j = 2 pi j/N.
    ``grid='reference'``: the case grid of ``run/TGV_Validate100/inputfile.py:31-44``,
    ``linspace(0, 2pi - dx, N)`` with dx = 2pi/(N-1) and u = (cos x sin y cos z,
    -sin x cos y cos z, 0).  Returns ``(uhat, (dx, dy, dz))`` with the spacing of the
    sample points.
    """
    if grid == "exact":
        d = 2 * np.pi / N
        x = np.arange(N) * d
    elif grid == "reference":
        x = np.linspace(0, 2 * np.pi - 2 * np.pi / (N - 1), N)
        d = x[1] - x[0]  # the spacing gridclass reads off the case arrays (src/Classes.py:562-564)
    else:
        raise ValueError(grid)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    if grid == "exact":
        u = np.sin(X) * np.cos(Y) * np.cos(Z)
        v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    else:
        u = np.cos(X) * np.sin(Y) * np.cos(Z)
        v = -np.sin(X) * np.cos(Y) * np.cos(Z)
    s = np.sqrt(float(N) ** 3)
    uh = np.stack([np.fft.rfftn(u), np.fft.rfftn(v), np.zeros((N, N, N // 2 + 1), complex)]) / s
    return uh, (d, d, d)


def decaying_isotropic(N: int, seed: int = 0, k0: float = 4.0, sigma: float = 4.0, urms: float = 1.0):
    """Seeded random-phase solenoidal field with E(k) ~ (k/k0)^sigma exp(-sigma/2 (k/k0)^2),
    i.e. k^4 exp(-2 (k/4)^2) for the defaults, following the Mansour-Wray construction
    of ``run/HIT/MansourWray/inputfile.py:69-103``.  k = 0 and the Nyquist planes are
    zero, the half spectrum is Hermitian (``rfftn(irfftn(.))``) and the field is
    scaled so that E = 3/2 urms^2.  Returns ``(3, N, N, N//2+1)`` complex128."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(N, 1.0 / N)
    k3 = np.arange(N // 2 + 1, dtype=float)
    K1, K2, K3 = np.meshgrid(k1, k1, k3, indexing="ij")
    K1[N // 2], K2[:, N // 2] = -N / 2, -N / 2
    ksq = K1 ** 2 + K2 ** 2 + K3 ** 2
    kmag = np.sqrt(ksq) + 1e-30
    E = (kmag / k0) ** sigma * np.exp(-sigma / 2.0 * ksq / k0 ** 2)
    th1, th2, phi = (rng.random(K1.shape) * 2 * np.pi for _ in range(3))
    amp = np.sqrt(E / (4 * np.pi * ksq + 1e-30))
    a = amp * np.exp(1j * th1) * np.cos(phi)
    b = amp * np.exp(1j * th2) * np.sin(phi)
    kp = np.sqrt(K1 ** 2 + K2 ** 2)
    comps = [(a * kmag * K2 + b * K1 * K3) / (kmag * kp + 1e-30),
             (-a * kmag * K1 + b * K2 * K3) / (kmag * kp + 1e-30),
             -b * kp / kmag]
    out = []
    for c in comps:
        c[0, 0, 0] = 0
        c[N // 2] = 0
        c[:, N // 2] = 0
        c[:, :, N // 2] = 0
        out.append(np.fft.rfftn(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))))
    div = K1 * out[0] + K2 * out[1] + K3 * out[2]
    out = [out[0] - K1 * div / (ksq + 1e-300), out[1] - K2 * div / (ksq + 1e-300), out[2] - K3 * div / (ksq + 1e-300)]
    for c in out:
        c[0, 0, 0] = 0
    u = np.stack(out)
    e = (np.sum(np.abs(u[:, :, :, 1:N // 2]) ** 2) * 2 + np.sum(np.abs(u[:, :, :, 0]) ** 2)) / float(N) ** 3
    return u * (urms * np.sqrt(3.0 / e))


def prolong(u: np.ndarray, N1: int, N2: int | None = None, N3: int | None = None) -> np.ndarray:
    """Spectral interpolation of a state (3, n1, n2, n3/2+1) onto a finer grid
    (3, N1, N2, N3/2+1): the signed modes |k_i| < n_i/2 are copied (the coarse grid's
    Nyquist entries dropped), scaled by sqrt(N1 N2 N3 / (n1 n2 n3)) so that grid values
    of the band-limited field are kept.  Used to start grids of several resolutions from
    one seeded field."""
    N2 = N1 if N2 is None else N2
    N3 = N1 if N3 is None else N3
    _, n1, n2, nc = u.shape
    n3 = 2 * (nc - 1)
    if N1 < n1 or N2 < n2 or N3 < n3:
        raise ValueError("the target grid must be at least as fine as the source grid")
    out = np.zeros((u.shape[0], N1, N2, N3 // 2 + 1), dtype=complex)
    s1 = np.r_[0:n1 // 2, n1 // 2 + 1:n1]
    d1 = np.r_[0:n1 // 2, N1 - n1 // 2 + 1:N1]
    s2 = np.r_[0:n2 // 2, n2 // 2 + 1:n2]
    d2 = np.r_[0:n2 // 2, N2 - n2 // 2 + 1:N2]
    k3 = np.arange(n3 // 2)
    scale = np.sqrt((N1 * N2 * N3) / (n1 * n2 * n3))
    for c in range(u.shape[0]):
        out[c][np.ix_(d1, d2, k3)] = u[c][np.ix_(s1, s2, k3)] * scale
    return out
