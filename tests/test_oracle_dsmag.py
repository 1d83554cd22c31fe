"""Pin the oracle's dynamic Smagorinsky closure (computeRHS_DSMAG,
src/RHSfunctions.py:1663-1862) by brute force, Parseval and its reduction to the
static Smagorinsky closure."""
import numpy as np

from oracle import spectral_oracle as so
from tests import _bruteforce as bf

FFT = so.FFTBackend()


def _half_inner(a, b, N3):
    """sum over the full spectrum of Re(a b*) from a half spectrum (weight 2 for
    0 < k3 < N3/2)."""
    w = np.full(a.shape[-1], 2.0)
    w[0] = 1.0
    w[N3 // 2:] = 0.0 if a.shape[-1] > N3 // 2 else w[N3 // 2:]
    return float(np.sum(w * np.real(a * np.conj(b))))


def test_cs2_equals_spectral_inner_products():
    """mean over the padded grid of M_ij L_ij (six components) is a Hermitian inner
    product of band-limited spectra, so Cs^2 = <M,L>/<M,M> computed spectrally must
    equal the oracle's physical-space ratio."""
    N = 16
    u = so.random_isotropic(N, seed=2, kp=2.0)
    st = so.State(so.Grid(N, N, N), *u, nu=0.01, turb_model=7)
    cs2, L, M, _ = so.dsmag_coefficient(st, FFT, *u)
    num = sum(_half_inner(m, l, N) for m, l in zip(M, L))
    den = sum(_half_inner(m, m, N) for m in M)
    assert abs(cs2 - num / den) < 1e-12 * abs(cs2)


def test_leonard_stress_bruteforce():
    """L_ij = G (u_i u_j) - (G u_i)(G u_j), G the sharp test filter at kL = kc/2,
    with every product a direct truncated convolution."""
    N, kc = 8, 6
    u = so.random_isotropic(N, seed=4, kp=2.0)
    st = so.State(so.Grid(N, N, N, kc=kc), *u, nu=0.01, turb_model=7)
    _, L, _, _ = so.dsmag_coefficient(st, FFT, *u)
    kL = kc // 2
    k1 = np.fft.fftfreq(N, 1.0 / N)[:, None, None]
    k3 = np.arange(N // 2 + 1)[None, None, :]
    G = ((np.abs(k1) < kL) & (np.abs(k1.transpose(1, 0, 2)) < kL) & (k3 < kL)).astype(float)
    C = [bf.half_to_full(a, N) for a in u]
    Cf = [bf.half_to_full(G * a, N) for a in u]
    keep = list(range(-(N // 2 - 1), N // 2))
    s = np.sqrt(N ** 3)
    for (i, j), l in zip([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)], L):
        full = bf.full_to_half(bf.conv(C[i], C[j]), N, N, N // 2 + 1, keep, N // 2 - 1, s)
        filt = bf.full_to_half(bf.conv(Cf[i], Cf[j]), N, N, N // 2 + 1, keep, N // 2 - 1, s)
        assert np.abs(l - (G * full - filt)).max() < 1e-13 * np.abs(full).max()


def test_m_reduces_inside_the_test_filter():
    """A field whose modes all pass the test filter has S^F = S, so
    M = -2 Delta^2 (alpha^2 - G) FFT(|S| S)."""
    N, kc = 16, 16
    u = so.random_isotropic(N, seed=3, kp=1.0)
    g = so.Grid(N, N, N, kc=kc)
    G = so.Grid(N, N, N, kc=kc // 2).Gf
    u = [G * a for a in u]
    st = so.State(g, *u, nu=0.01, turb_model=7)
    _, L, M, SS = so.dsmag_coefficient(st, FFT, *u)
    d2 = g.Delta ** 2
    for m, ss in zip(M, SS):
        assert np.abs(m - (-2.0 * d2 * (4.0 - G) * ss)).max() < 1e-12 * np.abs(m).max()


def test_dsmag_is_smagorinsky_at_the_dynamic_coefficient(monkeypatch):
    """With Cs^2 > 0 the dynamic RHS equals the static Smagorinsky RHS at Cs = sqrt(Cs^2)."""
    N = 16
    u = so.random_isotropic(N, seed=2, kp=2.0)
    g = so.Grid(N, N, N)
    st = so.State(g, *u, nu=0.013, turb_model=7)
    so.compute_rhs(st, FFT)
    assert st.cs2 > 0
    monkeypatch.setattr(so, "SMAG_CS", float(np.sqrt(st.cs2)))
    sm = so.State(g, *u, nu=0.013, turb_model=1)
    so.compute_rhs(sm, FFT)
    assert np.abs(st.compact() - sm.compact()).max() < 1e-12 * np.abs(sm.compact()).max()
    assert max(np.abs(a - b).max() for a, b in zip(st.w0, sm.w0)) < 1e-12 * np.abs(sm.w0[0]).max()
