"""Pin the oracle's filtered-DNS subgrid stress (computeSGS_DNS, src/Classes.py:861-929)
and its filters against brute force and closed forms."""
import numpy as np
from scipy import integrate

from oracle import spectral_oracle as so
from tests import _bruteforce as bf

FFT = so.FFTBackend()


def _mode_index(k, n):
    return k % n


def test_sgs_tau_is_filtered_convolution_difference():
    """tau_ij = G (u_i u_j)^ - (ubar_i ubar_j)^ with every product a brute-force
    truncated convolution and G the sharp mask |k_i| < kc (index form of grid.Gf)."""
    N, kc = 8, 3
    uh = so.random_isotropic(N, seed=5, kp=2.0)
    g = so.Grid(N, N, N, kc=kc)
    st = so.State(g, *uh, nu=0.01)
    _, tau = so.sgs_dns(st, FFT)
    # the sharp mask stated on signed integer wavenumbers
    k1 = np.fft.fftfreq(N, 1.0 / N)[:, None, None]
    k3 = np.arange(N // 2 + 1)[None, None, :]
    G = ((np.abs(k1) < kc) & (np.abs(k1.transpose(1, 0, 2)) < kc) & (k3 < kc)).astype(float)
    C = [bf.half_to_full(a, N) for a in uh]
    Cf = [bf.half_to_full(G * a, N) for a in uh]
    keep = list(range(-(N // 2 - 1), N // 2))
    s = np.sqrt(N ** 3)
    for (i, j), t in zip([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)], tau):
        full = bf.full_to_half(bf.conv(C[i], C[j]), N, N, N // 2 + 1, keep, N // 2 - 1, s)
        filt = bf.full_to_half(bf.conv(Cf[i], Cf[j]), N, N, N // 2 + 1, keep, N // 2 - 1, s)
        expect = G * full - filt
        assert np.abs(t - expect).max() < 1e-13 * np.abs(full).max()


def test_sgs_vanishes_without_filtering_and_w0_is_its_divergence():
    N = 16
    uh = so.random_isotropic(N, seed=6, kp=3.0)
    g = so.Grid(N, N, N)                  # kc = N/2: the field has no -N/2 modes, G u = u
    st = so.State(g, *uh, nu=0.01)
    w0, tau = so.sgs_dns(st, FFT)
    assert max(np.abs(t).max() for t in tau) == 0.0
    g4 = so.Grid(N, N, N, kc=4)
    st4 = so.State(g4, *uh, nu=0.01)
    w0, tau = so.sgs_dns(st4, FFT)
    assert min(np.abs(t).max() for t in tau) > 0.0
    # special case of the divergence form: an isotropic stress tau_ij = delta_ij phi
    # gives w0 = -(1 + i) k phi (the reference's projection term carries a factor i)
    phi = tau[0]
    z = np.zeros_like(phi)
    wi = so.sgs_divergence(g4, [phi, phi, phi, z, z, z])
    for kk, w in zip((g4.k1, g4.k2, g4.k3), wi):
        expect = -(1 + 1j) * kk * phi
        assert np.abs(w - expect).max() < 1e-12 * np.abs(expect).max()


def test_box_transfer_is_the_top_hat_average():
    """The box filter's transfer function equals the top-hat average of cos(k x) over
    [x - delta/2, x + delta/2] divided by cos(k x), computed by quadrature."""
    N = 16
    g = so.Grid(N, N, N)
    delta = 3 * 2 * np.pi / N
    G = so.box_transfer(g, delta)
    for (a, b, c) in [(1, 0, 0), (0, 3, 0), (0, 0, 5), (2, -3, 4), (-7, 7, 7)]:
        val = 1.0
        for kk in (a, b, c):
            avg, _ = integrate.quad(lambda s: np.cos(kk * s), -delta / 2, delta / 2)
            val *= avg / delta
        assert abs(G[_mode_index(a, N), _mode_index(b, N), c] - val) < 1e-13
    # a box filter of a field is the physical top-hat average: check one mode end to end
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.cos(2 * X + 3 * Z)
    uh = np.fft.rfftn(u) * G
    got = np.fft.irfftn(uh, s=(N, N, N), axes=(0, 1, 2))
    sx = np.sin(2 * delta / 2) / (2 * delta / 2)
    sz = np.sin(3 * delta / 2) / (3 * delta / 2)
    assert np.abs(got - sx * sz * u).max() < 1e-13
