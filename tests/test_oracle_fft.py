"""Pin the oracle's padding/truncation and transform conventions by brute force."""
import numpy as np
import pytest

from oracle import spectral_oracle as so
from tests import _bruteforce as bf


def _rand_field(N, seed):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((N, N, N))
    return np.fft.rfftn(u) / np.sqrt(N ** 3)


@pytest.mark.parametrize("N", [6, 8])
def test_pad_inverse_is_truncated_interpolant(N):
    """pad + normalised c2r * scale samples the Nyquist-truncated interpolant on the
    3N/2 grid, scaled by (3/2)^(-3/2) Nt^(-1/4) (the reference's ureal convention)."""
    uh = _rand_field(N, 1)
    M = 3 * N // 2
    Nt = N ** 3
    scale = np.sqrt(1.5 ** 3 * np.sqrt(Nt))
    ur = np.fft.irfftn(so.pad(uh), s=(M, M, M)) * scale
    C = bf.half_to_full(uh, N)
    x = np.arange(M) * 2 * np.pi / M
    pts = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1)
    expect = bf.synth(C, pts) * np.sqrt(Nt) * scale / M ** 3
    assert np.allclose(ur, expect, rtol=0, atol=1e-13 * np.abs(expect).max())


@pytest.mark.parametrize("N", [6, 8])
def test_unpad_forward_is_truncated_dft(N):
    """unpad(r2c(f)) on the padded grid equals the brute-force DFT at kept modes;
    Nyquist entries are zero."""
    M = 3 * N // 2
    rng = np.random.default_rng(2)
    f = rng.standard_normal((M, M, M))
    got = so.unpad(np.fft.rfftn(f))
    x = np.arange(M) * 2 * np.pi / M
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    k = bf.kgrid(N)
    for i in range(N):
        for j in range(N):
            for l in range(N // 2 + 1):
                if i == N // 2 or j == N // 2 or l == N // 2:
                    assert got[i, j, l] == 0
                    continue
                e = np.sum(f * np.exp(-1j * (k[i] * X + k[j] * Y + l * Z)))
                assert abs(got[i, j, l] - e) < 1e-11 * np.abs(f).sum()


def test_pad_roundtrip_and_shapes():
    N = 8
    uh = _rand_field(N, 3)
    p = so.pad(uh)
    assert p.shape == (12, 12, 7)
    back = so.unpad(p)
    ref = uh.copy()
    ref[N // 2] = 0
    ref[:, N // 2] = 0
    ref[:, :, N // 2] = 0
    assert np.array_equal(back, ref)
    p2 = so.pad_2x(uh)
    assert p2.shape == (16, 16, 9)
    back2 = so.unpad_2x(p2)
    ref2 = uh.copy()
    ref2[:, :, N // 2] = 0          # 2x rule keeps the -N/2 rows, drops k3 = N/2
    assert np.array_equal(back2, ref2)


def test_pad2x_inverse_is_interpolant_with_oddball():
    """pad_2x keeps the k = -N/2 rows: they land at index 2N - N/2."""
    N = 8
    uh = np.zeros((N, N, N // 2 + 1), complex)
    uh[N // 2, 1, 1] = 1.0 + 0.5j       # k = (-4, 1, 1)
    p = so.pad_2x(uh)
    nz = np.argwhere(p != 0)
    assert nz.tolist() == [[2 * N - N // 2, 1, 1]]


def test_separate_modes_partition():
    N = 8
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2 * N, 2 * N, N + 1)) + 1j * rng.standard_normal((2 * N, 2 * N, N + 1))
    p, q = so.separate_modes(x)
    assert np.array_equal(p + q, x)
    k = bf.kgrid(2 * N)
    K1, K2 = np.meshgrid(k, k, indexing="ij")
    inside = (np.abs(K1) < N // 2) & (np.abs(K2) < N // 2)
    for l in range(N + 1):
        expect = inside if l < N // 2 else np.zeros_like(inside)
        assert np.array_equal(p[:, :, l] != 0, expect)
    kf = 3
    p, q = so.separate_modes_testfilter(x, kf)
    assert np.array_equal(p + q, x)
    inside = (np.abs(K1) < kf) & (np.abs(K2) < kf)
    for l in range(N + 1):
        expect = inside if l < kf else np.zeros_like(inside)
        assert np.array_equal(p[:, :, l] != 0, expect)


def test_grid_wavenumbers_and_filter():
    g = so.Grid(8, 8, 8, kc=3)
    assert g.k1[:, 0, 0].tolist() == [0, 1, 2, 3, -4, -3, -2, -1]
    assert g.k3[0, 0, :].tolist() == [0, 1, 2, 3, 4]
    # Gf keeps |k1|,|k2| < kc and k3 < kc
    inside = (np.abs(g.k1) < 3) & (np.abs(g.k2) < 3) & (g.k3 < 3)
    assert np.array_equal(g.Gf.astype(bool), inside)
