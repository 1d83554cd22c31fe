"""Host-side case setup: spectral prolongation keeps the grid values of a band-limited
field and restricts back to the original modes (Nyquist entries dropped)."""
import numpy as np

from oracle import spectral_oracle as so
from pyspectral_amd import cases


def test_prolong_keeps_grid_values_and_restricts_back():
    u = cases.decaying_isotropic(16, seed=3)
    U = cases.prolong(u, 32, 48, 64)
    x = np.fft.irfftn(u[1], s=(16, 16, 16), axes=(0, 1, 2)) * np.sqrt(16 ** 3)
    X = np.fft.irfftn(U[1], s=(32, 48, 64), axes=(0, 1, 2)) * np.sqrt(32 * 48 * 64)
    assert np.abs(X[::2, ::3, ::4] - x).max() < 1e-13 * np.abs(x).max()
    back = np.stack([so.restrict(f, so.Grid(16, 16, 16)) for f in U])
    ref = u.copy()
    ref[:, 8] = 0
    ref[:, :, 8] = 0
    ref[..., 8] = 0
    assert np.abs(back - ref).max() < 1e-14 * np.abs(u).max()
