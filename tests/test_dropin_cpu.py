"""Host-side pieces of the drop-in that need no GPU: grid arrays, case setup,
error paths."""
import numpy as np
import pytest

from oracle import spectral_oracle as so


def test_gridclass_arrays_match_oracle():
    from pyspectral_amd.Classes import gridclass

    N = 12
    d = 2 * np.pi / (N - 1)
    x = np.linspace(0, 2 * np.pi - d, N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    g = gridclass(N, N, N, X, Y, Z, 4)
    d = x[1] - x[0]
    o = so.Grid(N, N, N, dx=d, dy=d, dz=d, kc=4)
    for name in ("k1", "k2", "k3", "ksqr", "ksqr_i", "Gf"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    assert g.dx == pytest.approx(d) and g.Delta == pytest.approx(o.Delta)


def test_cases_match_oracle_setup():
    from pyspectral_amd import cases

    for grid in ("exact", "reference"):
        uh, (dx, _, _) = cases.taylor_green(16, grid=grid)
        u, v, w, (odx, _, _) = so.taylor_green(16, 16, 16, grid=grid)
        assert dx == odx
        assert np.allclose(uh, np.stack([u, v, w]), rtol=0, atol=1e-15)
    u = cases.decaying_isotropic(16, seed=3)
    o = np.stack(so.random_isotropic(16, seed=3))
    assert np.allclose(u, o, rtol=1e-12, atol=1e-14)
    g = so.Grid(16, 16, 16)
    st = so.State(g, *u, nu=0.01)
    assert so.compute_energy(st) == pytest.approx(1.5, rel=1e-12)
    div = g.k1 * u[0] + g.k2 * u[1] + g.k3 * u[2]
    assert np.abs(div).max() < 1e-13


def test_solver_refuses_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from pyspectral_amd._lib import PsxError
    from pyspectral_amd.solver import SpectralSolver

    with pytest.raises(PsxError):
        SpectralSolver(16, 16, 16, 0, nu=0.01)


def test_unsupported_model_is_refused():
    torch = pytest.importorskip("torch")
    from pyspectral_amd import Classes

    N = 8
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    g = Classes.gridclass(N, N, N, X, Y, Z, 4)
    z = np.zeros((N, N, N // 2 + 1), complex)
    with pytest.raises(NotImplementedError):
        Classes.variables(0, 3, 0, 0, 0, 0, g, z, z, z, 0.0, 0.01, 0.01, -10, -10, -10, -10, -10, -0.01, 3)
    del torch
