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


@pytest.mark.parametrize("maker", ["cases", "oracle"])
def test_initial_spectrum_has_the_k4_gaussian_shape(maker):
    """The decaying-turbulence field's spectrum, checked against the closed form
    E(k) = C k^4 exp(-2 (k/4)^2) rather than against a second copy of the
    construction (``run/HIT/MansourWray/inputfile.py:69-90``):

    * per mode: the construction gives every interior mode (0 < k3 < N/2, off the
      Nyquist planes, k1^2 + k2^2 > 0) |u|^2 + |v|^2 + |w|^2 = |alpha|^2 + |beta|^2
      = E(|k|) / (4 pi |k|^2), whatever the random phases, so
      4 pi |k|^2 |u_k|^2 / (|k|^4 exp(-|k|^2 / 8)) is one constant over all of them;
    * per shell: the shell sums follow C k^4 exp(-2 (k/4)^2), integrated over the
      shell, up to the lattice's mode count (the count of integer points in a shell
      of width 1 departs from 4 pi k^2 by up to 15% at 3 <= k < 16), peak at k = 4,
      and total E = 3/2 (u_rms = 1)."""
    from pyspectral_amd import cases

    N = 48
    uh = cases.decaying_isotropic(N, seed=5) if maker == "cases" else np.stack(so.random_isotropic(N, seed=5))
    k1 = np.fft.fftfreq(N, 1.0 / N)
    K1, K2, K3 = np.meshgrid(k1, k1, np.arange(N // 2 + 1.0), indexing="ij")
    ksq = K1 ** 2 + K2 ** 2 + K3 ** 2
    e = (np.abs(uh) ** 2).sum(axis=0)
    shape = ksq ** 2 * np.exp(-ksq / 8.0)
    per_mode = shape / np.maximum(ksq, 1)
    # modes whose amplitude is below 1e-4 of the peak's carry the round-off of the
    # construction's Hermitian round trip (1e-16 of the peak) above 1e-12
    interior = ((K3 > 0) & (K3 < N // 2) & (np.abs(K1) < N // 2) & (np.abs(K2) < N // 2)
                & (K1 ** 2 + K2 ** 2 > 0) & (per_mode > 1e-8 * per_mode.max()))
    assert interior.sum() > 4000
    ratio = 4 * np.pi * ksq[interior] * e[interior] / shape[interior]
    assert ratio.max() / ratio.min() - 1 < 1e-10
    # a plausible slip changes the ratio's spread by orders of magnitude
    for wrong in (ksq ** 2 * np.exp(-ksq / 4.0), ksq * np.exp(-ksq / 8.0), shape / np.maximum(ksq, 1)):
        r = 4 * np.pi * ksq[interior] * e[interior] / wrong[interior]
        assert r.max() / r.min() - 1 > 0.5
    # shell sums (modes with 0 < k3 < N/2 count twice in the half spectrum)
    w = np.where((K3 > 0) & (K3 < N // 2), 2.0, 1.0)
    shell = np.rint(np.sqrt(ksq)).astype(int)
    Ek = np.bincount(shell.ravel(), weights=(0.5 * w * e).ravel()) / N ** 3
    assert Ek.sum() == pytest.approx(1.5, rel=1e-12)
    assert int(np.argmax(Ek)) == 4
    # E(k) integrated over each shell [k - 1/2, k + 1/2), as the binning does
    k = np.arange(3, N // 3)
    q = np.linspace(-0.5, 0.5, 2001)
    kk = k[:, None] + q[None, :]
    model = np.trapezoid(kk ** 4 * np.exp(-2 * (kk / 4.0) ** 2), q, axis=1)
    C = Ek[k].sum() / model.sum()
    dev = np.abs(Ek[k] / (C * model) - 1)
    assert dev.max() < 0.2


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
