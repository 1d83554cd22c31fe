"""Pins of the oracle's grid-space fields, restriction, padded gradients and the
learned-closure RHS: analytic band-limited fields with known derivatives, and the
reduction of the learned-closure RHS to the Smagorinsky RHS for the pointwise
Smagorinsky closure."""
import numpy as np

from oracle import mlp_oracle as mo
from oracle import spectral_oracle as so


def _field(g, x, y, z):
    """A band-limited field and its gradient: modes up to 3 on each axis."""
    f = (np.sin(2 * x + 1.0) * np.cos(3 * y) + 0.5 * np.cos(x - 2 * z + 0.3) + 0.25 * np.sin(3 * z) * np.sin(y)
         + 0.1 * np.cos(x + y + z))
    fx = 2 * np.cos(2 * x + 1.0) * np.cos(3 * y) - 0.5 * np.sin(x - 2 * z + 0.3) - 0.1 * np.sin(x + y + z)
    fy = -3 * np.sin(2 * x + 1.0) * np.sin(3 * y) + 0.25 * np.sin(3 * z) * np.cos(y) - 0.1 * np.sin(x + y + z)
    fz = np.sin(x - 2 * z + 0.3) + 0.75 * np.cos(3 * z) * np.sin(y) - 0.1 * np.sin(x + y + z)
    return f, (fx, fy, fz)


def _mesh(n1, n2, n3):
    ax = [2 * np.pi * np.arange(n) / n for n in (n1, n2, n3)]
    return np.meshgrid(*ax, indexing="ij")


def _velocity(g):
    X, Y, Z = _mesh(g.N1, g.N2, g.N3)
    fields, grads = [], []
    for sh in (0.0, 0.7, 1.9):  # three components: shifted copies of the field
        f, d = _field(g, X + sh, Y - sh, Z + 2 * sh)
        fields.append(np.fft.rfftn(f) / np.sqrt(g.Nt))
        grads.append(d)
    return fields, grads


def test_physical_is_the_grid_field():
    g = so.Grid(12, 16, 8)
    X, Y, Z = _mesh(12, 16, 8)
    f, _ = _field(g, X, Y, Z)
    fhat = np.fft.rfftn(f) / np.sqrt(g.Nt)
    assert np.abs(so.physical(g, fhat) - f).max() < 1e-13


def test_gradients_are_the_analytic_derivatives():
    g = so.Grid(12, 16, 8)
    fields, grads = _velocity(g)
    G = so.gradients(g, *fields)
    for i in range(3):
        for j in range(3):
            assert np.abs(G[:, 3 * i + j] - grads[i][j].ravel()).max() < 1e-12, (i, j)


def test_padded_gradients_are_the_derivatives_on_the_padded_grid():
    g = so.Grid(12, 16, 8)
    fields, _ = _velocity(g)
    st = so.State(g, *fields, nu=0.01, turb_model=1)
    G = so.padded_gradients(st, so.FFTBackend(), *fields)
    X, Y, Z = _mesh(18, 24, 12)
    for i, sh in enumerate((0.0, 0.7, 1.9)):
        _, d = _field(g, X + sh, Y - sh, Z + 2 * sh)
        for j in range(3):
            assert np.abs(G[:, 3 * i + j] - d[j].ravel()).max() < 1e-12, (i, j)


def test_restrict_keeps_grid_values_of_a_band_limited_field():
    fine, coarse = so.Grid(24, 32, 16), so.Grid(12, 16, 8)
    Xf, Yf, Zf = _mesh(24, 32, 16)
    Xc, Yc, Zc = _mesh(12, 16, 8)
    f_fine, _ = _field(fine, Xf, Yf, Zf)
    f_coarse, _ = _field(coarse, Xc, Yc, Zc)
    r = so.restrict(np.fft.rfftn(f_fine) / np.sqrt(fine.Nt), coarse)
    assert np.abs(so.physical(coarse, r) - f_coarse).max() < 1e-12
    # modes at or beyond the coarse Nyquist are dropped
    bump = np.zeros((24, 32, 9), complex)
    bump[6, 0, 0] = bump[0, 8, 0] = bump[0, 0, 4] = 1.0
    assert np.abs(so.restrict(bump, coarse)).max() == 0.0


def test_learned_rhs_with_smagorinsky_closure_is_rhs_smag():
    g = so.Grid(12, 12, 16)
    u, v, w = _random_state(g)
    fft = so.FFTBackend()
    a = so.State(g, u, v, w, nu=0.02, turb_model=1)
    b = so.State(g, u, v, w, nu=0.02, turb_model=1)
    so.rhs_smag(a, fft)
    c2 = (so.SMAG_CS * g.Delta) ** 2
    so.rhs_learned(b, fft, lambda G: mo.smag_pointwise(G, c2))
    ra, rb = a.compact(), b.compact()
    assert np.abs(ra - rb).max() < 1e-12 * np.abs(ra).max()
    for x, y in zip(a.w0, b.w0):
        assert np.abs(x - y).max() < 1e-12 * np.abs(a.w0[0]).max()
    # a plausible slip -- the stress transposed (12 <-> 13) -- is seen
    b2 = so.State(g, u, v, w, nu=0.02, turb_model=1)
    so.rhs_learned(b2, fft, lambda G: mo.smag_pointwise(G, c2)[:, [0, 1, 2, 4, 3, 5]])
    assert np.abs(b2.compact() - ra).max() > 1e-6 * np.abs(ra).max()


def _random_state(g):
    rng = np.random.default_rng(11)
    shape = (g.N1, g.N2, g.N3)
    out = []
    for _ in range(3):
        f = rng.standard_normal(shape)
        fh = np.fft.rfftn(f) / np.sqrt(g.Nt)
        kk = np.sqrt(g.k1 ** 2 + g.k2 ** 2 + g.k3 ** 2)
        out.append(fh * np.exp(-0.15 * kk ** 2))
    return out
