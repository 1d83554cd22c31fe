"""Filtered-DNS subgrid stress (computeSGS_DNS, src/Classes.py:861-929) and the filters
on the GPU against the oracle (relative L2 within 1e-12)."""
import numpy as np
import pytest

from oracle import spectral_oracle as so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FFT = so.FFTBackend()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _field(N1, N2, N3, seed):
    rng = np.random.default_rng(seed)
    g = so.Grid(N1, N2, N3)
    shp = (3, N1, N2, N3 // 2 + 1)
    u = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * np.exp(-0.02 * g.ksqr)
    u = np.stack([np.fft.rfftn(np.fft.irfftn(c, s=(N1, N2, N3), axes=(0, 1, 2)), axes=(0, 1, 2)) for c in u])
    div = g.k1 * u[0] + g.k2 * u[1] + g.k3 * u[2]
    return np.stack([u[0] - g.k1 * div / g.ksqr, u[1] - g.k2 * div / g.ksqr, u[2] - g.k3 * div / g.ksqr])


@pytest.mark.parametrize("shape,kc", [((16, 16, 16), 5), ((32, 16, 32), 6), ((16, 16, 128), 7)])
@pytest.mark.parametrize("kind", ["sharp", "box"])
def test_sgs_dns_matches_oracle(shape, kc, kind):
    from pyspectral_amd.solver import SpectralSolver

    N1, N2, N3 = shape
    print(f"\nsgs shape={shape} kc={kc} kind={kind}", flush=True)
    u = _field(N1, N2, N3, 3 + N3)
    g = so.Grid(N1, N2, N3, kc=kc)
    delta = 3.0 * 2 * np.pi / N1
    st = so.State(g, *u, nu=0.01)
    w_ref, t_ref = so.sgs_dns(st, FFT, kind, delta)
    s = SpectralSolver(N1, N2, N3, 0, nu=0.01, kc=kc)
    w, t = s.sgs_dns(torch.from_numpy(u).cuda(), kind, delta, tau=True)
    torch.cuda.synchronize()
    ew = _rel(w.cpu().numpy(), np.stack(w_ref))
    et = _rel(t.cpu().numpy(), np.stack(t_ref))
    print(f"w0 rel {ew:.3e} tau rel {et:.3e}")
    assert et < 1e-12 and ew < 1e-12
    f = s.filter(torch.from_numpy(u).cuda(), kind, delta).cpu().numpy()
    assert _rel(f, so.filter_transfer(g, kind, delta) * u) < 1e-14
    s.close()


def test_dropin_sgs_and_filter():
    from pyspectral_amd import Classes

    N, kc = 16, 5
    u = _field(N, N, N, 9)
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    grid = Classes.gridclass(N, N, N, X, Y, Z, kc)
    main = Classes.variables(0, 0, 0, 0, 0, 0, grid, u[0], u[1], u[2], 0.0, 0.01, 0.01, -10, -10, -10, -10, -10,
                             -0.01, 3)
    ut = Classes.utilitiesClass()
    wu, wv, ww = ut.computeSGS_DNS(main, grid, None)
    og = so.Grid(N, N, N, kc=kc)
    w_ref, _ = so.sgs_dns(so.State(og, *u, nu=0.01), FFT)
    assert wu.shape == (N, N, N // 2 + 1, 1)
    assert _rel(np.stack([wu[..., 0], wv[..., 0], ww[..., 0]]), np.stack(w_ref)) < 1e-12
    fu = ut.myFilter(u[0], 3)
    assert np.array_equal(fu, so.Grid(N, N, N, kc=3).Gf * u[0])
