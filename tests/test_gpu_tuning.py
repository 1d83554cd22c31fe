"""Kernel-shape tuning: every tile width and z-pass variant gives the oracle's RHS, and
the autotuner picks among them."""
import numpy as np
import pytest

from oracle import spectral_oracle as so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FFT = so.FFTBackend()


def _field(N1, N2, N3, seed):
    rng = np.random.default_rng(seed)
    g = so.Grid(N1, N2, N3)
    shp = (3, N1, N2, N3 // 2 + 1)
    u = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * np.exp(-0.02 * g.ksqr)
    u = np.stack([np.fft.rfftn(np.fft.irfftn(c, s=(N1, N2, N3), axes=(0, 1, 2)), axes=(0, 1, 2)) for c in u])
    div = g.k1 * u[0] + g.k2 * u[1] + g.k3 * u[2]
    return np.stack([u[0] - g.k1 * div / g.ksqr, u[1] - g.k2 * div / g.ksqr, u[2] - g.k3 * div / g.ksqr])


@pytest.mark.parametrize("model", [1, 0])
def test_every_kernel_shape_matches_oracle(model):
    from pyspectral_amd.solver import SpectralSolver

    shape = (32, 16, 128)
    u = _field(*shape, 77)
    st = so.State(so.Grid(*shape), *u, nu=0.01, turb_model=model)
    so.compute_rhs(st, FFT)
    ref = st.compact()
    s = SpectralSolver(*shape, model, nu=0.01)
    ud = torch.from_numpy(u).cuda()
    # lengths 48 (x) and 24 (y) take every width; 24 and 32 leave a partial last tile
    # of the 64 k3 columns
    for T in (4, 8, 16, 24, 32):
        for zv in (0, 1):
            for order in (0, 1, 2):
                print(f"\nmodel={model} tile={T} z={zv} order={order}", flush=True)
                s.set_tuning({"x_inverse_tile": T, "y_inverse_tile": T, "y_forward_tile": T, "z_variant": zv,
                              "x_inverse_order": order, "y_inverse_order": order, "y_forward_order": order,
                              "stress_sums_in_epilogue": order & 1})
                r = s.rhs(ud, 0.0).cpu().numpy()
                err = np.linalg.norm(r - ref) / np.linalg.norm(ref)
                assert err < 1e-12, (T, zv, order, err)
    tuned = s.autotune(ud, reps=2)
    assert set(tuned) == set(s.TUNING_KEYS)
    assert all(tuned[k] in (4, 8, 16, 24, 32) for k in s.TUNING_KEYS[:3]) and tuned["z_variant"] in (0, 1)
    r = s.rhs(ud, 0.0).cpu().numpy()
    assert np.linalg.norm(r - ref) / np.linalg.norm(ref) < 1e-12
    s.close()


@pytest.mark.parametrize("shape", [(16, 16, 32), (32, 48, 48), (48, 32, 64)])
@pytest.mark.parametrize("model", [1, 0])
def test_short_z_lengths_match_oracle(shape, model):
    """Padded z lengths 48, 72 and 96 (the small grids of the ensembles and the 64^3
    LES): the lane-group z pass (z_variant 0: 16, 24 or 32 lanes per position; 4
    positions per wavefront, 8 per three-wavefront workgroup, 2 per wavefront) and the
    block-wide one (z_variant 1) both give the oracle's RHS, and one RK4 step with
    either gives the same state to round-off.  The 24 k3 columns of the z = 48 grids
    take the 24-column pencil tiles by default."""
    from pyspectral_amd.solver import SpectralSolver

    u = _field(*shape, sum(shape) + model)
    st = so.State(so.Grid(*shape), *u, nu=0.01, turb_model=model)
    so.compute_rhs(st, FFT)
    ref = st.compact()
    s = SpectralSolver(*shape, model, nu=0.01)
    ud = torch.from_numpy(u).cuda()
    out = {}
    for zv in (0, 1):
        s.set_tuning({"z_variant": zv})
        r = s.rhs(ud, 0.0).cpu().numpy()
        err = np.linalg.norm(r - ref) / np.linalg.norm(ref)
        print(f"\n{shape} model={model} z_variant={zv}: {err:.2e}", flush=True)
        assert err < 1e-12, (zv, err)
        q, qs = ud.clone(), ud.clone()
        s.step(q, qs, 0.0)
        out[zv] = q.cpu().numpy()
    assert np.abs(out[0] - out[1]).max() <= 1e-13 * np.abs(out[1]).max()
    s.close()


@pytest.mark.parametrize("model", [1, 0, 7])
def test_chunked_pipeline_is_bitwise_identical(model):
    """Splitting the y/z passes into x-row chunks on two streams changes no arithmetic:
    RHS, steps and graph replays equal the unchunked ones bit for bit."""
    from pyspectral_amd.solver import SpectralSolver

    shape = (32, 32, 48)
    u = torch.from_numpy(_field(*shape, 5)).cuda()
    ref = SpectralSolver(*shape, model, nu=0.01, cfl=0.5)
    r0 = ref.rhs(u, 0.0)
    q0, s0 = u.clone(), u.clone()
    for _ in range(2):
        ref.step(q0, s0)
    for n, streams, order in ((2, 2, 0), (3, 2, 0), (4, 2, 0), (2, 1, 0), (4, 1, 0), (2, 2, 1), (1, 0, 1)):
        print(f"\nmodel={model} chunks={n} streams={streams} order={order}", flush=True)
        s = SpectralSolver(*shape, model, nu=0.01, cfl=0.5)
        s.set_tuning({"yz_chunks": n, "yz_streams": streams, "x_inverse_order": order, "y_inverse_order": order,
                      "y_forward_order": order})
        assert torch.equal(s.rhs(u, 0.0), r0)
        q, qs = u.clone(), u.clone()
        for _ in range(2):
            s.step(q, qs)
        assert torch.equal(q, q0) and torch.equal(qs, s0)
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            qg, sg = u.clone(), u.clone()
            dt = torch.zeros(1, dtype=torch.float64, device="cuda")
            tt = torch.zeros(1, dtype=torch.float64, device="cuda")
            for _ in range(2):
                s.step_graph(qg, sg, dt, tt)
        stream.synchronize()
        assert torch.equal(qg, q0) and torch.equal(sg, s0)
        s.close()
    ref.close()
