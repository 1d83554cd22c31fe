"""Error contract and settings on the GPU: non-finite states, graphs re-captured after
parameter changes, the torch custom ops, and the weave variant's missing rotation.

NaN and infinite values are written into the state as ordinary numbers: every kernel
reads and writes the same memory as for a finite field.
"""
import math

import numpy as np
import pytest

from oracle import spectral_oracle as so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FFT = so.FFTBackend()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _field(N, seed):
    return np.stack(so.random_isotropic(N, seed=seed, kp=max(2.0, N / 8)))


@pytest.mark.parametrize("model", [0, 1, "Dynamic tau-model"])
@pytest.mark.parametrize("bad", ["nan", "inf"])
def test_nonfinite_state_gives_reference_dt_and_status(model, bad):
    """A NaN or infinite Fourier coefficient makes the grid velocity NaN (an infinite one
    meets 0 and -inf in the inverse transform), so dt is NaN, as np.amax and np.minimum
    give in src/Classes.py:687-698; the non-finite status is raised and stays set until
    cleared."""
    from pyspectral_amd import _lib
    from pyspectral_amd.solver import SpectralSolver

    N = 16
    print(f"\nnonfinite model={model} {bad}", flush=True)
    u = _field(N, 5)
    s = SpectralSolver(N, N, N, model, nu=0.01, cfl=0.5)
    q = _dev(u)
    qs = q.clone()
    s.step(q, qs, 0.0)
    assert s.finite()
    qs[1, 3, 2, 4] = complex(float(bad), 0.0)  # the stale velocity the CFL limit reads
    stale = _host(qs)
    dt = _host(s.step(q, qs, 0.0))[0]
    assert math.isnan(dt)
    assert not s.finite()
    # sticky: a later call reports it without waiting
    rc = s.lib.psx_stats(s.ctx, q.data_ptr(), torch.empty(3, dtype=torch.float64, device="cuda").data_ptr())
    assert rc == _lib.PSX_E_NONFINITE
    s.clear_status()
    assert s.finite()
    # the oracle's dt for the same state
    g = so.Grid(N, N, N)
    st = so.State(g, *stale, nu=0.01, cfl=0.5)
    with np.errstate(invalid="ignore"):
        so.compute_dt(st)
    assert math.isnan(st.dt)
    s.close()


def test_stats_of_a_nonfinite_state_raise_the_status():
    from pyspectral_amd.solver import SpectralSolver

    N = 16
    u = _field(N, 6)
    u[0, 1, 1, 1] = np.nan
    s = SpectralSolver(N, N, N, 1, nu=0.01)
    out = _host(s.stats(_dev(u)))
    assert np.isnan(out[0]) and np.isnan(out[1])
    assert not s.finite()
    s.close()


def test_status_after_graph_replays():
    """The flag written by a replayed graph is readable after the replays."""
    from pyspectral_amd.solver import SpectralSolver

    N = 16
    u = _field(N, 8)
    s = SpectralSolver(N, N, N, 1, nu=0.01, cfl=0.5)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        q = _dev(u)
        qs = q.clone()
        dt = torch.zeros(1, dtype=torch.float64, device="cuda")
        t = torch.zeros(1, dtype=torch.float64, device="cuda")
        for _ in range(2):
            s.step_graph(q, qs, dt, t)
        side.synchronize()
        assert s.finite()
        qs[0, 2, 2, 2] = complex(float("nan"), 0.0)
        for _ in range(2):
            s.step_graph(q, qs, dt, t)
        side.synchronize()
        assert not s.finite()
        assert math.isnan(_host(dt)[0])
    s.close()


@pytest.mark.parametrize("model", [1, 2])
def test_graph_recaptured_after_set_params(model):
    """A captured step holds nu by value: after psx_set_params the next replay must use
    the new viscosity (compared bitwise with an eager step under the new nu)."""
    from pyspectral_amd.solver import SpectralSolver

    N = 16
    u = _field(N, 9)
    s = SpectralSolver(N, N, N, model, nu=0.01, cfl=0.5, Ct=0.2)
    e = SpectralSolver(N, N, N, model, nu=0.03, cfl=0.5, Ct=0.2)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        q = _dev(u)
        qs = q.clone()
        dt = torch.zeros(1, dtype=torch.float64, device="cuda")
        t = torch.zeros(1, dtype=torch.float64, device="cuda")
        s.step_graph(q, qs, dt, t)  # captured with nu = 0.01
        qe, qse = q.clone(), qs.clone()
        t0 = float(_host(t)[0])
        s.set_params(nu=0.03)
        s.step_graph(q, qs, dt, t)
        e.step(qe, qse, t0)
        side.synchronize()
    assert np.array_equal(_host(q), _host(qe))
    assert np.array_equal(_host(qs), _host(qse))
    s.close()
    e.close()


def test_torch_custom_ops_match_the_c_abi():
    """pyspectral::rhs, rk4_step_ and stats give what the C-ABI calls give."""
    from pyspectral_amd import ops  # noqa: F401  (registers the ops)
    from pyspectral_amd.solver import SpectralSolver

    N = 16
    u = _field(N, 10)
    s = SpectralSolver(N, N, N, 1, nu=0.01, cfl=0.5)
    x = _dev(u)
    r_op = torch.ops.pyspectral.rhs(s.id, x, 0.2)
    r_abi = s.rhs(x, 0.2)
    assert torch.equal(r_op, r_abi)
    st_op = torch.ops.pyspectral.stats(s.id, x)
    st_abi = s.stats(x)
    assert torch.equal(st_op, st_abi)
    q1, qs1 = x.clone(), x.clone()
    q2, qs2 = x.clone(), x.clone()
    d1 = torch.zeros(1, dtype=torch.float64, device="cuda")
    d2 = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.ops.pyspectral.rk4_step_(s.id, q1, qs1, d1, 0.0)
    s.step(q2, qs2, 0.0, d2)
    assert torch.equal(q1, q2) and torch.equal(qs1, qs2) and torch.equal(d1, d2)
    # and the op result is the oracle's RHS
    st = so.State(so.Grid(N, N, N), *u, nu=0.01, turb_model=1, t=0.2)
    so.compute_rhs(st, FFT)
    assert _rel(_host(r_op), st.compact()) < 1e-12
    s.close()


def test_weave_variant_ignores_rotation():
    """variables(weave=1, turb_model=0, rotate=1) runs computeRHS_NOSGS_WEAVE, whose fused
    assembly has no Coriolis terms (src/RHSfunctionsWEAVE.py:44-66): its RHS and steps
    equal the oracle's plain RHS without rotation."""
    from pyspectral_amd import PySpec_3d
    from pyspectral_amd.Classes import FFTclass, gridclass, utilitiesClass, variables

    N, nu = 16, 0.01
    u = _field(N, 12)
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    grid = gridclass(N, N, N, X, Y, Z, N // 2)
    fft = FFTclass(N, N, N)
    util = utilitiesClass()
    main = variables(1, 0, 1, 0.3, -0.4, 0.9, grid, u[0], u[1], u[2], 0.0, 0.01, nu, -10, -10, -10, -10, -10,
                     -0.01, 3, myFFT=fft)
    main.computeRHS(main, grid, fft, util)
    st = so.State(so.Grid(N, N, N), *u, nu=nu, rotate=0)
    so.compute_rhs(st, FFT)
    got = np.stack([main.Q[c::3, c::3, c::3] for c in range(3)])
    assert _rel(got, st.compact()) < 1e-12
    # the fused step of the same case also runs without rotation
    main.Qd.copy_(_dev(u))
    main.Ud.copy_(_dev(u))
    PySpec_3d.advanceQ_RK4(main, grid, fft, util)
    st = so.State(so.Grid(N, N, N), *u, nu=nu, rotate=0, dt=0.01)
    so.advance_rk4(st, FFT)
    assert _rel(_host(main.Qd), st.compact()) < 1e-12
    # the plain computeRHS_NOSGS of a rotating case does apply the Coriolis terms
    main2 = variables(0, 0, 1, 0.3, -0.4, 0.9, grid, u[0], u[1], u[2], 0.0, 0.01, nu, -10, -10, -10, -10, -10,
                      -0.01, 3, myFFT=fft)
    main2.computeRHS(main2, grid, fft, util)
    st = so.State(so.Grid(N, N, N), *u, nu=nu, rotate=1, Om=(0.3, -0.4, 0.9))
    so.compute_rhs(st, FFT)
    got = np.stack([main2.Q[c::3, c::3, c::3] for c in range(3)])
    assert _rel(got, st.compact()) < 1e-12
