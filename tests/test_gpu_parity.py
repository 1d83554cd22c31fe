"""GPU parity: libpsx's HIP path against the oracle on the same seeded inputs.

Tolerances: single RHS evaluations agree to 1e-12 relative L2 (FFT summation
order is the only difference); the 100-step Taylor-Green run of the plain
method agrees to 1e-10 relative L2 on uhat (the stated parity bar).
"""
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


def _solver(N, model, nu, **kw):
    from pyspectral_amd.solver import SpectralSolver

    return SpectralSolver(N, N, N, model, nu=nu, **kw)


def _field(N, seed, nyquist_noise=False):
    u = np.stack(so.random_isotropic(N, seed=seed, kp=max(2.0, N / 8)))
    if nyquist_noise:
        rng = np.random.default_rng(seed + 100)
        for c in range(3):
            u[c, N // 2] += 0.1 * rng.standard_normal((N, N // 2 + 1))
            u[c, :, N // 2] += 0.1 * rng.standard_normal((N, N // 2 + 1))
            u[c, :, :, N // 2] += 0.1 * rng.standard_normal((N, N))
    return u


MODELS = [0, 1, 2, "Static tau-model", "Dynamic tau-model", 7]
TAU_KW = dict(Ct=0.2, dt0=0.05, dt1=0.5)


@pytest.mark.parametrize("N", [16, 32, 48])
@pytest.mark.parametrize("model", MODELS)
def test_rhs_matches_oracle(N, model):
    print(f"\nrhs model={model} N={N}", flush=True)
    nu, t = 0.013, 0.37
    u = _field(N, 3 + N, nyquist_noise=True)
    g = so.Grid(N, N, N)
    st = so.State(g, *u, nu=nu, turb_model=model, t=t, **TAU_KW)
    so.compute_rhs(st, FFT)
    ref = st.compact()
    s = _solver(N, model, nu, Ct=0.2, tau_a=0.05, tau_b=0.5)
    ud = _dev(u)
    um = torch.empty_like(ud)
    w0 = torch.empty_like(ud) if model != 0 else None
    r = _host(s.rhs(ud, t, uhat_model=um, w0=w0))
    err = _rel(r, ref)
    print(f"rhs relative L2 = {err:.3e}")
    assert err < 1e-12, err
    # main.uhat as the model leaves it
    exp_u = np.stack([st.uhat, st.vhat, st.what])
    assert _rel(_host(um), exp_u) == 0.0
    if model != 0:
        w_ref = np.stack(st.w0)
        assert _rel(_host(w0), w_ref) < 1e-11, _rel(_host(w0), w_ref)
    if model == 7:
        cs2 = s.tau_values()[0]
        assert abs(cs2 - st.cs2) <= 1e-11 * abs(st.cs2), (cs2, st.cs2)
    if model == "Dynamic tau-model":
        tau = s.tau_values()
        assert abs(tau[0] - st.tau) <= 1e-10 * abs(st.tau), (tau, st.tau)
        LE, pf, fp = st.tau_parts
        assert abs(complex(tau[2], tau[3]) - LE) <= 1e-10 * abs(LE)
        assert abs(complex(tau[4], tau[5]) - pf) <= 1e-10 * abs(pf)
        assert abs(complex(tau[6], tau[7]) - fp) <= 1e-10 * abs(fp)


def _field_shape(N1, N2, N3, seed):
    """Seeded smooth solenoidal Hermitian field on an N1 x N2 x N3 grid (any shape)."""
    rng = np.random.default_rng(seed)
    g = so.Grid(N1, N2, N3)
    shp = (3, N1, N2, N3 // 2 + 1)
    u = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * np.exp(-0.02 * g.ksqr)
    u = np.stack([np.fft.rfftn(np.fft.irfftn(c, s=(N1, N2, N3), axes=(0, 1, 2)), axes=(0, 1, 2)) for c in u])
    div = g.k1 * u[0] + g.k2 * u[1] + g.k3 * u[2]
    u = np.stack([u[0] - g.k1 * div / g.ksqr, u[1] - g.k2 * div / g.ksqr, u[2] - g.k3 * div / g.ksqr])
    return u


# z lengths 64 R0 (3/2 rule: 3 N3 / 2, t-model: 2 N3, CFL: N3) take the one-wavefront z pass
@pytest.mark.parametrize("shape,models", [((16, 16, 128), MODELS), ((16, 32, 64), MODELS),
                                          ((16, 16, 256), [0, 1, 2, "Static tau-model"]), ((16, 16, 512), [0])])
def test_rhs_long_z_matches_oracle(shape, models):
    _check_rhs_and_dt(shape, models)


# x lengths 64, 128 and 256: the t-model family's PL pass runs the 2x-grid x lengths 128,
# 256 and 512 with each thread's four entries in registers through the first and last
# stages (radix plans 4 8 4, 4 4 4 4 and 4 8 4 4)
@pytest.mark.parametrize("shape,models", [((64, 16, 16), [2, "Static tau-model", "Dynamic tau-model"]),
                                          ((128, 16, 32), [2, "Dynamic tau-model"]),
                                          ((256, 16, 16), [2, "Dynamic tau-model"])])
def test_rhs_long_x_tau_family_matches_oracle(shape, models):
    _check_rhs_and_dt(shape, models)


def _check_rhs_and_dt(shape, models):
    N1, N2, N3 = shape
    nu, t = 0.011, 0.29
    u = _field_shape(N1, N2, N3, 41 + N3)
    for model in models:
        print(f"\nrhs shape={shape} model={model}", flush=True)
        g = so.Grid(N1, N2, N3)
        st = so.State(g, *u, nu=nu, turb_model=model, t=t, **TAU_KW)
        so.compute_rhs(st, FFT)
        ref = st.compact()
        from pyspectral_amd.solver import SpectralSolver

        s = SpectralSolver(N1, N2, N3, model, nu=nu, Ct=0.2, tau_a=0.05, tau_b=0.5)
        w0 = torch.empty_like(_dev(u)) if model != 0 else None
        r = _host(s.rhs(_dev(u), t, w0=w0))
        err = _rel(r, ref)
        print(f"rhs relative L2 = {err:.3e}", flush=True)
        assert err < 1e-12, (model, err)
        if model != 0:
            assert _rel(_host(w0), np.stack(st.w0)) < 1e-11
        st2 = so.State(g, *u, nu=nu, cfl=0.5)
        so.compute_dt(st2)
        s2 = SpectralSolver(N1, N2, N3, 1, nu=nu, cfl=0.5)
        dt = _host(s2.compute_dt(_dev(u)))[0]
        assert abs(dt - st2.dt) < 1e-14 * st2.dt, (dt, st2.dt)
        s.close()
        s2.close()


# Without a w0 output the Smagorinsky-form y-forward folds products and stresses from
# 12 fields into 9 (D = N + tau and the k1-polynomial coefficients of the stress's
# projection term); with one it keeps the 12.  Both agree with the oracle.
@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 32, 32), (48, 48, 48), (16, 16, 128), (16, 32, 64),
                                   (32, 16, 48), (16, 16, 256)])
def test_smagorinsky_rhs_folded_matches_oracle(shape):
    N1, N2, N3 = shape
    nu, t = 0.009, 0.41
    print(f"\nfolded smagorinsky rhs shape={shape}", flush=True)
    u = _field(N1, 7 + N1, nyquist_noise=True) if N1 == N2 == N3 else _field_shape(N1, N2, N3, 53 + N3)
    st = so.State(so.Grid(N1, N2, N3), *u, nu=nu, turb_model=1, t=t)
    so.compute_rhs(st, FFT)
    ref = st.compact()
    from pyspectral_amd.solver import SpectralSolver

    s = SpectralSolver(N1, N2, N3, 1, nu=nu)
    r_fold = _host(s.rhs(_dev(u), t))
    w0 = torch.empty_like(_dev(u))
    r_full = _host(s.rhs(_dev(u), t, w0=w0))
    e_fold, e_pair = _rel(r_fold, ref), _rel(r_fold, r_full)
    print(f"folded vs oracle {e_fold:.3e}, folded vs 12-field {e_pair:.3e}", flush=True)
    assert e_fold < 1e-12 and e_pair < 1e-13, (e_fold, e_pair)
    s.close()


@pytest.mark.parametrize("model", MODELS)
def test_rhs_with_linear_forcing(model):
    N, nu, a, t = 16, 0.012, 0.1, 0.37
    u = _field(N, 31, nyquist_noise=True)
    st = so.State(so.Grid(N, N, N), *u, nu=nu, turb_model=model, force=a, t=t, **TAU_KW)
    so.compute_rhs(st, FFT)
    s = _solver(N, model, nu, force=a, Ct=0.2, tau_a=0.05, tau_b=0.5)
    r = _host(s.rhs(_dev(u), t))
    assert _rel(r, st.compact()) < 1e-12


@pytest.mark.parametrize("model", MODELS)
def test_rhs_with_rotation(model):
    """Coriolis terms (rotate = 1, src/RHSfunctions.py:33-50,653-656,760-763,875-878):
    every model's RHS with a rotation vector against the oracle, at a time where the
    t-model memory is active."""
    N, nu, t = 16, 0.012, 0.37
    Om = (0.3, -0.2, 0.5)
    u = _field(N, 37, nyquist_noise=True)
    st = so.State(so.Grid(N, N, N), *u, nu=nu, turb_model=model, t=t, rotate=1, Om=Om, **TAU_KW)
    so.compute_rhs(st, FFT)
    s = _solver(N, model, nu, Ct=0.2, tau_a=0.05, tau_b=0.5, rotate=1, Om=Om)
    r = _host(s.rhs(_dev(u), t))
    err = _rel(r, st.compact())
    print(f"\nrotation model={model}: rhs relative L2 = {err:.3e}")
    assert err < 1e-12, err
    s.close()


def test_dynamic_tau_model_with_rotation_steps_and_set_params():
    """The dynamic tau-model with rotation: its dynamic coefficient and two RK4 steps
    against the oracle, then rotation switched on by psx_set_params on a context created
    without it (src/RHSfunctions.py:653-656,760-763,875-878)."""
    N, nu, Om = 16, 0.01, (0.4, 0.1, -0.6)
    u = _field(N, 41, nyquist_noise=True)
    st = so.State(so.Grid(N, N, N), *u, nu=nu, turb_model="Dynamic tau-model", cfl=0.5, rotate=1, Om=Om,
                  **TAU_KW)
    s = _solver(N, "Dynamic tau-model", nu, cfl=0.5, Ct=0.2, tau_a=0.05, tau_b=0.5)
    s.set_params(rotate=1, Om=Om)
    q = _dev(u)
    qs = q.clone()
    t = 0.1
    for _ in range(2):
        st.t = t
        so.advance_rk4(st, FFT)
        s.step(q, qs, t)
        t += st.dt
    assert _rel(_host(q), st.compact()) < 1e-12
    assert abs(s.tau_values()[0] - st.tau) < 1e-10 * max(abs(st.tau), 1e-30)


def test_rhs_tgv_closed_form():
    N, nu = 32, 1.0 / 400
    u, v, w, _ = so.taylor_green(N, N, N)
    s = _solver(N, 0, nu)
    r = _host(s.rhs(_dev(np.stack([u, v, w])), 0.0))
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    fu = -3 * nu * np.sin(X) * np.cos(Y) * np.cos(Z) - np.sin(2 * X) * np.cos(2 * Z) / 8
    fv = 3 * nu * np.cos(X) * np.sin(Y) * np.cos(Z) - np.sin(2 * Y) * np.cos(2 * Z) / 8
    fw = (np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z) / 8
    ref = np.stack([np.fft.rfftn(f) / np.sqrt(N ** 3) for f in (fu, fv, fw)])
    assert _rel(r, ref) < 1e-13


@pytest.mark.parametrize("N", [16, 32])
def test_compute_dt_matches_oracle(N):
    u = _field(N, 7, nyquist_noise=True)
    g = so.Grid(N, N, N)
    st = so.State(g, *u, nu=0.01, cfl=0.5)
    so.compute_dt(st)
    s = _solver(N, 1, 0.01, cfl=0.5)
    dt = _host(s.compute_dt(_dev(u)))[0]
    assert abs(dt - st.dt) < 1e-14 * st.dt, (dt, st.dt)


@pytest.mark.parametrize("model,N,steps,cfl", [(0, 16, 5, 0.5), (1, 16, 5, 0.5), (0, 32, 3, -0.002),
                                                 (1, 32, 3, 0.5), (2, 16, 4, 0.5), ("Static tau-model", 16, 3, 0.5),
                                                 ("Dynamic tau-model", 16, 3, 0.5), (7, 16, 3, 0.5)])
def test_steps_match_oracle(model, N, steps, cfl):
    print(f"\nstep model={model} N={N}", flush=True)
    nu = 0.01
    u = _field(N, 11, nyquist_noise=True)
    g = so.Grid(N, N, N)
    st = so.State(g, *u, nu=nu, turb_model=model, cfl=cfl, dt=0.002, **TAU_KW)
    s = _solver(N, model, nu, cfl=cfl, Ct=0.2, tau_a=0.05, tau_b=0.5)
    q = _dev(u)
    qs = q.clone()
    t = 0.1
    for _ in range(steps):
        st.t = t
        so.advance_rk4(st, FFT)
        dt = s.step(q, qs, t)
        assert abs(_host(dt)[0] - st.dt) <= 1e-13 * st.dt
        t += st.dt
    assert _rel(_host(q), st.compact()) < 1e-12
    stale = np.stack([st.uhat, st.vhat, st.what])
    assert _rel(_host(qs), stale) < 1e-12


def test_tgv_100_steps_parity():
    """Plain method, 32^3 TGV, nu = 1/400, dt = 0.01, 100 steps: <= 1e-10 relative L2."""
    N, nu, dt = 32, 1.0 / 400, 0.01
    u, v, w, _ = so.taylor_green(N, N, N)
    g = so.Grid(N, N, N)
    st = so.State(g, u, v, w, nu=nu, dt=dt)
    s = _solver(N, 0, nu, cfl=-dt)
    q = _dev(np.stack([u, v, w]))
    qs = q.clone()
    for _ in range(100):
        so.advance_rk4(st, FFT)
        s.step(q, qs, 0.0)
    err = _rel(_host(q), st.compact())
    print(f"\nTGV 100-step relative L2 = {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("N", [16, 32])
def test_stats_and_spectrum_match_oracle(N):
    u = _field(N, 5, nyquist_noise=True)
    g = so.Grid(N, N, N, kc=N // 3)
    st = so.State(g, *u, nu=0.01)
    s = _solver(N, 0, 0.01, kc=N // 3)
    ud = _dev(u)
    E, Om, Er = _host(s.stats(ud))
    assert abs(E - so.compute_energy(st)) < 1e-13 * abs(E)
    assert abs(Om - so.compute_enstrophy(st)) < 1e-13 * abs(Om)
    assert abs(Er - so.compute_energy_resolved(st)) < 1e-13 * abs(Er)
    for res in (False, True):
        k, spec = so.compute_spectrum(st, resolved=res)
        got = _host(s.spectrum(ud, resolved=res))
        assert got.shape == spec.shape
        assert np.allclose(got, spec, rtol=1e-12, atol=1e-15 * spec.max())
