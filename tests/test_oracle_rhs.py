"""Pin the oracle's plain (DNS) and Smagorinsky right-hand sides, stepper, CFL dt and
diagnostics against closed forms, invariants and brute force."""
import numpy as np
import pytest

from oracle import spectral_oracle as so
from tests import _bruteforce as bf

FFT = so.FFTBackend()


def _state(uh, model=0, nu=0.01, N=8, **kw):
    g = so.Grid(N, N, N, **{k: kw.pop(k) for k in list(kw) if k in ("dx", "dy", "dz", "kc")})
    return so.State(g, uh[0], uh[1], uh[2], nu=nu, turb_model=model, **kw)


def _rand_div_free(N, seed):
    return so.random_isotropic(N, seed=seed, kp=2.0)


def test_products_are_truncated_convolutions():
    """The six 3/2-padded products equal brute-force truncated convolutions."""
    N = 8
    uh = _rand_div_free(N, 11)
    st = _state(uh, N=N)
    got = so._products_32(st, FFT, *uh)
    C = [bf.half_to_full(a, N) for a in uh]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    keep = list(range(-(N // 2 - 1), N // 2))
    s = np.sqrt(N ** 3)
    for (i, j), g in zip(pairs, got):
        full = bf.conv(C[i], C[j])
        expect = bf.full_to_half(full, N, N, N // 2 + 1, keep, N // 2 - 1, s)
        assert np.abs(g - expect).max() < 1e-13 * np.abs(expect).max()


def test_tgv_rhs_closed_form():
    """TGV u=(sin x cos y cos z, -cos x sin y cos z, 0): RHS = -3 nu u +
    (-1/8 sin2x cos2z, -1/8 sin2y cos2z, 1/8 (cos2x+cos2y) sin2z)."""
    N, nu = 16, 1.0 / 400
    u, v, w, _ = so.taylor_green(N, N, N)
    st = _state((u, v, w), nu=nu, N=N)
    so.rhs_nosgs(st, FFT)
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    fu = -3 * nu * np.sin(X) * np.cos(Y) * np.cos(Z) - np.sin(2 * X) * np.cos(2 * Z) / 8
    fv = 3 * nu * np.cos(X) * np.sin(Y) * np.cos(Z) - np.sin(2 * Y) * np.cos(2 * Z) / 8
    fw = (np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z) / 8
    ref = np.stack([np.fft.rfftn(f) / np.sqrt(N ** 3) for f in (fu, fv, fw)])
    assert np.linalg.norm(st.compact() - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("model", [0, 1])
def test_projection_and_nyquist(model):
    """DNS: k.RHS = 0 and RHS = 0 on the Nyquist entries.  Smagorinsky: the NS part
    is solenoidal; Nyquist entries see only -nu k^2 uhat."""
    N = 8
    uh = list(_rand_div_free(N, 5))
    rng = np.random.default_rng(9)
    for a in uh:  # give the Nyquist entries nonzero values the solver must handle
        a[N // 2, :, :] = rng.standard_normal((N, N // 2 + 1))
        a[:, N // 2, :] = rng.standard_normal((N, N // 2 + 1))
    st = _state(uh, model=model, N=N, nu=0.02)
    so.compute_rhs(st, FFT)
    R = st.compact()
    g = st.grid
    nyq = np.zeros(g.k1.shape, bool)
    nyq[N // 2] = nyq[:, N // 2] = nyq[:, :, N // 2] = True
    if model == 0:
        assert np.all(R[:, nyq] == 0)
        div = g.k1 * R[0] + g.k2 * R[1] + g.k3 * R[2]
        assert np.abs(div).max() < 1e-13 * np.abs(R).max() * N
    else:
        for c in range(3):
            expect = -st.nu * g.ksqr * uh[c]
            assert np.allclose(R[c][nyq], expect[nyq], rtol=1e-14, atol=0)


def test_nonlinear_term_conserves_energy():
    """Galerkin energy conservation: sum_k Re(conj(u).RHS) = 0 when nu = 0."""
    N = 8
    uh = _rand_div_free(N, 7)
    st = _state(uh, N=N, nu=0.0)
    so.rhs_nosgs(st, FFT)
    R = st.compact()
    dE = sum(np.real(so._half_energy(R[c], N // 2) * 0) for c in range(3))  # shape check
    dE = 0.0
    for c in range(3):
        a, r = uh[c], R[c]
        dE += np.real(np.sum(r[:, :, 1:N // 2] * np.conj(a[:, :, 1:N // 2]) * 2)
                      + np.sum(r[:, :, 0] * np.conj(a[:, :, 0])))
    scale = np.linalg.norm(R) * np.linalg.norm(np.stack(uh))
    assert abs(dE) < 1e-14 * scale


def test_stepper_linear_amplification():
    """Kolmogorov mode u=(A sin 2y,0,0) has no nonlinear term: after n steps
    uhat_n = R(-nu k^2 dt)^n uhat_0, R(z) = 1+z+z^2/2+z^3/6+z^4/24.  For the
    low-storage form Q = Q0 + c_i dt RHS this fixes c = (1/4, 1/3, 1/2, 1)."""
    N, nu, dt, A = 8, 0.05, 0.1, 0.7
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = A * np.sin(2 * Y)
    s = np.sqrt(N ** 3)
    uh = [np.fft.rfftn(u) / s, np.zeros((N, N, N // 2 + 1), complex), np.zeros((N, N, N // 2 + 1), complex)]
    st = _state(uh, N=N, nu=nu, dt=dt)
    for _ in range(5):
        so.advance_rk4(st, FFT)
    z = -nu * 4 * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.allclose(st.compact()[0], uh[0] * R ** 5, rtol=0, atol=1e-15)


def test_compute_dt_closed_form():
    """CFL dt from max(|u|/dx+|v|/dy+|w|/dz) of the physical field (brute-force synthesis)."""
    N = 8
    uh = _rand_div_free(N, 3)
    cfl, nu = 0.5, 0.01
    st = _state(uh, N=N, nu=nu, cfl=cfl)
    so.compute_dt(st)
    x = np.arange(N) * 2 * np.pi / N
    pts = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1)
    vel = [bf.synth(bf.half_to_full(a, N), pts) for a in uh]
    dx = 2 * np.pi / N
    mv = np.max(sum(np.abs(v) / dx for v in vel))
    dt = cfl / (mv + 1e-10)
    dt = min(dt, 0.634 / (3 / dx ** 2) / nu * cfl / 1.35)
    assert abs(st.dt - dt) < 1e-13 * dt
    st2 = _state(uh, N=N, nu=nu, dt=0.003)
    so.compute_dt(st2)
    assert st2.dt == 0.003


def test_stats_parseval():
    """Energy and enstrophy equal physical-space means; the shell spectrum sums to 2E."""
    N = 8
    uh = _rand_div_free(N, 4)
    st = _state(uh, N=N)
    x = np.arange(N) * 2 * np.pi / N
    pts = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1)
    C = [bf.half_to_full(a, N) for a in uh]
    vel = [bf.synth(c, pts) for c in C]
    E = 0.5 * np.mean(sum(v ** 2 for v in vel))
    assert abs(so.compute_energy(st) - E) < 1e-13 * E
    K = N // 2 - 1
    ks = np.arange(-K, K + 1)
    KX, KY, KZ = np.meshgrid(ks, ks, ks, indexing="ij")
    d = lambda c, k: 1j * k * c  # noqa: E731
    om = [d(C[2], KY) - d(C[1], KZ), d(C[0], KZ) - d(C[2], KX), d(C[1], KX) - d(C[0], KY)]
    Om = 0.5 * np.mean(sum(bf.synth(o, pts) ** 2 for o in om))
    assert abs(so.compute_enstrophy(st) - Om) < 1e-12 * Om
    k, spec = so.compute_spectrum(st)
    assert abs(0.5 * spec.sum() - E) < 1e-13 * E
    # a single mode lands in bin rint(|k|)
    one = np.zeros((N, N, N // 2 + 1), complex)
    one[1, 2, 2] = 1.0
    st1 = _state((one, 0 * one, 0 * one), N=N)
    k, spec = so.compute_spectrum(st1)
    assert np.argmax(spec[:, 0]) == 3 and np.count_nonzero(spec[:, 0]) == 1


def test_smagorinsky_shear_closed_form():
    """u=(A sin my,0,0): on the padded grid tau_12 = -(Cs Delta)^2 |Am cos my| Am cos my
    pointwise, the other stresses vanish and w0 = (-i k2 tau_12, 0, 0)."""
    N, A, m = 16, 0.7, 2
    M = 3 * N // 2
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    s = np.sqrt(N ** 3)
    z = np.zeros((N, N, N // 2 + 1), complex)
    uh = (np.fft.rfftn(A * np.sin(m * Y)) / s, z, z.copy())
    st = _state(uh, model=1, N=N)
    tau, Sr, nut = so.smag_stress_hat(st, FFT, *uh, return_real=True)
    alpha2 = 1.5 ** -3 * (N ** 3) ** -0.5
    xm = np.arange(M) * 2 * np.pi / M
    _, Ym, _ = np.meshgrid(xm, xm, xm, indexing="ij")
    CsD = so.SMAG_CS * st.grid.Delta
    t12 = -2.0 * nut * Sr[3] / alpha2
    expect = -CsD ** 2 * np.abs(A * m * np.cos(m * Ym)) * A * m * np.cos(m * Ym)
    assert np.abs(t12 - expect).max() < 1e-13 * np.abs(expect).max()
    for i in (0, 1, 2, 4, 5):
        assert np.abs(tau[i]).max() < 1e-14 * np.abs(tau[3]).max()
    wu, wv, ww = so.sgs_divergence(st.grid, tau)
    assert np.allclose(wu, -1j * st.grid.k2 * tau[3], atol=1e-15)
    assert np.abs(wv).max() < 1e-14 and np.abs(ww).max() < 1e-14


def test_smagorinsky_stress_spectra_by_explicit_sums():
    """The Smagorinsky stress spectra of a random field (N = 8) equal explicit sums:
    the strain synthesised on the 3/2 grid (M = 12) from its Fourier coefficients
    (Nyquist dropped), -2 (Cs Delta)^2 |S| S_ij formed pointwise, analysed by a direct
    DFT over the M^3 points and kept at |k1|, |k2| < N/2, 0 <= k3 < N/2.  This pins the
    aliasing the cubic stress picks up on the 3/2 grid, the truncation and the scaled
    units ((3/2)^(-3/2) Nt^(-1/4) times physical for the strain, squared for the
    stress) without the oracle's pad/unpad or FFTs."""
    N = 8
    M = 3 * N // 2
    uh = [a.copy() for a in _rand_div_free(N, 23)]
    rng = np.random.default_rng(5)
    for a in uh:  # Nyquist entries, which the 3/2 pad must drop
        a[N // 2] += 0.3 * rng.standard_normal((N, N // 2 + 1))
        a[:, N // 2] += 0.3 * rng.standard_normal((N, N // 2 + 1))
        a[:, :, N // 2] += 0.3 * rng.standard_normal((N, N))
    st = _state(uh, model=1, N=N)
    tau = so.smag_stress_hat(st, FFT, *uh)
    C = [bf.half_to_full(a, N) for a in uh]  # |k| <= N/2 - 1: the Nyquist modes dropped
    K = N // 2 - 1
    ks = np.arange(-K, K + 1)
    kk = np.meshgrid(ks, ks, ks, indexing="ij")
    dC = [[1j * kk[j] * C[i] for j in range(3)] for i in range(3)]  # coefficients of du_i/dx_j
    xm = np.arange(M) * 2 * np.pi / M
    pts = np.stack(np.meshgrid(xm, xm, xm, indexing="ij"), axis=-1)
    g = [[bf.synth(dC[i][j], pts) for j in range(3)] for i in range(3)]
    S = [[0.5 * (g[i][j] + g[j][i]) for j in range(3)] for i in range(3)]
    Smag = np.sqrt(2.0 * sum(S[i][j] ** 2 for i in range(3) for j in range(3)))
    CsD2 = (so.SMAG_CS * st.grid.Delta) ** 2
    Nt = float(N) ** 3
    alpha = 1.5 ** -1.5 * Nt ** -0.25
    E = np.exp(-1j * np.outer(xm, ks))            # analysis, k1 and k2 signed
    Ez = np.exp(-1j * np.outer(xm, np.arange(0, K + 1)))
    for (i, j), t in zip([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)], tau):
        phys = -2.0 * CsD2 * Smag * S[i][j]
        T = np.einsum("xyz,xa,yb,zc->abc", phys, E, E, Ez)
        expect = np.zeros_like(t)
        for a in range(2 * K + 1):
            for b in range(2 * K + 1):
                expect[ks[a] % N, ks[b] % N, :K + 1] = alpha ** 2 * T[a, b]
        assert np.abs(t - expect).max() < 1e-12 * np.abs(expect).max(), (i, j)
        # the Nyquist rows and plane carry nothing
        assert np.abs(t[N // 2]).max() == 0 and np.abs(t[:, N // 2]).max() == 0 and np.abs(t[:, :, N // 2]).max() == 0


def test_linear_forcing_amplification():
    """With linear forcing f = a u the Kolmogorov mode (no nonlinear term) grows or
    decays as R((a - nu k^2) dt)^n exactly, and for every model with forcing
    RHS(force) - RHS(0) = force * u."""
    N, nu, dt, a, A = 8, 0.05, 0.1, 0.3, 0.7
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    s = np.sqrt(N ** 3)
    uh = [np.fft.rfftn(A * np.sin(2 * Y)) / s, np.zeros((N, N, N // 2 + 1), complex),
          np.zeros((N, N, N // 2 + 1), complex)]
    st = _state(uh, N=N, nu=nu, dt=dt, force=a)
    for _ in range(4):
        so.advance_rk4(st, FFT)
    z = (a - nu * 4) * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.abs(st.compact()[0] - uh[0] * R ** 4).max() < 1e-14 * np.abs(uh[0]).max()
    u = _rand_div_free(N, 21)
    for model in (0, 1, 2, "Static tau-model", "Dynamic tau-model", 7):
        s0 = _state(u, model=model, N=N)
        s1 = _state(u, model=model, N=N, force=a)
        so.compute_rhs(s0, FFT)
        so.compute_rhs(s1, FFT)
        expect = a * np.stack([s0.uhat, s0.vhat, s0.what])
        assert np.abs(s1.compact() - s0.compact() - expect).max() < 1e-13


def test_coriolis_terms_are_the_projected_rotation_force():
    """With rotate = 1 the RHS gains the Leray projection of 2 u x Omega (the Coriolis
    force with the pressure that keeps it solenoidal), evaluated on the velocity as the
    model uses it, and it does no work on any mode: Re(conj(u_k) . delta_k) = 0.  At
    t = 0 (no t-model memory term) this holds for every model except the dynamic
    tau-model, which the GPU build does not run with rotation."""
    N = 8
    Om = (0.3, -0.2, 0.5)
    uh = _rand_div_free(N, 23)
    for model in (0, 1, 2, 7):
        s0 = _state(uh, model=model, N=N)
        s1 = _state(uh, model=model, N=N, rotate=1, Om=Om)
        so.compute_rhs(s0, FFT)
        so.compute_rhs(s1, FFT)
        d = s1.compact() - s0.compact()
        g = s0.grid
        u = np.stack([s0.uhat, s0.vhat, s0.what])
        f = 2.0 * np.cross(np.moveaxis(u, 0, -1), np.asarray(Om))   # 2 u x Omega per mode
        f = np.moveaxis(f, -1, 0)
        k = np.stack([g.k1, g.k2, g.k3])
        proj = f - k * (k * f).sum(0) / g.ksqr
        assert np.abs(d - proj).max() < 1e-13 * np.abs(f).max(), model
        assert np.abs((np.conj(u) * d).sum(0).real).max() < 1e-13 * np.abs(f).max() ** 2
