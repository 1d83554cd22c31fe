"""Pin the oracle's 2x-padded t-model / tau-model family by brute-force convolution.

Every product here is a direct-summation convolution of centred coefficient
arrays, so index conventions (2x padding keeping the -N/2 rows, the p/q split,
unpad_2x) are checked against set definitions rather than array slicing."""
import numpy as np
import pytest

from oracle import spectral_oracle as so
from tests import _bruteforce as bf

FFT = so.FFTBackend()
N = 8


def _field(seed):
    return so.random_isotropic(N, seed=seed, kp=2.0)


def _bf_pl_and_plql(uh, nu, q_inside):
    """Brute-force PL (on signed k up to N-1) and PLQL (on the N grid incl. -N/2)."""
    K = N // 2 - 1
    C = [bf.half_to_full(a, N, K=K) for a in uh]
    s = np.sqrt(N ** 3)
    # products u_i u_j on their full support |k| <= N-2: centred arrays with offset 2K
    prod = {}
    for i in range(3):
        for j in range(i, 3):
            prod[(i, j)] = bf.conv(C[i], C[j])
    KP = 2 * K
    ks = np.arange(-KP, KP + 1).astype(float)
    k1, k2, k3 = np.meshgrid(ks, ks, ks, indexing="ij")
    ksq = k1 * k1 + k2 * k2 + k3 * k3 + 1e-50
    uu, vv, ww = prod[(0, 0)], prod[(1, 1)], prod[(2, 2)]
    uv, uw, vw = prod[(0, 1)], prod[(0, 2)], prod[(1, 2)]
    ph = (-k1 * k1 * uu - k2 * k2 * vv - k3 * k3 * ww - 2 * k1 * k2 * uv - 2 * k1 * k3 * uw
          - 2 * k2 * k3 * vw) / ksq
    Cp = [np.pad(c, K) for c in C]   # velocity on the same centred support
    PL = [-1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw - 1j * k1 * ph - nu * (ksq - 1e-50) * Cp[0],
          -1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw - 1j * k2 * ph - nu * (ksq - 1e-50) * Cp[1],
          -1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww - 1j * k3 * ph - nu * (ksq - 1e-50) * Cp[2]]
    inside = q_inside(k1, k2, k3)
    PLq = [np.where(inside, 0, p) for p in PL]
    # u_j * PLq_i, read back on the N grid including the -N/2 rows
    keep = list(range(-N // 2, N // 2))
    half = {}
    for i in range(3):
        for j in range(3):
            full = bf.conv(Cp[j], PLq[i])
            half[(j, i)] = bf.full_to_half(full, N, N, N // 2 + 1, keep, N // 2 - 1, s)
    PLh = [bf.full_to_half(p, N, N, N // 2 + 1, keep, N // 2 - 1, s) for p in PL]
    return PLh, half


def _plql_from(half, g):
    k1, k2, k3 = g.k1, g.k2, g.k3
    A = lambda j, i: half[(j, i)]  # noqa: E731  (u_j * PLq_i)
    pterm = 2 * g.ksqr_i * (k1 * k1 * A(0, 0) + k2 * k2 * A(1, 1) + k3 * k3 * A(2, 2)
                            + k1 * k2 * (A(0, 1) + A(1, 0)) + k1 * k3 * (A(0, 2) + A(2, 0))
                            + k2 * k3 * (A(1, 2) + A(2, 1)))
    out = []
    kk = (k1, k2, k3)
    for i in range(3):
        r = 0
        for j in range(3):
            r = r - 1j * kk[j] * A(j, i) - 1j * kk[j] * A(i, j)
        out.append(r + 1j * kk[i] * pterm)
    return out


@pytest.mark.parametrize("model", [2, "Static tau-model"])
def test_tmodel_family_bruteforce(model):
    uh = _field(21)
    nu, t = 0.013, 0.37
    g = so.Grid(N, N, N)
    st = so.State(g, *uh, nu=nu, turb_model=model, t=t, Ct=0.2, dt0=0.05, dt1=0.5)
    so.compute_rhs(st, FFT)
    PLh, half = _bf_pl_and_plql(uh, nu, lambda a, b, c: (np.abs(a) < N // 2) & (np.abs(b) < N // 2)
                                & (np.abs(c) < N // 2))
    Q = _plql_from(half, g)
    coef = 0.2 * t if model == 2 else max(0.0, 0.05 + 0.5 * t)
    R = st.compact()
    for c in range(3):
        expect = PLh[c] + coef * Q[c]
        assert np.abs(R[c] - expect).max() < 1e-12 * np.abs(expect).max()


def test_tmodel_linear_in_ct_t():
    uh = _field(22)
    g = so.Grid(N, N, N)
    out = []
    for t in (0.0, 0.5, 1.0):
        st = so.State(g, *uh, nu=0.01, turb_model=2, t=t)
        so.compute_rhs(st, FFT)
        out.append(st.compact())
    assert np.allclose(out[2] - out[0], 2 * (out[1] - out[0]), rtol=0, atol=1e-13 * np.abs(out[2]).max())
    # the -N/2 rows of Q receive a (passive) nonzero RHS, the k3 = N/2 plane none
    R = out[1]
    assert np.abs(R[:, N // 2]).max() > 0 and np.all(R[:, :, :, N // 2] == 0)


def test_dynamic_tau_reductions_bruteforce():
    """Dynamic tau: the PLQL of the filtered field with the test-filter split and the
    three Hermitian-weighted reductions, rebuilt from brute-force convolutions."""
    uh = _field(23)
    nu = 0.011
    g = so.Grid(N, N, N, kc=5)   # kf = int(2*5/3) = 3
    st = so.State(g, *uh, nu=nu, turb_model="Dynamic tau-model")
    kf = st.kf
    assert kf == 3
    so.compute_rhs(st, FFT)
    # brute-force pieces
    PLh, half = _bf_pl_and_plql(uh, nu, lambda a, b, c: (np.abs(a) < N // 2) & (np.abs(b) < N // 2)
                                & (np.abs(c) < N // 2))
    Q = _plql_from(half, g)
    uf = [so.test_filter(g, kf, a) for a in uh]
    PLf, halff = _bf_pl_and_plql(uf, nu, lambda a, b, c: (np.abs(a) < kf) & (np.abs(b) < kf)
                                 & (np.abs(c) < kf))
    Qf = _plql_from(halff, g)
    w = lambda a, b: (np.sum(a[:, :, 1:N // 2] * np.conj(b[:, :, 1:N // 2]) * 2)  # noqa: E731
                      + np.sum(a[:, :, 0] * np.conj(b[:, :, 0])))
    Nt = N ** 3
    PLQLE_filt = sum(w(so.test_filter(g, kf, Qf[c]), uf[c]) for c in range(3)) / Nt
    filt_PLQLE = sum(w(Q[c], uf[c]) for c in range(3)) / Nt
    LE, pf, fp = st.tau_parts
    assert abs(pf - PLQLE_filt) < 1e-12 * abs(PLQLE_filt)
    assert abs(fp - filt_PLQLE) < 1e-12 * abs(filt_PLQLE)
    tau = max(0.0, abs(LE) / (abs(pf) - abs(fp) + 1e-5))
    R = st.compact()
    for c in range(3):
        expect = PLh[c] + tau * Q[c]
        assert np.abs(R[c] - expect).max() < 1e-11 * np.abs(expect).max()


def _bf_leonard_nl(uh):
    """-i k_j u_i u_j + i k_i p with p = -k_m k_n u_m u_n / k^2, the form the dynamic
    tau-model's Leonard term takes as written (src/RHSfunctions.py:814-819), from
    direct convolutions; centred arrays on |k_i| <= N - 2 and their wavenumbers."""
    K = N // 2 - 1
    C = [bf.half_to_full(a, N, K=K) for a in uh]
    prod = {(i, j): bf.conv(C[i], C[j]) for i in range(3) for j in range(i, 3)}
    ks = np.arange(-2 * K, 2 * K + 1).astype(float)
    k1, k2, k3 = np.meshgrid(ks, ks, ks, indexing="ij")
    ksq = k1 * k1 + k2 * k2 + k3 * k3 + 1e-50
    uu, vv, ww = prod[(0, 0)], prod[(1, 1)], prod[(2, 2)]
    uv, uw, vw = prod[(0, 1)], prod[(0, 2)], prod[(1, 2)]
    p = (-k1 * k1 * uu - k2 * k2 * vv - k3 * k3 * ww - 2 * k1 * k2 * uv - 2 * k1 * k3 * uw
         - 2 * k2 * k3 * vw) / ksq
    nl = [-1j * k1 * uu - 1j * k2 * uv - 1j * k3 * uw + 1j * k1 * p,
          -1j * k1 * uv - 1j * k2 * vv - 1j * k3 * vw + 1j * k2 * p,
          -1j * k1 * uw - 1j * k2 * vw - 1j * k3 * ww + 1j * k3 * p]
    return nl, (k1, k2, k3)


def test_dynamic_tau_leonard_reduction_bruteforce():
    """Dynamic tau: the Leonard reduction LE = <G(N(u)) - G(N(G u)), G u> over the half
    spectrum (weight 2 for 0 < k3 < N/2), with G the sharp test filter |k_i| < kf, rebuilt
    from direct convolutions: this pins the oracle's 2x padding, products, test filter
    on the 2x grid and Hermitian weights for LE as for its two companion reductions."""
    uh = _field(29)
    g = so.Grid(N, N, N, kc=5)   # kf = 3
    st = so.State(g, *uh, nu=0.013, turb_model="Dynamic tau-model")
    kf = st.kf
    so.compute_rhs(st, FFT)
    uf = [so.test_filter(g, kf, a) for a in uh]
    nl, (k1, k2, k3) = _bf_leonard_nl(uh)
    nlf, _ = _bf_leonard_nl(uf)
    inside = (np.abs(k1) < kf) & (np.abs(k2) < kf) & (np.abs(k3) < kf)
    K = N // 2 - 1
    Cf = [np.pad(bf.half_to_full(a, N, K=K), K) for a in uf]  # same support as nl
    w = np.where(k3 == 0, 1.0, np.where((k3 > 0) & (k3 < N // 2), 2.0, 0.0))
    LE = sum(np.sum(w * np.where(inside, a - b, 0) * np.conj(c)) for a, b, c in zip(nl, nlf, Cf))
    got = st.tau_parts[0]
    assert abs(LE) > 0
    assert abs(got - LE) < 1e-12 * abs(LE), (got, LE)


def test_separate_modes_per_axis_cutoffs():
    """p holds exactly the modes with |k1| < N1/2, |k2| < N2/2 and 0 <= k3 < N3/2 of the
    2x grid, each axis at its own N (src/padding.py:65-79), stated on integer wavenumbers."""
    N1, N2, N3 = 8, 12, 16
    rng = np.random.default_rng(2)
    shp = (2 * N1, 2 * N2, N3 + 1)
    u = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    up, uq = so.separate_modes(u)
    k1 = np.fft.fftfreq(2 * N1, 1.0 / (2 * N1))[:, None, None]
    k2 = np.fft.fftfreq(2 * N2, 1.0 / (2 * N2))[None, :, None]
    k3 = np.arange(N3 + 1)[None, None, :]
    inside = (np.abs(k1) < N1 / 2) & (np.abs(k2) < N2 / 2) & (k3 < N3 / 2)
    assert np.array_equal(up != 0, inside)
    assert np.array_equal(up + uq, u)
    kf = 3
    fp, fq = so.separate_modes_testfilter(u, kf)
    inside_f = (np.abs(k1) < kf) & (np.abs(k2) < kf) & (k3 < kf)
    assert np.array_equal(fp != 0, inside_f)
