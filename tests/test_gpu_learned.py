"""Learned-closure path on the GPU against the oracle: grid-space fields, gradients
and spectral restriction (float64 to 1e-13, float32 gradients to float32 rounding);
the learned-closure RHS with the Smagorinsky fallback everywhere against the
Smagorinsky RHS (stresses pass through float32: 1e-6 of the subgrid term); with the
network, against the oracle's learned-closure RHS fed by the same network kernel on
the oracle's gradients (1e-6 of the subgrid term); graph replays bitwise against eager
steps; filtered-DNS training pairs against the oracle; a short training run."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
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


def _field(N1, N2, N3, seed, nyquist=True):
    """Seeded smooth field with content up to the Nyquist entries."""
    rng = np.random.default_rng(seed)
    g = so.Grid(N1, N2, N3)
    shp = (3, N1, N2, N3 // 2 + 1)
    u = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) * np.exp(-0.01 * g.ksqr)
    if not nyquist:
        u[:, N1 // 2] = 0
        u[:, :, N2 // 2] = 0
        u[..., N3 // 2] = 0
    return u


SHAPES = [(16, 16, 16), (32, 16, 48), (16, 16, 64), (16, 32, 128)]


@pytest.mark.parametrize("shape", SHAPES)
def test_physical_and_gradients_match_oracle(shape):
    from pyspectral_amd.solver import SpectralSolver

    N1, N2, N3 = shape
    print(f"\nphysical/gradients shape={shape}", flush=True)
    g = so.Grid(N1, N2, N3)
    u = _field(N1, N2, N3, 7 + N3)
    s = SpectralSolver(N1, N2, N3, 1, nu=0.01)
    # 4 fields: one chunk of three and one of one
    f4 = np.concatenate([u, u[:1] * (0.5 - 0.25j)])
    ph = _host(s.physical(_dev(f4)))
    ref = np.stack([so.physical(g, f) for f in f4])
    assert np.abs(ph - ref).max() < 1e-13 * np.abs(ref).max(), np.abs(ph - ref).max()
    gr = _host(s.gradients(_dev(u)))
    gref = so.gradients(g, *u)
    err = np.abs(gr - gref).max() / np.abs(gref).max()
    print(f"gradients max error / max = {err:.2e}")
    assert err < 1e-7
    s.close()


@pytest.mark.parametrize("fine,coarse", [((32, 32, 32), (16, 16, 16)), ((48, 32, 64), (32, 16, 32))])
def test_restrict_matches_oracle(fine, coarse):
    from pyspectral_amd.solver import SpectralSolver

    print(f"\nrestrict {fine} -> {coarse}", flush=True)
    u = _field(*fine, 3)
    s = SpectralSolver(*coarse, 0, nu=0.01)
    r = _host(s.restrict(_dev(u), fine))
    ref = np.stack([so.restrict(f, so.Grid(*coarse)) for f in u])
    assert np.abs(r - ref).max() <= 1e-15 * np.abs(ref).max()
    s.close()


def _learned_solver(shape, nu, net, bound, **kw):
    from pyspectral_amd.solver import SpectralSolver

    s = SpectralSolver(*shape, "learned", nu=nu, **kw)
    s.set_closure(net, bound)
    return s


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 32, 32), (16, 16, 128)])
def test_learned_fallback_everywhere_is_smagorinsky(shape):
    from pyspectral_amd.learned import MLP

    print(f"\nlearned (Smagorinsky everywhere) shape={shape}", flush=True)
    nu, t = 0.013, 0.1
    u = _field(*shape, 21, nyquist=False)
    g = so.Grid(*shape)
    st = so.State(g, *u, nu=nu, turb_model=1, t=t)
    so.rhs_smag(st, FFT)
    ref, w_ref = st.compact(), np.stack(st.w0)
    net = MLP(seed=1)
    s = _learned_solver(shape, nu, net, -1.0)
    ud = _dev(u)
    w0 = torch.empty_like(ud)
    r = _host(s.rhs(ud, t, w0=w0))
    scale = np.linalg.norm(w_ref)
    err_r = np.linalg.norm(r - ref) / scale
    err_w = np.linalg.norm(_host(w0) - w_ref) / scale
    print(f"rhs error / |w0| = {err_r:.2e}, w0 error / |w0| = {err_w:.2e}")
    assert err_r < 1e-6 and err_w < 1e-6
    s.close()
    net.close()


def _trained_like_net(G, seed=2):
    """A network with input statistics of G and non-trivial random weights."""
    from pyspectral_amd.learned import MLP

    net = MLP(seed=seed)
    rng = np.random.default_rng(seed)
    net.set_norm(G.mean(0), G.std(0), rng.normal(0, 0.05, 6), rng.uniform(0.05, 0.2, 6))
    p = net.params()
    p[9344:] = rng.normal(0, 0.3, 6)
    net.set_params(p)
    return net


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 32, 32), (16, 16, 128)])
@pytest.mark.parametrize("bound", [float("inf"), 1.5])
def test_learned_network_rhs_matches_oracle(shape, bound):
    print(f"\nlearned (network, bound={bound}) shape={shape}", flush=True)
    nu, t = 0.013, 0.1
    u = _field(*shape, 31, nyquist=False)
    g = so.Grid(*shape)
    st = so.State(g, *u, nu=nu, turb_model=1, t=t)
    G = so.padded_gradients(st, FFT, *u)
    net = _trained_like_net(G)
    c2 = (0.16 * g.Delta) ** 2
    G32 = G.astype(np.float32)
    tau_net = _host(net.forward(_dev(G32))).astype(np.float64)
    use = np.abs(mo.normalise_inputs(G32, net.norm)).max(axis=1) <= bound
    tau = np.where(use[:, None], tau_net, mo.smag_pointwise(G32, c2).astype(np.float32).astype(np.float64))
    so.rhs_learned(st, FFT, lambda _G: tau)
    ref, w_ref = st.compact(), np.stack(st.w0)
    print(f"network at {use.mean():.3f} of the points")
    s = _learned_solver(shape, nu, net, bound)
    ud = _dev(u)
    w0 = torch.empty_like(ud)
    r = _host(s.rhs(ud, t, w0=w0))
    scale = np.linalg.norm(w_ref)
    err_r = np.linalg.norm(r - ref) / scale
    err_w = np.linalg.norm(_host(w0) - w_ref) / scale
    print(f"rhs error / |w0| = {err_r:.2e}, w0 error / |w0| = {err_w:.2e}")
    assert err_r < 1e-6 and err_w < 1e-6
    # and the network stress differs from the Smagorinsky one (the test sees the network)
    if bound == float("inf"):
        st2 = so.State(g, *u, nu=nu, turb_model=1, t=t)
        so.rhs_smag(st2, FFT)
        assert np.linalg.norm(np.stack(st2.w0) - w_ref) > 1e-3 * scale
    s.close()
    net.close()


@pytest.mark.parametrize("clip", [False, True])
def test_closure_kernel_matches_oracle(clip):
    """The closure as the LES evaluates it (network, range guard, backscatter clip)
    against the oracle's: decisions agree wherever the oracle's is clear, network points
    to the BF16 emulation's accuracy, Smagorinsky points to float32 rounding."""
    rng = np.random.default_rng(17)
    n = 20000
    G = rng.normal(0, 1.0, (n, 9)) * rng.uniform(0.2, 2.0, 9)
    net = _trained_like_net(G, seed=5)
    c2, bound = 0.03, 2.5
    G32 = G.astype(np.float32)
    y = _host(net.closure(_dev(G32), c2, bound, clip))
    P = mo.unpack(net.params())
    ref, use = mo.closure(P, net.norm, bound, c2, clip)(G)
    tn = mo.forward(P, G32, net.norm)[-1]
    sm = mo.smag_pointwise(G32, c2)
    scale = np.abs(tn).max()
    clear = np.ones(n, bool)
    if clip:
        Pi = mo.backscatter(tn, G32)
        clear = np.abs(Pi) > 0.02 * np.abs(tn).sum(1) * np.abs(G32).sum(1)
    print(f"\nnetwork at {use.mean():.3f} of the points, {clear.mean():.3f} clear")
    a, b = use & clear, ~use & clear
    assert a.sum() > 100 and b.sum() > 100
    assert np.abs(y[a] - tn[a]).max() < 2e-2 * scale
    assert np.abs(y[b] - sm[b]).max() < 1e-5 * np.abs(sm[b]).max()
    net.close()


def test_learned_rhs_with_clip_matches_oracle():
    shape = (16, 16, 16)
    nu, t = 0.013, 0.1
    u = _field(*shape, 33, nyquist=False)
    g = so.Grid(*shape)
    st = so.State(g, *u, nu=nu, turb_model=1, t=t)
    G = so.padded_gradients(st, FFT, *u)
    net = _trained_like_net(G)
    c2 = (0.16 * g.Delta) ** 2
    tau = _host(net.closure(_dev(G.astype(np.float32)), c2, 3.0, True)).astype(np.float64)
    so.rhs_learned(st, FFT, lambda _G: tau)
    ref, w_ref = st.compact(), np.stack(st.w0)
    from pyspectral_amd.solver import SpectralSolver

    s = SpectralSolver(*shape, "learned", nu=nu)
    s.set_closure(net, 3.0, clip_backscatter=True)
    ud = _dev(u)
    w0 = torch.empty_like(ud)
    r = _host(s.rhs(ud, t, w0=w0))
    scale = np.linalg.norm(w_ref)
    err = np.linalg.norm(r - ref) / scale
    print(f"\nrhs error / |w0| = {err:.2e}")
    assert err < 1e-6 and np.linalg.norm(_host(w0) - w_ref) / scale < 1e-6
    s.close()
    net.close()


@pytest.mark.parametrize("shape", [(16, 16, 128), (16, 32, 256), (16, 16, 32), (16, 32, 48), (32, 16, 64)])
@pytest.mark.parametrize("bound,clip", [(float("inf"), False), (1.5, True)])
def test_fused_closure_pass_matches_separate_network_kernel(shape, bound, clip):
    """Padded z lengths 192 and 384 (one wavefront per position) and 48, 72 and 96 (4,
    2 and 2 positions per wavefront) run the network inside the z pass; the separate
    network kernel between a gradient and a products pass gives the same RHS (the two
    pair their pencils differently, so float casts of the gradients may differ in the
    last bit at a few points)."""
    from pyspectral_amd.solver import SpectralSolver

    print(f"\nfused vs separate closure shape={shape} bound={bound} clip={clip}", flush=True)
    u = _field(*shape, 41, nyquist=False)
    st = so.State(so.Grid(*shape), *u, nu=0.01, turb_model=1)
    net = _trained_like_net(so.padded_gradients(st, FFT, *u))
    out = []
    for separate in (0, 1):
        s = SpectralSolver(*shape, "learned", nu=0.01)
        s.set_closure(net, bound, clip_backscatter=clip)
        s.set_tuning({"closure_separate": separate})
        ud = _dev(u)
        w0 = torch.empty_like(ud)
        out.append((_host(s.rhs(ud, 0.0, w0=w0)), _host(w0)))
        s.close()
    (r0, w0a), (r1, w0b) = out
    scale = np.linalg.norm(w0b)
    err = max(np.linalg.norm(r0 - r1), np.linalg.norm(w0a - w0b)) / scale
    print(f"fused vs separate: error / |w0| = {err:.2e}")
    assert err < 1e-6
    net.close()


@pytest.mark.parametrize("shape,separate", [((16, 16, 128), 0), ((16, 16, 128), 1), ((32, 16, 48), 1),
                                            ((32, 16, 48), 0)])
def test_learned_rhs_folded_matches_twelve_fields(shape, separate):
    """Without a w0 output the learned closure's y-forward folds its 12 fields into 9;
    the RHS equals the 12-field one (which the oracle pins) to rounding."""
    from pyspectral_amd.solver import SpectralSolver

    print(f"\nfolded learned rhs shape={shape} separate={separate}", flush=True)
    u = _field(*shape, 43, nyquist=False)
    st = so.State(so.Grid(*shape), *u, nu=0.01, turb_model=1)
    net = _trained_like_net(so.padded_gradients(st, FFT, *u))
    s = SpectralSolver(*shape, "learned", nu=0.01)
    s.set_closure(net, 1.5, clip_backscatter=True)
    s.set_tuning({"closure_separate": separate})
    ud = _dev(u)
    r_fold = _host(s.rhs(ud, 0.0))
    r_full = _host(s.rhs(ud, 0.0, w0=torch.empty_like(ud)))
    err = _rel(r_fold, r_full)
    print(f"folded vs 12-field: {err:.2e}")
    assert err < 1e-13
    s.close()
    net.close()


def test_learned_graph_replays_match_eager():
    from pyspectral_amd import cases

    N = 16
    uh = cases.decaying_isotropic(N, seed=5, k0=2.0)
    g = so.Grid(N, N, N)
    st = so.State(g, *uh, nu=0.01, turb_model=1)
    net = _trained_like_net(so.padded_gradients(st, FFT, *uh))
    kw = dict(cfl=0.5)
    eager = _learned_solver((N, N, N), 0.01, net, 2.0, **kw)
    graphed = _learned_solver((N, N, N), 0.01, net, 2.0, **kw)
    qe = torch.from_numpy(uh).cuda()
    se = qe.clone()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        qg = torch.from_numpy(uh).cuda()
        sg = qg.clone()
        dtg = torch.zeros(1, dtype=torch.float64, device="cuda")
        tg = torch.zeros(1, dtype=torch.float64, device="cuda")
    stream.synchronize()
    for step in range(4):
        eager.step(qe, se, 0.0)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            graphed.step_graph(qg, sg, dtg, tg)
        stream.synchronize()
        _ = (qe * 2).sum().item()
        assert torch.equal(qg, qe), f"state differs after replay {step}"
        assert torch.equal(sg, se), f"stale velocity differs after replay {step}"
    assert torch.isfinite(torch.view_as_real(qe)).all()
    eager.close()
    graphed.close()
    net.close()


def test_training_pairs_match_oracle():
    from pyspectral_amd.learned import training_pairs
    from pyspectral_amd.solver import SpectralSolver

    Nd, Nl = 32, 16
    print(f"\ntraining pairs {Nd}^3 -> {Nl}^3", flush=True)
    u = np.stack(so.random_isotropic(Nd, seed=9, kp=3.0))
    dns = SpectralSolver(Nd, Nd, Nd, 0, nu=0.01)
    les = SpectralSolver(Nl, Nl, Nl, "learned", nu=0.01)
    delta = 2 * (2 * np.pi / Nl)
    x, t = training_pairs(dns, _dev(u), les, delta)
    gd, gl = so.Grid(Nd, Nd, Nd), so.Grid(Nl, Nl, Nl)
    Gb = so.box_transfer(gd, delta)
    st = so.State(gd, *u, nu=0.01)
    _, tau = so.sgs_dns(st, FFT, "box", delta)
    ub = [so.restrict(Gb * f, gl) for f in u]
    xr = so.gradients(gl, *ub)
    tr = np.stack([so.physical(gl, so.restrict(f, gl)).ravel() for f in tau], axis=1)
    xe = np.abs(_host(x) - xr).max() / np.abs(xr).max()
    te = np.abs(_host(t) - tr).max() / np.abs(tr).max()
    print(f"inputs error {xe:.2e}, targets error {te:.2e}")
    assert xe < 1e-6 and te < 1e-6
    dns.close()
    les.close()


def test_fit_learns_the_subgrid_stress():
    from pyspectral_amd.learned import MLP, fit, training_pairs
    from pyspectral_amd.solver import SpectralSolver

    Nd, Nl = 64, 32
    u = np.stack(so.random_isotropic(Nd, seed=13, kp=6.0))
    dns = SpectralSolver(Nd, Nd, Nd, 0, nu=0.01)
    les = SpectralSolver(Nl, Nl, Nl, "learned", nu=0.01)
    delta = 2 * (2 * np.pi / Nl)
    x, t = training_pairs(dns, _dev(u), les, delta)
    net = MLP(seed=4)
    losses = fit(net, x, t, steps=400, batch=4096, lr=3e-3)
    first, last = np.mean(losses[:10]), np.mean(losses[-20:])
    y = _host(net.forward(x))
    tt = _host(t).astype(np.float64)
    X = _host(x).astype(np.float64)
    sm = mo.smag_pointwise(X, (0.16 * delta) ** 2)

    def corr(a, b):
        a, b = a - a.mean(), b - b.mean()
        return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))

    c_net = [corr(y[:, c], tt[:, c]) for c in range(6)]
    c_sm = [corr(sm[:, c], tt[:, c]) for c in range(3, 6)]
    print(f"\nloss {first:.3f} -> {last:.3f}; correlation network {np.round(c_net, 3)}, "
          f"Smagorinsky (off-diagonal) {np.round(c_sm, 3)}")
    assert last < 0.5 * first
    assert min(c_net[3:]) > max(c_sm) and min(c_net[3:]) > 0.7
    dns.close()
    les.close()
    net.close()


def test_dissipation_calibration_matches_the_targets():
    """Targets with a known positive drain (1.7 x the Smagorinsky stress of the inputs):
    after calibration the clipped closure's mean subgrid dissipation equals theirs."""
    from pyspectral_amd.learned import MLP, calibrate_dissipation, fit, subgrid_dissipation
    from pyspectral_amd.solver import SpectralSolver

    N = 32
    les = SpectralSolver(N, N, N, "learned", nu=0.01)
    x = les.gradients(_dev(np.stack(so.random_isotropic(N, seed=4, kp=4.0))))
    X = _host(x).astype(np.float64)
    c2 = (0.16 * 2 * np.pi / N) ** 2
    t = _dev((1.7 * mo.smag_pointwise(X, c2)).astype(np.float32))
    net = MLP(seed=8)
    fit(net, x, t, steps=150, batch=4096, lr=3e-3)
    a = calibrate_dissipation(net, x, t, c2, 50.0, True)
    d_cl = subgrid_dissipation(_host(net.closure(x, c2, 50.0, True)), X)
    d_ref = subgrid_dissipation(_host(t), X)
    print(f"\ncalibration factor {a:.3f}; dissipation {d_cl:.4e} vs targets {d_ref:.4e}")
    assert d_ref > 0 and abs(d_cl / d_ref - 1) < 1e-2
    les.close()
    net.close()
