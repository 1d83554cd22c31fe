"""Post-processing transforms and exact restarts on the GPU.

* ``psx_spectral`` (rfftn / sqrt(Nt) of real grid fields) against numpy, and as the
  inverse of ``psx_physical``;
* ``FFTFromFile``'s six transform objects (3/2, 2x and 3x grids,
  ``src/postProcessingFunctions.py:67-108``) against numpy's irfftn (pyFFTW's
  normalised backward call) and rfftn;
* a run stopped after k steps and restarted from the ``npsol`` file written then ends
  bit for bit where the uninterrupted run ends.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,nf", [((16, 32, 64), 5), ((48, 16, 32), 3), ((32, 32, 32), 1)])
def test_spectral_matches_numpy_and_inverts_physical(shape, nf):
    from pyspectral_amd.solver import SpectralSolver

    rng = np.random.default_rng(sum(shape) + nf)
    phys = rng.standard_normal((nf,) + shape)
    s = SpectralSolver(*shape, 0, nu=0.0)
    got = s.spectral(torch.from_numpy(phys).cuda())
    ref = np.stack([np.fft.rfftn(p) for p in phys]) / np.sqrt(np.prod(shape))
    err = np.linalg.norm(got.cpu().numpy() - ref) / np.linalg.norm(ref)
    back = s.physical(got).cpu().numpy()
    err_b = np.linalg.norm(back - phys) / np.linalg.norm(phys)
    print(f"\nspectral {shape} x{nf}: vs numpy {err:.2e}, round trip {err_b:.2e}")
    assert err < 1e-14 and err_b < 1e-14
    s.close()


def test_fftfromfile_transforms_match_numpy(tmp_path):
    from pyspectral_amd import postProcessingFunctions as pp

    N = 32
    k = np.zeros((N, N, N // 2 + 1))
    np.savez(tmp_path / "grid.npz", k1=k, k2=k, k3=k)
    with np.load(tmp_path / "grid.npz") as gf:
        fft = pp.FFTFromFile(gf, 4)
    assert fft.scale == pytest.approx(np.sqrt((1.5) ** 3 * np.sqrt(N ** 3)))
    rng = np.random.default_rng(3)
    for suffix, M in (("", 3 * N // 2), ("2", 2 * N), ("3", 3 * N)):
        inv, fwd = getattr(fft, "ifftT_obj" + suffix), getattr(fft, "fft_obj" + suffix)
        spec = rng.standard_normal((M, M, M // 2 + 1)) + 1j * rng.standard_normal((M, M, M // 2 + 1))
        x = inv(spec)
        ref = np.fft.irfftn(spec, s=(M, M, M), axes=(0, 1, 2))
        e_inv = np.linalg.norm(x - ref) / np.linalg.norm(ref)
        real = rng.standard_normal((M, M, M))
        y = fwd(real)
        ref_f = np.fft.rfftn(real)
        e_fwd = np.linalg.norm(y - ref_f) / np.linalg.norm(ref_f)
        print(f"\nFFTFromFile grid {M}^3: ifftT_obj{suffix} {e_inv:.2e}, fft_obj{suffix} {e_fwd:.2e}")
        assert x.shape == (M, M, M) and y.shape == (M, M, M // 2 + 1)
        assert e_inv < 1e-14 and e_fwd < 1e-14
        # device tensors in, device tensors out
        yd = fwd(torch.from_numpy(real).cuda())
        assert yd.is_cuda and np.array_equal(yd.cpu().numpy(), y)


@pytest.mark.parametrize("turb_model", [1, 2])
def test_restart_from_npsol_is_exact(turb_model, tmp_path):
    """Stop after 5 steps (the savehook writes npsol4.npz), restart from that file for 6
    more: state, stale velocity, time and the per-step energies equal the uninterrupted
    11-step run's bit for bit (CFL-limited dt, and the t-model's time-dependent term)."""
    from pyspectral_amd import cases
    from pyspectral_amd import postProcessingFunctions as pp
    from pyspectral_amd.PySpec_3d import run_case

    N, nu = 16, 0.01
    uh = cases.decaying_isotropic(N, seed=4)
    x = np.arange(N) * 2 * np.pi / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")

    def case(outdir, hook=None):
        g = dict(turb_model=turb_model, N1=N, N2=N, N3=N, kc=N // 2, nu=nu, t=0.0, et=10.0, save_freq=4,
                 dt=0.01, cfl=0.5, x=X, y=Y, z=Z, uhat=uh[0], vhat=uh[1], what=uh[2])
        if hook:
            g["savehook"] = hook
        return g

    full = run_case(case(str(tmp_path / "a")), outdir=str(tmp_path / "a"), verbose=False, max_steps=11)
    out_b = str(tmp_path / "b")

    def hook(main, grid, utilities, iteration, myFFT):
        pp.save_npsol(main, iteration, out_b)

    run_case(case(out_b, hook), outdir=out_b, verbose=False, max_steps=5)
    path = os.path.join(out_b, "npsol4.npz")
    with np.load(path) as f:
        assert {"uhat", "vhat", "what", "t", "Qhat", "t_Qhat", "iteration"} <= set(f.files)
    res = run_case(case(str(tmp_path / "c")), outdir=str(tmp_path / "c"), verbose=False, max_steps=6, resume=path)
    a, c = full["main"], res["main"]
    torch.cuda.synchronize()
    assert torch.equal(a.Qd, c.Qd) and torch.equal(a.Ud, c.Ud)
    assert a.t == c.t
    assert np.array_equal(full["Energy"][5:], res["Energy"])
    assert np.array_equal(full["t"][5:], res["t"])
