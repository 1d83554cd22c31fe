"""The Python drop-in (variables / advanceQ_RK4 / computeRHS_* / run_case) on the GPU,
against the oracle and the reference's shipped TGV fixture."""
import os

import numpy as np
import pytest

from oracle import spectral_oracle as so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(N, nu, dt, grid="reference", turb_model=0, et=1.0, **extra):
    from pyspectral_amd import cases

    uh, (dx, dy, dz) = cases.taylor_green(N, grid=grid)
    if grid == "reference":
        x = np.linspace(0, 2 * np.pi - 2 * np.pi / (N - 1), N)
    else:
        x = np.arange(N) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    g = dict(turb_model=turb_model, N1=N, N2=N, N3=N, kc=N // 2, nu=nu, t=0.0, et=et, save_freq=10 ** 9,
             dt=dt, nthreads=1, x=X, y=Y, z=Z, uhat=uh[0], vhat=uh[1], what=uh[2])
    g.update(extra)
    return g, (dx, dy, dz)


@pytest.mark.parametrize("turb_model", [0, 1, 2])
def test_run_case_matches_oracle(turb_model, tmp_path):
    from pyspectral_amd.PySpec_3d import run_case

    N, nu, dt = 16, 0.01, 0.01
    g, (dx, dy, dz) = _case(N, nu, dt, turb_model=turb_model, et=0.095)
    hooks = []
    g["savehook"] = lambda main, grid, utilities, iteration, myFFT: hooks.append((iteration, main.uhat.copy()))
    g["save_freq"] = 3
    res = run_case(g, outdir=str(tmp_path / "3DSolution"), verbose=False)
    grid = so.Grid(N, N, N, dx=dx, dy=dy, dz=dz, kc=N // 2)
    st = so.State(grid, g["uhat"], g["vhat"], g["what"], nu=nu, turb_model=turb_model, dt=dt)
    seen = []
    ref = so.run(st, so.FFTBackend(), et=0.095,
                 callback=lambda s, it, t: seen.append((it - 1, s.uhat.copy())) if (it - 1) % 3 == 0 else None)
    assert np.allclose(res["Energy"], ref["Energy"], rtol=1e-12, atol=0)
    assert np.allclose(res["Enstrophy"], ref["Enstrophy"], rtol=1e-12, atol=0)
    assert np.allclose(res["t"], ref["t"], rtol=1e-14, atol=0)
    assert [h[0] for h in hooks] == [s[0] for s in seen]
    for (_, a), (_, b) in zip(hooks, seen):
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    main = res["main"]
    Q = main.Q
    assert np.linalg.norm(Q - st.Q) <= 1e-12 * np.linalg.norm(st.Q)
    assert os.path.exists(tmp_path / "3DSolution" / "stats.npz")


def test_computeRHS_side_effects():
    from pyspectral_amd import Classes, RHSfunctions

    N, nu = 16, 0.02
    g, (dx, dy, dz) = _case(N, nu, 0.01)
    grid = Classes.gridclass(N, N, N, g["x"], g["y"], g["z"], N // 2)
    myFFT = Classes.FFTclass(N, N, N)
    main = Classes.variables(0, 1, 0, 0, 0, 0, grid, g["uhat"], g["vhat"], g["what"], 0.0, 0.01, nu, -10, -10,
                             -10, -10, -10, -0.01, 3, myFFT=myFFT)
    RHSfunctions.computeRHS_SMAG(main, grid, myFFT, Classes.utilitiesClass())
    og = so.Grid(N, N, N, dx=dx, dy=dy, dz=dz, kc=N // 2)
    st = so.State(og, g["uhat"], g["vhat"], g["what"], nu=nu, turb_model=1)
    so.rhs_smag(st, so.FFTBackend())
    assert np.linalg.norm(main.Q - st.Q) <= 1e-12 * np.linalg.norm(st.Q)
    assert np.linalg.norm(main.w0_u[..., 0] - st.w0[0]) <= 1e-12 * np.linalg.norm(st.w0[0]) + 1e-300


def test_tgv_re100_fixture_history(golden_dir):
    """The reference's TGV Re=100 history (64^3 case grid, dt = 0.010005, t in [0, 20])
    reproduced end to end by the GPU drop-in."""
    from pyspectral_amd.PySpec_3d import run_case

    d = np.load(os.path.join(golden_dir, "tgv_re100_benchmark_sol_64.npz"))
    E_ref, t_ref = d["Energy"].real, d["t"]
    dt = float(t_ref[1] - t_ref[0])
    g, _ = _case(64, 0.01, dt, et=t_ref[-1] - 0.5 * dt)
    res = run_case(g, outdir="gpurun_out/tgv_fixture", verbose=False)
    E = res["Energy"]
    n = min(len(E), len(E_ref))
    rel = np.abs(E[1:n] - E_ref[1:n]) / E_ref[1:n]
    print(f"\nTGV Re=100 vs fixture: {n} samples, max rel dE = {rel.max():.3e}, final rel = {rel[-1]:.3e}")
    # The fixture's provenance is not recorded; its first sample agrees to ~1.5e-7 and
    # the difference then grows linearly (~3e-7 per step), reaching a few 1e-3.
    assert n >= 1990
    assert rel[0] < 5e-7
    assert rel.max() < 5e-3


def test_solution_files_round_trip(tmp_path):
    """A Smagorinsky run writing npsol<it>.npz through the savehook; every file the run
    leaves loads without pickles and the loaders give back the saved state."""
    from pyspectral_amd import postProcessingFunctions as pp
    from pyspectral_amd.PySpec_3d import run_case

    N, nu, dt = 16, 0.01, 0.01
    g, _ = _case(N, nu, dt, turb_model=1, et=0.045)
    out = str(tmp_path / "3DSolution")
    seen = {}

    def hook(main, grid, utilities, iteration, myFFT):
        seen[iteration] = main.uhat.copy()
        pp.save_npsol(main, iteration, out)

    g["savehook"], g["save_freq"] = hook, 2
    run_case(g, outdir=out, verbose=False)
    with np.load(os.path.join(out, "runinfo.npz"), allow_pickle=False) as info, \
            np.load(os.path.join(out, "grid.npz"), allow_pickle=False) as gf:
        for it, u in seen.items():
            with np.load(os.path.join(out, f"npsol{it}.npz"), allow_pickle=False) as sol:
                v = pp.variablesFromFile(gf, sol, info, kc=N // 2)
                assert np.array_equal(v.uhat, u)
                assert v.w0_u.shape == u.shape
                gg = pp.gridFromFile(gf, sol, info, N // 2)
                assert gg.N1 == N
    stats = pp.load_stats(os.path.join(out, "stats.npz"))
    assert set(stats) == {"Energy", "Dissipation", "Energy_resolved", "Dissipation_resolved", "t", "Enstrophy"}
