"""Ensembles of uneven grid sizes: every member advances exactly as a single
SpectralSolver of its grid (same kernels, batched), eagerly and from one hipGraph."""
import numpy as np
import pytest

from oracle import spectral_oracle as so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _closure_net():
    from pyspectral_amd.learned import MLP

    net = MLP(seed=7)
    net.set_norm(np.zeros(9), np.full(9, 2.0), np.zeros(6), np.full(6, 0.02))
    return net


@pytest.mark.parametrize("model", [1, 0, "learned"])
def test_ensemble_members_match_single_solvers(model):
    from pyspectral_amd.ensemble import Ensemble
    from pyspectral_amd.solver import SpectralSolver

    sizes = [16, 32, 16, 48, 32, 64]
    nus = [2e-3, 5e-3, 8e-4, 1e-3, 3e-3, 6e-4]
    print(f"\nensemble model={model} sizes={sizes}", flush=True)
    ens = Ensemble([(n, n, n) for n in sizes], nus, model=model, cfl=0.5)
    net = _closure_net() if model == "learned" else None
    if net is not None:
        ens.set_closure(net, 3.0)
    q = ens.new_state()
    fields = [so.random_isotropic(n, seed=10 + i, kp=3.0 + i % 3) for i, n in enumerate(sizes)]
    for i, f in enumerate(fields):
        ens.member(q, i).copy_(torch.from_numpy(np.stack(f)))
    qs = q.clone()
    dt = ens.new_dt()
    singles = [SpectralSolver(n, n, n, model, nu=nu, cfl=0.5) for n, nu in zip(sizes, nus)]
    if net is not None:
        for s in singles:
            s.set_closure(net, 3.0)
    sq = [torch.from_numpy(np.stack(f)).cuda() for f in fields]
    sqs = [x.clone() for x in sq]
    for step in range(3):
        ens.step(q, qs, dt)
        for i, s in enumerate(singles):
            d = s.step(sq[i], sqs[i])
            torch.cuda.synchronize()
            assert float(d.item()) == float(dt[ens.slot(i)].item()), (step, i)
            assert torch.equal(ens.member(q, i), sq[i]), (step, i)
            assert torch.equal(ens.member(qs, i), sqs[i]), (step, i)
    st = ens.stats(q).cpu().numpy()
    for i, s in enumerate(singles):
        assert np.array_equal(st[i], s.stats(sq[i]).cpu().numpy())
        s.close()
    ens.close()


def test_ensemble_graph_replays_match_eager():
    from pyspectral_amd.ensemble import Ensemble

    sizes = [32, 16, 48, 16]
    nus = [1e-3, 4e-3, 2e-3, 6e-4]
    e1 = Ensemble([(n, n, n) for n in sizes], nus, model=1, cfl=0.5)
    e2 = Ensemble([(n, n, n) for n in sizes], nus, model=1, cfl=0.5)
    q1 = e1.new_state()
    for i, n in enumerate(sizes):
        e1.member(q1, i).copy_(torch.from_numpy(np.stack(so.random_isotropic(n, seed=i, kp=3.0))))
    s1, d1 = q1.clone(), e1.new_dt()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        q2, s2, d2 = q1.clone(), s1.clone(), e2.new_dt()
    stream.synchronize()
    for step in range(4):
        e1.step(q1, s1, d1)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            e2.step_graph(q2, s2, d2)
        stream.synchronize()
        _ = (q1.abs() * 2).sum().item()   # eager work between replays
        assert torch.equal(q1, q2), step
        assert torch.equal(s1, s2), step
        assert torch.equal(d1, d2), step
    e1.close()
    e2.close()


def test_ensemble_autotune_keeps_the_results():
    """psx_ensemble_autotune times each size class's kernel shapes on ensemble steps
    from a copy of the state (the state itself is unchanged); the replay captured after
    it matches the step of an untuned ensemble to round-off, and a graph captured before
    the tuning is re-captured rather than replayed with the old shapes."""
    from pyspectral_amd.ensemble import Ensemble
    from pyspectral_amd.solver import SpectralSolver

    sizes = [32, 16, 48, 16, 64]
    nus = [1e-3, 4e-3, 2e-3, 6e-4, 1e-3]
    e1 = Ensemble([(n, n, n) for n in sizes], nus, model=1, cfl=0.5)
    e2 = Ensemble([(n, n, n) for n in sizes], nus, model=1, cfl=0.5)
    q1 = e1.new_state()
    for i, n in enumerate(sizes):
        e1.member(q1, i).copy_(torch.from_numpy(np.stack(so.random_isotropic(n, seed=i, kp=3.0))))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        q2, s2, d2 = q1.clone(), q1.clone(), e2.new_dt()
        e2.step_graph(q2, s2, d2)          # captured with the built-in shapes
        q2.copy_(q1)
        s2.copy_(q1)
        q_before = q1.clone()
        tuned = e2.autotune(q1, reps=2)
        assert torch.equal(q1, q_before)
        e2.step_graph(q2, s2, d2)          # re-captured with the tuned shapes
    stream.synchronize()
    print("\n", tuned, flush=True)
    assert len(tuned) == 4 and all(set(t) == set(SpectralSolver.TUNING_KEYS) for t in tuned)
    s1, d1 = q1.clone(), e1.new_dt()
    e1.step(q1, s1, d1)
    torch.cuda.synchronize()
    # the CFL z pass of a retuned class may use the other kernel variant: dt to round-off
    assert (d1 - d2).abs().max().item() <= 1e-14 * d1.abs().max().item()
    err = (q1 - q2).abs().max().item() / q1.abs().max().item()
    assert err < 1e-13, err
    e1.close()
    e2.close()
