"""hipGraph replay of the full step (psx_step_graph) against eager psx_step.

Every replay is compared with the eager result, with eager work and
synchronisations between the replays, and the device-held time must advance
exactly as the host loop's t += dt."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", [0, 1, 2, "Dynamic tau-model", 7])
def test_graph_replays_match_eager(model):
    from pyspectral_amd import cases
    from pyspectral_amd.solver import SpectralSolver

    N = 16
    uh = cases.decaying_isotropic(N, seed=5, k0=2.0)
    kw = dict(nu=0.01, cfl=0.5, Ct=0.3)
    eager = SpectralSolver(N, N, N, model, **kw)
    graphed = SpectralSolver(N, N, N, model, **kw)
    qe = torch.from_numpy(uh).cuda()
    se = qe.clone()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        qg = torch.from_numpy(uh).cuda()
        sg = qg.clone()
        dtg = torch.zeros(1, dtype=torch.float64, device="cuda")
        tg = torch.full((1,), 0.2, dtype=torch.float64, device="cuda")
    stream.synchronize()
    t = 0.2
    for step in range(6):
        dte = eager.step(qe, se, t)
        torch.cuda.synchronize()
        t += float(dte.item())
        with torch.cuda.stream(stream):
            graphed.step_graph(qg, sg, dtg, tg)
        stream.synchronize()
        # eager work on the default stream between replays
        _ = (qe * 2).sum().item()
        assert torch.equal(qg, qe), f"state differs after replay {step}"
        assert torch.equal(sg, se), f"stale velocity differs after replay {step}"
        assert float(dtg.item()) == float(dte.item())
        assert abs(float(tg.item()) - t) <= 1e-15 * abs(t)
    eager.close()
    graphed.close()


def test_graph_needs_a_stream():
    from pyspectral_amd._lib import PsxError
    from pyspectral_amd.solver import SpectralSolver

    s = SpectralSolver(16, 16, 16, 0, nu=0.01)
    q = torch.zeros((3, 16, 16, 9), dtype=torch.complex128, device="cuda")
    with pytest.raises(PsxError):
        s.step_graph(q, q.clone(), torch.zeros(1, dtype=torch.float64, device="cuda"),
                     torch.zeros(1, dtype=torch.float64, device="cuda"))
    np.testing.assert_equal(q.abs().max().item(), 0.0)
