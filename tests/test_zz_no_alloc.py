"""psx_stats_noalloc (the diagnostics entry point that allocates nothing) against
psx_stats and the oracle: two calls in a row, the second on a smaller grid, so the
second reuses the static partials array that the first filled."""
import ctypes as C

import numpy as np
import pytest

from oracle import spectral_oracle as so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _stats_noalloc(lib, u, members, N, kc, stream=None):
    out = torch.empty(3 * members, dtype=torch.float64, device="cuda")
    rc = lib.psx_stats_noalloc(u.data_ptr(), members, N, N, N, 2 * np.pi, 2 * np.pi, 2 * np.pi, kc,
                               out.data_ptr(), stream)
    assert rc == 0, rc
    return out


def test_stats_noalloc_matches_psx_stats_and_oracle():
    from pyspectral_amd import _lib
    from pyspectral_amd.solver import SpectralSolver

    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    for N, members in ((48, 2), (16, 1)):  # the second call is smaller
        print(f"\nstats_noalloc N={N} members={members}", flush=True)
        fields = [np.stack(so.random_isotropic(N, seed=40 + m, kp=3.0)) for m in range(members)]
        u = torch.from_numpy(np.ascontiguousarray(np.stack(fields))).to("cuda")
        got = _stats_noalloc(lib, u, members, N, N // 3, C.c_void_p(stream))
        s = SpectralSolver(N, N, N, 1, nu=0.01, kc=N // 3)
        for m in range(members):
            ref = s.stats(u[m].contiguous())
            torch.cuda.synchronize()
            assert torch.equal(got[3 * m:3 * m + 3], ref), (N, m)
            st = so.State(so.Grid(N, N, N, kc=N // 3), *fields[m], nu=0.01)
            o = np.array([so.compute_energy(st), so.compute_enstrophy(st), so.compute_energy_resolved(st)])
            assert np.allclose(got[3 * m:3 * m + 3].cpu().numpy(), o, rtol=1e-13, atol=0)
        s.close()


def test_stats_noalloc_rejects_bad_arguments():
    from pyspectral_amd import _lib

    lib = _lib.load()
    out = torch.empty(3, dtype=torch.float64, device="cuda")
    u = torch.zeros((3, 16, 16, 9), dtype=torch.complex128, device="cuda")
    assert lib.psx_stats_noalloc(u.data_ptr(), 1, 15, 16, 16, 1.0, 1.0, 1.0, 4, out.data_ptr(), None) == -1
    assert lib.psx_stats_noalloc(u.data_ptr(), 0, 16, 16, 16, 1.0, 1.0, 1.0, 4, out.data_ptr(), None) == -1
    assert lib.psx_stats_noalloc(None, 1, 16, 16, 16, 1.0, 1.0, 1.0, 4, out.data_ptr(), None) == -1
    assert lib.psx_stats_noalloc(u.data_ptr(), 1, 16, 16, 16, 0.0, 1.0, 1.0, 4, out.data_ptr(), None) == -1
