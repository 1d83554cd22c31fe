"""PyTorch custom ops over libpsx (``pyspectral::rhs``, ``pyspectral::rk4_step_``,
``pyspectral::stats``).  A solver is named by the integer id of a
:class:`pyspectral_amd.solver.SpectralSolver`."""
from __future__ import annotations

import torch

from . import solver as _solver


@torch.library.custom_op("pyspectral::rhs", mutates_args=())
def rhs(solver_id: int, uhat: torch.Tensor, t: float) -> torch.Tensor:
    """computeRHS_<model>(uhat) at step-start time t."""
    return _solver.get(solver_id).rhs(uhat, t)


@rhs.register_fake
def _rhs_fake(solver_id, uhat, t):
    return torch.empty_like(uhat)


@torch.library.custom_op("pyspectral::rk4_step_", mutates_args=("q", "qs", "dt"))
def rk4_step_(solver_id: int, q: torch.Tensor, qs: torch.Tensor, dt: torch.Tensor, t: float) -> None:
    """advanceQ_RK4 in place: q <- new state, qs <- stale stage-4 input, dt <- step."""
    _solver.get(solver_id).step(q, qs, t, dt)


@torch.library.custom_op("pyspectral::stats", mutates_args=())
def stats(solver_id: int, uhat: torch.Tensor) -> torch.Tensor:
    """(E, Omega, E_resolved) of uhat."""
    return _solver.get(solver_id).stats(uhat)


@stats.register_fake
def _stats_fake(solver_id, uhat):
    return uhat.new_empty(3, dtype=torch.float64)
