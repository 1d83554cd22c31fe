"""Drop-in driver: ``advanceQ_RK4`` and the time loop of ``src/PySpec_3d.py``.

The reference runs a case by ``execfile``-ing its driver inside the case
script's globals (``run/*/inputfile.py:60``).  :func:`run_case` takes those
globals as a dict (or module namespace) and runs the same loop: defaults of
``src/PySpec_3d.py:9-72``, ``3DSolution/grid.npz`` and ``runinfo.npz``, the
``while t <= et`` loop with ``savehook`` every ``save_freq`` iterations,
per-step stats on the stale velocity and ``3DSolution/stats.npz`` at the end.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np
import torch

from .Classes import FFTclass, gridclass, utilitiesClass, variables


def advanceQ_RK4(main, grid, myFFT, utilities):
    """``src/PySpec_3d.py:96-105``: CFL dt on the stale ``main.uhat``, then the four
    low-storage stages ``Q = Q0 + dt*c_i*RHS(Q)``, c = (1/4, 1/3, 1/2, 1), fused on the
    GPU.  Afterwards ``main.Q`` holds the new state and ``main.uhat`` the stage-4
    input, as in the reference."""
    utilities.preAdvanceQ_hook(main, grid, myFFT)
    dt = main.solver.step(main.Qd, main.Ud, main.t, main._dt_dev)
    main.dt = float(dt.item())
    main.t_state = main.t + main.dt  # the time of the new state in main.Qd
    utilities.postAdvanceQ_hook(main, grid, myFFT)


_DEFAULTS = dict(weave=0, dim=3, rotate=0, Om1=0, Om2=0, Om3=0, Ct=-10, dt0=-10, dt1=-10,
                 dt0_subintegrations=-10, dt1_subintegrations=-10, L1=2.0 * np.pi, L2=2.0 * np.pi,
                 L3=2.0 * np.pi)


def run_case(g: dict, outdir: str = "3DSolution", verbose: bool = True, max_steps: int | None = None,
             resume: str | None = None):
    """Run a reference-style case from its globals ``g`` (needs N1, N2, N3, kc, nu, t,
    et, dt, save_freq, turb_model, x, y, z, uhat, vhat, what; optional savehook,
    cfl, nthreads and the defaults above).  ``max_steps`` bounds the steps of this call.
    ``resume``: an ``npsol`` file written by ``postProcessingFunctions.save_npsol``
    after a step; the run continues from its state, stale velocity, time and iteration
    exactly as the run that wrote it would have.  Returns the stats dict written to
    ``stats.npz``."""
    g = dict(g)
    for k, v in _DEFAULTS.items():
        g.setdefault(k, v)
    g.setdefault("cfl", -g["dt"])
    g.setdefault("nthreads", 1)
    utilities = utilitiesClass()
    myFFT = FFTclass(g["N1"], g["N2"], g["N3"], g["nthreads"])
    grid = gridclass(g["N1"], g["N2"], g["N3"], g["x"], g["y"], g["z"], g["kc"], g["L1"], g["L2"], g["L3"])
    main = variables(g["weave"], g["turb_model"], g["rotate"], g["Om1"], g["Om2"], g["Om3"], grid, g["uhat"],
                     g["vhat"], g["what"], g["t"], g["dt"], g["nu"], g["Ct"], g["dt0"], g["dt0_subintegrations"],
                     g["dt1"], g["dt1_subintegrations"], g["cfl"], g["dim"], myFFT=myFFT)
    os.makedirs(outdir, exist_ok=True)
    grid_k = grid.k1, grid.k2, grid.k3
    np.savez(os.path.join(outdir, "grid"), k1=grid_k[0], k2=grid_k[1], k3=grid_k[2], x=grid.x, y=grid.y, z=grid.z)
    np.savez(os.path.join(outdir, "runinfo"), turb_model=g["turb_model"], dt=g["dt"], save_freq=g["save_freq"],
             nu=g["nu"])
    t0 = time.time()
    t = float(g["t"])
    iteration = 0
    if resume is None:
        main.U2Q()
    else:
        with np.load(resume) as f:
            if "Qhat" not in f.files:
                raise ValueError(f"{resume} holds no Qhat: it was not written for a restart")
            main.Qd.copy_(torch.from_numpy(np.ascontiguousarray(f["Qhat"])))
            main.Ud.copy_(torch.from_numpy(np.ascontiguousarray(np.stack([f["uhat"], f["vhat"], f["what"]]))))
            t = float(f["t_Qhat"])
            iteration = int(f["iteration"]) + 1  # the file was written after that step
    et = float(g["et"])
    save_freq = int(g["save_freq"])
    savehook = g.get("savehook")
    s = utilities._stats(main)
    Energy, Enstrophy, Energy_resolved, t_hist = [s[0]], [s[1]], [s[2]], [t]
    first = iteration
    while t <= et:
        if max_steps is not None and iteration - first >= max_steps:
            break
        main.t = t
        main.iteration = iteration
        advanceQ_RK4(main, grid, myFFT, utilities)
        t += main.dt
        if savehook is not None and iteration % save_freq == 0:
            savehook(main, grid, utilities, iteration, myFFT)
        iteration += 1
        enstrophy, energy, dissipation, lambda_k, tau_k, Re_lambda = utilities.computeAllStats(main, grid)
        Energy.append(energy)
        Enstrophy.append(enstrophy)
        Energy_resolved.append(utilities.computeEnergy_resolved(main, grid))
        t_hist.append(t)
        if verbose:
            sys.stdout.write("Wall Time= %.4f   t=%.6g   Energy = %.10g  eps = %.10g\n"
                             % (time.time() - t0, t, energy, dissipation))
    main.t = t
    E, Er, th = np.array(Energy), np.array(Energy_resolved), np.array(t_hist)
    stats = dict(Energy=E, Dissipation=(E[1:] - E[:-1]) / (th[1:] - th[:-1]), Energy_resolved=Er,
                 Dissipation_resolved=(Er[1:] - Er[:-1]) / (th[1:] - th[:-1]), t=th, Enstrophy=np.array(Enstrophy))
    np.savez(os.path.join(outdir, "stats"), **stats)
    stats["main"], stats["grid"] = main, grid
    return stats
