"""Drop-in ``computeRHS_*(main, grid, myFFT, utilities)`` (``src/RHSfunctions.py``).

Each function has the reference's signature and side effects: ``main.Q2U()``,
then the right-hand side of ``main.uhat`` written into ``main.Q``, with
``main.uhat`` left as the model leaves it (Nyquist entries zeroed for the models
that truncate, ``src/RHSfunctions.py:10-12``) and, for the closures,
``main.w0_{u,v,w}`` set.  The work runs in libpsx's HIP kernels on ``main.Qd`` /
``main.Ud``; which closure runs is fixed by the context ``variables`` created
for ``turb_model``.
"""
from __future__ import annotations

import torch


def _rhs(main, want_w0: bool):
    main.Q2U()
    uin = main.Ud.clone()
    w0 = None
    if want_w0:
        if main.w0d is None:
            main.w0d = torch.empty_like(main.Ud)
        w0 = main.w0d
    main.solver.rhs(uin, main.t, out=main.Qd, uhat_model=main.Ud, w0=w0)


def computeRHS_NOSGS(main, grid, myFFT, utilities):
    """``src/RHSfunctions.py:3-50``."""
    _rhs(main, False)


def computeRHS_NOSGS_WEAVE(main, grid, myFFT, utilities):
    """``src/RHSfunctionsWEAVE.py:5-77``: the arithmetic of ``computeRHS_NOSGS`` with the
    assembly fused into one loop (``:44-66``), which has no Coriolis terms: rotation is
    off whatever ``main.rotate`` says."""
    s = main.solver
    if s.params.rotate:
        s.set_params(rotate=0)
        try:
            _rhs(main, False)
        finally:
            s.set_params(rotate=1)
    else:
        _rhs(main, False)


def computeRHS_SMAG(main, grid, myFFT, utilities):
    """``src/RHSfunctions.py:146-237``."""
    _rhs(main, True)


def computeRHS_tmodel(main, grid, myFFT, utilities):
    """``src/RHSfunctions.py:240-363``."""
    _rhs(main, True)


def computeRHS_staumodel(main, grid, myFFT, utilities):
    """``src/RHSfunctions.py:366-497``."""
    _rhs(main, True)


def computeRHS_Dtaumodel(main, grid, myFFT, utilities):
    """``src/RHSfunctions.py:592-878``."""
    _rhs(main, True)


def computeRHS_DSMAG(main, grid, myFFT, utilities):
    """``src/RHSfunctions.py:1663-1862`` (dynamic Smagorinsky, turb_model 7)."""
    _rhs(main, True)
