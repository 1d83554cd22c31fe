"""pyspectral_amd: MI355X-native pseudo-spectral RHS + low-storage RK4 for PySpectral.

The hot path of PySpectral's serial solver (``PySpectral/src``) runs as
hand-written gfx950 HIP kernels in ``libpsx.so`` behind the C ABI of
``include/psx.h``; this package holds the ctypes binding, PyTorch custom ops and
the Python drop-in for the reference's entry points.
"""
from ._lib import PsxError, load  # noqa: F401

__all__ = ["PsxError", "load", "SpectralSolver"]


def __getattr__(name):
    if name == "SpectralSolver":
        from .solver import SpectralSolver

        return SpectralSolver
    raise AttributeError(name)
