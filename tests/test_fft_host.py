"""The FFT butterflies, radix plans and Stockham stage/twiddle indexing of
csrc/fft_core.hpp, run on the host (libpsx_hosttest.so) against numpy.fft for
every supported length, and the radix-R butterflies the z pass's register step uses."""
import ctypes as C
import os

import numpy as np
import pytest

from pyspectral_amd import build_lib

LENGTHS = [8, 12, 16, 24, 32, 36, 48, 64, 72, 96, 128, 144, 192, 256, 288, 384, 512, 576, 768]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(build_lib.HOSTTEST_OUT):
        build_lib.build_hosttest(verbose=False)
    lib = C.CDLL(build_lib.HOSTTEST_OUT)
    lib.psx_host_fft.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.psx_host_fft.restype = C.c_int
    lib.psx_host_dft.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.psx_host_dft.restype = C.c_int
    return lib


@pytest.mark.parametrize("L", LENGTHS)
def test_host_fft_matches_numpy(lib, L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    for sign in (-1, 1):
        buf = np.ascontiguousarray(x.astype(np.complex128))
        rc = lib.psx_host_fft(L, sign, buf.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 0
        ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * L
        err = np.abs(buf - ref).max() / np.abs(ref).max()
        assert err < 5e-15, (L, sign, err)


@pytest.mark.parametrize("R", [2, 3, 4, 6, 8, 12])
def test_host_butterflies_match_numpy(lib, R):
    rng = np.random.default_rng(100 + R)
    x = rng.standard_normal(R) + 1j * rng.standard_normal(R)
    for sign in (-1, 1):
        buf = np.ascontiguousarray(x.astype(np.complex128))
        assert lib.psx_host_dft(R, sign, buf.ctypes.data_as(C.POINTER(C.c_double))) == 0
        ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * R
        assert np.abs(buf - ref).max() < 5e-15 * np.abs(ref).max(), (R, sign)
