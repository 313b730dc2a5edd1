"""The fftz module (fftz.f90:32-127) through bl_fft_z / bl_ifft_z, which run
the solver's z-DFT kernels, against numpy's r2c / c2r: forward unnormalised,
inverse divided by N (FFTW's c2r / N, as compute_ifft_z does), for prime,
odd, even and power-of-two lengths including the 191 of the 814x257x193
workload.  The c2r ignores the imaginary parts of the zero and (even N)
Nyquist terms, as FFTW and numpy do."""
import sys

import numpy as np
import pytest

from .blgpu_import import blgpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [2, 6, 7, 16, 33, 64, 191, 254])
def test_fftz_matches_numpy(N):
    rng = np.random.default_rng(N)
    n1 = 37
    U = np.asfortranarray(rng.standard_normal((n1, N)))
    f = blgpu.FFTZ(N)
    H = f.forward(U)
    Hn = np.fft.rfft(U, axis=1)
    ef = np.abs(H - Hn).max() / np.abs(Hn).max()
    # Hermitian-consistent spectrum back to real, then a spectrum with
    # imaginary parts where c2r discards them
    G = rng.standard_normal((n1, N // 2 + 1)) + 1j * rng.standard_normal((n1, N // 2 + 1))
    Ui = f.inverse(G)
    Un = np.fft.irfft(G, n=N, axis=1)
    ei = np.abs(Ui - Un).max() / np.abs(Un).max()
    er = np.abs(f.inverse(H) - U).max() / np.abs(U).max()
    print("fftz N", N, "forward", ef, "inverse", ei, "round trip", er, flush=True)
    sys.stdout.flush()
    assert ef < 1e-13 and ei < 1e-13 and er < 1e-13
    f.close()
