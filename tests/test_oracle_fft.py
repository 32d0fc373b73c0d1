"""The oracle's FFT (stand-in for cuFFT's C2C transforms, cwes3.cuh:175-193)
pinned to the DFT definition: brute O(N^2) sums, numpy.fft, Parseval,
impulse/tone known answers and the round trip.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def brute(x, sign):
    N = x.size
    j = np.arange(N)
    return np.array([np.sum(x * np.exp(sign * 2j * np.pi * j * k / N)) for k in range(N)])


@pytest.mark.parametrize("N", [2, 4, 8, 16, 64, 256])
@pytest.mark.parametrize("sign", [+1, -1])
def test_fft_matches_brute_dft(oracle_lib, N, sign):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ref = brute(x, sign)
    assert np.max(np.abs(O.fft(x, sign) - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("N", [1024, 1 << 14])
def test_fft_matches_numpy(oracle_lib, N):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    # CUFFT_INVERSE (sign +1) is N * numpy's ifft; CUFFT_FORWARD is numpy's fft
    assert np.max(np.abs(O.fft(x, +1) - N * np.fft.ifft(x))) < 1e-12 * np.sqrt(N) * np.linalg.norm(x)
    assert np.max(np.abs(O.fft(x, -1) - np.fft.fft(x))) < 1e-12 * np.sqrt(N) * np.linalg.norm(x)


def test_fft_known_answers(oracle_lib):
    N = 512
    d = np.zeros(N, complex)
    d[0] = 1
    assert np.allclose(O.fft(d, -1), np.ones(N), atol=1e-15)
    m = 37
    tone = np.exp(-2j * np.pi * m * np.arange(N) / N)
    X = O.fft(tone, +1)  # IDFT+ of e^{-2 pi i m j/N} is N at k = m
    expect = np.zeros(N, complex)
    expect[m] = N
    assert np.max(np.abs(X - expect)) < 1e-10
    rng = np.random.default_rng(3)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    X = O.fft(x, -1)
    assert np.sum(np.abs(X) ** 2) == pytest.approx(N * np.sum(np.abs(x) ** 2), rel=1e-12)  # Parseval
    back = O.fft(O.fft(x, +1) / N, -1)  # the SSFM's to-frequency-and-back
    assert np.max(np.abs(back - x)) < 1e-13
