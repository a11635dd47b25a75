"""Writes tests/golden/h2h_numpy.npz: numpy.fft (pocketfft, FP64) outputs on the
closed-form inputs and size lists of the reference's correctness gate
(benchmarks/verify/main.go:17-79 builds the inputs; verify_correctness.py:16-63
compares go-fft against numpy.fft on them with rtol=1e-9, atol=1e-7).

The fixture is data only: inputs are regenerated from their closed forms by
the tests, so the file holds the expected outputs. numpy is an independent FP64
comparator; nothing from the reference is executed.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

COMPLEX_SIZES = [256, 1024, 4096, 1000, 1080, 1920, 1009, 1296, 10007]
REAL_SIZES = [256, 1024, 4096, 1000, 1080, 1920]
FFT2_SHAPES = [(64, 64), (128, 128)]


def cmplx(n):
    i = np.arange(n)
    return ((i * 7 + 1) % 13) * 0.1 + 1j * (((i * 3 + 2) % 11) * 0.1)


def realv(n):
    i = np.arange(n)
    return ((i * 7 + 1) % 13) * 0.1


def halfspec(n):
    return cmplx(n // 2 + 1)


def main():
    out = {}
    for n in COMPLEX_SIZES:
        out["complex_%d" % n] = np.fft.fft(cmplx(n))
    for n in REAL_SIZES:
        out["real_%d" % n] = np.fft.rfft(realv(n))
        out["ireal_%d" % n] = np.fft.irfft(halfspec(n), n)
    for r, c in FFT2_SHAPES:
        out["fft2_%dx%d" % (r, c)] = np.fft.fft2(cmplx(r * c).reshape(r, c)).ravel()
    np.savez_compressed(os.path.join(HERE, "h2h_numpy.npz"), **out)


if __name__ == "__main__":
    main()
