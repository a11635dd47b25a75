"""The oracle against the committed golden vectors: numpy FP64 outputs on the
reference correctness gate's inputs (tests/golden/make_golden.py). CPU only;
the GPU path is checked against the same vectors in test_gpu_golden.py."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, h2h_complex, h2h_real, rel_l2

GOLD = np.load(os.path.join(ROOT, "tests", "golden", "h2h_numpy.npz"))
RTOL, ATOL = 1e-9, 1e-7  # verify_correctness.py:16-17


def bound(n):
    return 1e-14 * max(1.0, math.log2(n))


@pytest.mark.parametrize("n", [256, 1024, 4096, 1000, 1080, 1920, 1009, 1296, 10007])
def test_complex(oracle, n):
    want = GOLD["complex_%d" % n]
    got = oracle.fft(h2h_complex(n))
    assert np.allclose(got, want, rtol=RTOL, atol=ATOL)
    assert rel_l2(got, want) <= bound(n)


@pytest.mark.parametrize("n", [256, 1024, 4096, 1000, 1080, 1920])
def test_real(oracle, n):
    want = GOLD["real_%d" % n]
    got = oracle.rfft(h2h_real(n))
    assert np.allclose(got, want, rtol=RTOL, atol=ATOL) and rel_l2(got, want) <= bound(n)
    wanti = GOLD["ireal_%d" % n]
    goti = oracle.irfft(h2h_complex(n // 2 + 1), n)
    assert np.allclose(goti, wanti, rtol=RTOL, atol=ATOL) and rel_l2(goti, wanti) <= bound(n)


@pytest.mark.parametrize("shape", [(64, 64), (128, 128)])
def test_fft2(oracle, shape):
    r, c = shape
    want = GOLD["fft2_%dx%d" % shape]
    got = oracle.fftn(h2h_complex(r * c), [r, c])
    assert np.allclose(got, want, rtol=RTOL, atol=ATOL) and rel_l2(got, want) <= bound(r * c)
