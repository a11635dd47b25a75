"""The GPU path against the committed golden vectors (numpy FP64 on the
reference correctness gate's inputs), through the reference-mirroring API."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, h2h_complex, h2h_real, rel_l2

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "h2h_numpy.npz"))


def bound(n):
    return 1e-14 * max(1.0, math.log2(n))


@pytest.fixture(scope="module")
def api():
    import torch
    assert torch.cuda.is_available()
    from fft_amd import api
    return api


@pytest.mark.parametrize("n", [256, 1024, 4096, 1000, 1080, 1920, 1009, 1296, 10007])
def test_complex(api, n):
    got = api.FFT(h2h_complex(n))
    assert rel_l2(got, GOLD["complex_%d" % n]) <= bound(n)


@pytest.mark.parametrize("n", [256, 1024, 4096, 1000, 1080, 1920])
def test_real(api, n):
    assert rel_l2(api.RFFT(h2h_real(n)), GOLD["real_%d" % n]) <= bound(n)
    assert rel_l2(api.IRFFT(h2h_complex(n // 2 + 1), n), GOLD["ireal_%d" % n]) <= bound(n)


@pytest.mark.parametrize("shape", [(64, 64), (128, 128)])
def test_fft2(api, shape):
    r, c = shape
    got = api.FFT2(h2h_complex(r * c), shape)
    assert rel_l2(got, GOLD["fft2_%dx%d" % shape]) <= bound(r * c)
