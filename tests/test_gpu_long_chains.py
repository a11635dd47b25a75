"""GPU parity for convolutions longer than two passes per FFT.

A Rader prime whose n-1 is not 13-smooth, and a Bluestein length, both near
600000: their convolution lengths exceed a million points, so each FFT of the
convolution chain runs in three passes (five launches). The complex128 Bluestein
path keeps the reference's recurrence twiddles over those passes. Checked
against the oracle forward and inverse: complex128 within rel-L2 1e-14*log2(n),
complex64 within 1e-6.
"""
import math

import numpy as np
import pytest

from conftest import cgauss, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


@pytest.mark.parametrize("n", [600011, 600014])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_long_chain(dev, oracle, n, prec):
    D, torch = dev
    x = cgauss(n, seed=n % 1000).reshape(1, n)
    dt = torch.complex128 if prec == "c128" else torch.complex64
    nd = np.complex128 if prec == "c128" else np.complex64
    p = D.C2CPlan(n, 1, dt)
    assert p.num_launches() >= 5, p.describe()
    bar = 1e-14 * math.log2(n) if prec == "c128" else 1e-6
    for inverse in (False, True):
        want = oracle.fft_batch(x, inverse=inverse)
        got = p.exec(torch.from_numpy(x.astype(nd).ravel()).to("cuda"), inverse=inverse)
        e = rel_l2(got.cpu().numpy().astype(np.complex128).reshape(1, n), want)
        assert e <= bar, (n, prec, inverse, e, p.describe())
