"""GPU parity: FFTN/FFT2 and RFFT2/IRFFT2 against the oracle."""
import math

import numpy as np
import pytest

from conftest import cgauss, rel_l2, rgauss

pytestmark = pytest.mark.gpu


def bound64(n):
    return 1e-14 * max(1.0, math.log2(max(n, 2)))


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


SHAPES = [[2, 2], [3, 4], [4, 3], [1, 5], [5, 1], [2, 3, 4], [3, 1, 2], [2, 2, 2, 2], [6, 5],
          [17, 8], [64, 64], [256, 256], [4096, 64], [64, 4096], [1009, 16], [16, 1009], [32, 32, 32],
          [1, 1], [7]]


@pytest.mark.parametrize("shape", SHAPES)
def test_fftn_fp64(dev, oracle, shape):
    D, torch = dev
    total = int(np.prod(shape))
    x = cgauss(total, seed=total)
    p = D.NDPlan(shape, 0, 1, torch.complex128)
    xt = torch.from_numpy(x).to("cuda")
    got = p.exec(xt).cpu().numpy()
    assert rel_l2(got, oracle.fftn(x, shape)) <= bound64(total), shape
    goti = p.exec(xt, inverse=True).cpu().numpy()
    assert rel_l2(goti, oracle.fftn(x, shape, inverse=True)) <= bound64(total), shape


def test_fft2_4096_fp32(dev, oracle):
    D, torch = dev
    shape = [4096, 4096]
    x = cgauss(4096 * 4096, seed=0)
    p = D.NDPlan(shape, 0, 1, torch.complex64)
    got = p.exec(torch.from_numpy(x.astype(np.complex64)).to("cuda")).cpu().numpy()
    want = np.fft.fft2(x.reshape(shape)).ravel()
    assert rel_l2(got, want) <= 1e-6


# Strided axes long enough for several passes, over rows that are not a
# whole number of 128-byte lines or of column groups (such passes place
# consecutive column groups on one XCD).
@pytest.mark.parametrize("shape", [[40000, 3], [20000, 5], [3, 40000, 2], [65536, 3]])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_fftn_long_strided_axes(dev, oracle, shape, prec):
    D, torch = dev
    total = int(np.prod(shape))
    x = cgauss(total, seed=total % 997)
    dt, nd = (torch.complex128, np.complex128) if prec == "c128" else (torch.complex64, np.complex64)
    p = D.NDPlan(shape, 0, 1, dt)
    xt = torch.from_numpy(x.astype(nd)).to("cuda")
    bar = bound64(total) if prec == "c128" else 1e-6
    for inverse in (False, True):
        got = p.exec(xt, inverse=inverse).cpu().numpy().astype(np.complex128)
        e = rel_l2(got, oracle.fftn(x, shape, inverse=inverse))
        assert e <= bar, (shape, prec, inverse, e, p.describe())


@pytest.mark.parametrize("shape", [(4, 6), (3, 4), (5, 8), (2, 2), (4, 5), (1, 6), (6, 1), (7, 3),
                                   (64, 128), (512, 1024), (1024, 1024), (20000, 6), (16384, 10)])
def test_rfft2_irfft2_fp64(dev, oracle, shape):
    D, torch = dev
    r, c = shape
    x = rgauss(r * c, seed=r * c)
    p = D.NDPlan(list(shape), 1, 1, torch.float64)
    X = p.rfft2(torch.from_numpy(x).to("cuda"))
    want = oracle.rfft2(x, shape)
    assert rel_l2(X.cpu().numpy(), want) <= bound64(r * c), shape
    back = p.irfft2(torch.from_numpy(want).to("cuda")).cpu().numpy()
    assert rel_l2(back, oracle.irfft2(want, shape)) <= bound64(r * c), shape
    # short spectrum zero-pads (rfft2_test.go:85-100)
    short = want[:-1].copy()
    bs = p.irfft2(torch.from_numpy(short).to("cuda"), bins=len(short)).cpu().numpy()
    assert rel_l2(bs, oracle.irfft2(short, shape)) <= bound64(r * c), shape


def test_nd_shape_errors(dev):
    D, torch = dev
    from fft_amd import GofftError
    with pytest.raises(GofftError, match="shape lengths must be positive"):
        D.NDPlan([0, 4], 0, 1, torch.complex128)
    with pytest.raises(GofftError, match="shape lengths must be positive"):
        D.NDPlan([-1, 4], 0, 1, torch.complex128)
