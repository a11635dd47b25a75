"""GPU: the radix-16 stage probe (gofft_probe_dft16), VALU butterflies and
the dense DFT-matrix product on the matrix cores, against numpy's FFT: one
application is the unitary DFT of each 16-point line, four give the input
back (F^4 = I), on a ragged count of lines (not a multiple of a wave's 32)."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import _lib as L
    return L, torch


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("transforms", [1, 77, 4096 + 45])
def test_probe_matches_fft(dev, method, transforms):
    L, torch = dev
    rng = np.random.default_rng(transforms)
    x = (rng.standard_normal((transforms, 16)) + 1j * rng.standard_normal((transforms, 16))).astype(np.complex64)
    xt = torch.from_numpy(x.ravel()).to("cuda")
    for reps, want in ((1, np.fft.fft(x.astype(np.complex128), axis=1) / 4.0), (4, x), (2, None)):
        out = torch.empty_like(xt)
        rc = L.lib().gofft_probe_dft16(method, ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       transforms, reps, None)
        assert rc == 0, L.last_error()
        got = out.cpu().numpy().reshape(transforms, 16)
        if want is None:  # F^2 x = x[-k]: index reversal mod 16
            want = x[:, (-np.arange(16)) % 16]
        assert rel_l2(got, want) <= 1e-6, (method, transforms, reps, rel_l2(got, want))
