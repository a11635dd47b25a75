"""GPU parity: real transforms, PSD and Spectrogram against the oracle."""
import math

import numpy as np
import pytest

from conftest import real_signal, rel_l2, rgauss

pytestmark = pytest.mark.gpu


def bound64(n):
    return 1e-14 * max(1.0, math.log2(max(n, 2)))


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


# multi-pass half-lengths (untangle fused into the last pass): 1 << 17,
# 26880 (m = 2^7*3*5*7), 49152 (m = 2^13*3), 62500 (m = 2*5^6), 1 << 20;
# 1 << 22 (m = 2^21: two passes, 2048 x 1024 forward and 1024 x 2048 inverse,
# each paired pass with half its columns)
RLEN = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 17, 31, 32, 1000, 1024, 2046, 4096, 1 << 14, 1 << 17,
        1 << 20, 1 << 22, 20014, 2 * 1009, 26880, 49152, 62500,
        40028, 10007, 3 << 17, 2 * 3 ** 12]  # Bluestein half length, odd prime, multi-pass 3-smooth halves


@pytest.mark.parametrize("n", RLEN)
def test_rfft_irfft_fp64(dev, oracle, n):
    D, torch = dev
    batch = 3
    x = rgauss(n * batch, seed=n).reshape(batch, n)
    p = D.R2CPlan(n, batch, torch.float64)
    X = p.forward(torch.from_numpy(x.ravel()).to("cuda"))
    got = X.cpu().numpy().reshape(batch, n // 2 + 1)
    want = np.stack([oracle.rfft(r) for r in x])
    assert rel_l2(got, want) <= bound64(n), (n, rel_l2(got, want))
    # inverse of the oracle's spectrum
    back = p.inverse(torch.from_numpy(want.ravel()).to("cuda")).cpu().numpy().reshape(batch, n)
    wantb = np.stack([oracle.irfft(w, n) for w in want])
    assert rel_l2(back, wantb) <= bound64(n), (n, rel_l2(back, wantb))


@pytest.mark.parametrize("n", [1024, 1 << 20, 1 << 22, 1000, 20014, 1 << 17, 49152, 62500, 40028, 3 << 17])
def test_rfft_fp32(dev, oracle, n):
    D, torch = dev
    x = rgauss(n, seed=n)
    p = D.R2CPlan(n, 1, torch.float32)
    got = p.forward(torch.from_numpy(x.astype(np.float32)).to("cuda")).cpu().numpy()
    assert rel_l2(got, oracle.rfft(x)) <= 1e-6
    X = oracle.rfft(x)
    back = p.inverse(torch.from_numpy(X.astype(np.complex64)).to("cuda")).cpu().numpy()
    assert rel_l2(back, oracle.irfft(X, n)) <= 1e-6


def test_irfft_short_spectrum_and_dc_nyquist(dev, oracle):
    D, torch = dev
    for n in [4, 8, 1024, 1 << 18, 7, 62500, 26880]:
        p = D.R2CPlan(n, 1, torch.float64)
        spec = np.fft.rfft(real_signal(n)) + 0j
        spec[0] += 3j  # imaginary DC / Nyquist are ignored (realplan.go:243-250)
        if n % 2 == 0:
            spec[-1] += 2j
        for bins in [1, 2, n // 2, n // 2 + 1]:
            got = p.inverse(torch.from_numpy(spec[:bins].copy()).to("cuda"), bins=bins).cpu().numpy()
            want = oracle.irfft(spec[:bins], n)
            assert np.abs(got - want).max() <= 1e-12 * max(1, np.abs(want).max()), (n, bins)


@pytest.mark.parametrize("n", [32, 15, 1024, 1 << 20, 1000, 62500, 26880])
def test_psd(dev, oracle, n):
    D, torch = dev
    x = rgauss(n, seed=3)
    p = D.PSDPlan(n, 1, 0.1, torch.float64)
    got = p.exec(torch.from_numpy(x).to("cuda")).cpu().numpy()
    assert rel_l2(got, oracle.psd(x, 0.1)) <= bound64(n)


@pytest.mark.parametrize("seg,ov", [(1024, 512), (256, 100), (8, 4), (255, 10), (1024, 0)])
def test_spectrogram(dev, oracle, seg, ov):
    D, torch = dev
    x = rgauss(40000, seed=seg)
    w = oracle.window("hann", seg)
    for dtype, bar in [(torch.float64, 1e-13), (torch.float32, 1e-6)]:
        p = D.SpectrogramPlan(len(x), seg, ov, w, 2.0, dtype)
        xt = torch.from_numpy(x.astype(np.float64 if dtype == torch.float64 else np.float32)).to("cuda")
        got = p.exec(xt).cpu().numpy()[: p.frames * p.bins].reshape(p.frames, p.bins)
        want = oracle.spectrogram(x, seg, ov, w, 2.0)
        assert got.shape == want.shape
        assert rel_l2(got, want) <= bar, (seg, ov, dtype, rel_l2(got, want))


# Complex64 batches of power-of-two half length whose line groups fill the
# resident grid take the table forms (kernels.hip k_real_fwd_tab,
# k_real_inv_tab): persistent workgroups, a last group of partial lines,
# windows, odd hops (scalar sample loads), the PSD, plain bins and the
# inverse.
@pytest.mark.parametrize("n,batch", [(128, 50000), (1024, 6000), (512, 12000)])
def test_real_table_forms(dev, oracle, n, batch):
    D, torch = dev
    x = rgauss(n * batch, seed=n + batch).reshape(batch, n)
    p = D.R2CPlan(n, batch, torch.float32)
    assert "forward table form" in p.describe() and "inverse table form" in p.describe(), p.describe()
    got = p.forward(torch.from_numpy(x.astype(np.float32).ravel()).to("cuda")).cpu().numpy()
    want = np.stack([oracle.rfft(r) for r in x])
    assert rel_l2(got.reshape(batch, n // 2 + 1), want) <= 1e-6
    back = p.inverse(torch.from_numpy(want.astype(np.complex64).ravel()).to("cuda")).cpu().numpy()
    wantb = np.stack([oracle.irfft(w, n) for w in want])
    assert rel_l2(back.reshape(batch, n), wantb) <= 1e-6
    q = D.PSDPlan(n, batch, 0.1, torch.float32)
    assert "forward table form" in q.describe(), q.describe()
    got = q.exec(torch.from_numpy(x.astype(np.float32).ravel()).to("cuda")).cpu().numpy()
    want = np.stack([oracle.psd(r, 0.1) for r in x])
    assert rel_l2(got.reshape(batch, n // 2 + 1), want) <= 1e-6


@pytest.mark.parametrize("seg,ov", [(1024, 512), (256, 100), (256, 101), (128, 64)])
def test_spectrogram_table_form(dev, oracle, seg, ov):
    D, torch = dev
    x = rgauss(1 << 22, seed=seg + ov)
    w = oracle.window("hann", seg)
    p = D.SpectrogramPlan(len(x), seg, ov, w, 2.0, torch.float32)
    assert "forward table form" in p.describe(), p.describe()
    got = p.exec(torch.from_numpy(x.astype(np.float32)).to("cuda")).cpu().numpy()[: p.frames * p.bins]
    want = oracle.spectrogram(x, seg, ov, w, 2.0)
    assert rel_l2(got.reshape(p.frames, p.bins), want) <= 1e-6, (seg, ov)


def test_spectrogram_argument_errors(dev):
    D, torch = dev
    from fft_amd import GofftError
    with pytest.raises(GofftError, match="segment length must be positive"):
        D.SpectrogramPlan(100, 0, 0, [])
    with pytest.raises(GofftError, match="window length must equal segment length"):
        D.SpectrogramPlan(100, 4, 0, [1.0, 1.0])
    with pytest.raises(GofftError, match="overlap must satisfy"):
        D.SpectrogramPlan(100, 4, 4, [1.0] * 4)
    p = D.SpectrogramPlan(10, 16, 0, [1.0] * 16)
    assert p.frames == 0
