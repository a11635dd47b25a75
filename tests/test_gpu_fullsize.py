"""GPU checks at the full sizes of the spectral and multi-dimensional
workloads: every output against the oracle (FP64 restatement of the
reference, its N-D and spectrogram line fan-out on 16 host threads) on the
same input, plus properties the oracle does not supply: round trips,
Parseval's identity, and single bins against direct sums with exact integer
phases.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


_SIGNALS = {}


def signal(n, seed=0):
    """Sum of 64 sinusoids (f ~ U[0,0.5), a ~ U[0,1], phase 0) plus N(0, 0.1^2)
    noise, as a float64 host array (built once per size and seed)."""
    import torch
    key = (n, seed)
    if key not in _SIGNALS:
        rng = np.random.default_rng(seed)
        f = rng.uniform(0, 0.5, 64)
        a = rng.uniform(0, 1, 64)
        t = torch.arange(n, dtype=torch.float64, device="cuda")
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        for fj, aj in zip(f, a):
            x += float(aj) * torch.sin((2 * math.pi * float(fj)) * t)
        _SIGNALS.clear()
        _SIGNALS[key] = x.cpu().numpy() + rng.normal(0, 0.1, n)
    return _SIGNALS[key]


def _dft_sum(x, ph, den):
    """sum_j x_j exp(-2 pi i ph_j / den) in FP64 on the GPU (torch elementwise
    work and a tree sum, no FFT), ph the exact integer phase indices."""
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda")
    ang = ph.to(torch.float64) * (-2 * math.pi / den)
    re = xt.real if xt.is_complex() else xt
    im = xt.imag if xt.is_complex() else torch.zeros_like(xt)
    c, s_ = torch.cos(ang), torch.sin(ang)
    sr = float((re * c - im * s_).sum())
    si = float((re * s_ + im * c).sum())
    return complex(sr, si)


def dft_bin_1d(x, k):
    import torch
    n = len(x)
    ph = (torch.arange(n, dtype=torch.int64, device="cuda") * k) % n
    return _dft_sum(x, ph, n)


def dft_bin_2d(x, u, v):
    import torch
    rows, cols = x.shape
    pr = (torch.arange(rows, dtype=torch.int64, device="cuda") * u) % rows
    pc = (torch.arange(cols, dtype=torch.int64, device="cuda") * v) % cols
    # exact phase index of u*r/rows + v*c/cols over the common denominator
    ph = ((pr[:, None] * cols + pc[None, :] * rows) % (rows * cols)).reshape(-1)
    return _dft_sum(x.reshape(-1), ph, rows * cols)


def bin_tol(prec):
    """Error of one bin relative to ||x||_2 + |X_k| (the DC bin of a
    positive signal carries errors aligned with its own size)."""
    return 1e-13 if prec in ("f64", "c128") else 5e-6


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_spectrogram_2_26(dev, oracle, prec):
    D, torch = dev
    n, seg, ov = 1 << 26, 1024, 512
    x = signal(n)
    win = oracle.window("hann", seg)  # symmetric
    rt = torch.float64 if prec == "f64" else torch.float32
    sp = D.SpectrogramPlan(n, seg, ov, list(win), 1.0, rt)
    assert sp.frames == (n - seg) // (seg - ov) + 1 == 131071
    xt = torch.from_numpy(x).to("cuda", rt)
    out = sp.exec(xt).cpu().numpy().astype(np.float64).reshape(sp.frames, sp.bins)
    xs = x if prec == "f64" else x.astype(np.float32).astype(np.float64)
    bar = 1e-14 * 10 if prec == "f64" else 1e-6
    # every frame against the oracle's Spectrogram (freq.go:107-127)
    want = oracle.spectrogram(xs, seg, ov, win, 1.0, threads=16)
    assert want.shape == out.shape
    err = np.linalg.norm(out - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() <= bar, (prec, int(err.argmax()), float(err.max()))
    # and a few frames through the oracle's PSD of the windowed frame
    rng = np.random.default_rng(1)
    for t in [0, sp.frames - 1] + list(rng.choice(sp.frames, size=4, replace=False)):
        frame = xs[t * (seg - ov): t * (seg - ov) + seg] * win
        assert rel_l2(out[t], oracle.psd(frame, 1.0)) <= bar, (prec, t)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_psd_2_26(dev, oracle, prec):
    D, torch = dev
    n = 1 << 26
    x = signal(n, seed=2)
    rt = torch.float64 if prec == "f64" else torch.float32
    p = D.PSDPlan(n, 1, 0.5, rt)
    xt = torch.from_numpy(x).to("cuda", rt)
    got = p.exec(xt).cpu().numpy().astype(np.float64)
    xs = x if prec == "f64" else x.astype(np.float32).astype(np.float64)
    # the whole PSD against the oracle's (freq.go:69-90); |X|^2 doubles the
    # spectrum's relative error
    want = oracle.psd(xs, 0.5)
    bar = 2e-14 * 25 if prec == "f64" else 1e-6
    assert rel_l2(got, want) <= bar, (prec, rel_l2(got, want))
    # Parseval: sum_k psd_k = d * sum_j x_j^2
    want_sum = 0.5 * float(np.sum(xs * xs))
    tol = 1e-12 if prec == "f64" else 1e-5
    assert abs(got.sum() - want_sum) <= tol * want_sum
    scale = math.sqrt(float(np.sum(xs * xs)))
    for k in [0, 1, 12345678, n // 2]:
        X = dft_bin_1d(xs, k)
        c = 0.5 / n * (1 if k in (0, n // 2) else 2)
        delta = bin_tol(prec) * (scale + abs(X))  # bound on |X_got - X|
        assert abs(got[k] - c * abs(X) ** 2) <= c * delta * (2 * abs(X) + delta), (prec, k, got[k])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rfft_irfft_2_24(dev, oracle, prec):
    D, torch = dev
    n = 1 << 24
    x = signal(n, seed=3)
    rt = torch.float64 if prec == "f64" else torch.float32
    ct = torch.complex128 if prec == "f64" else torch.complex64
    p = D.R2CPlan(n, 1, rt)
    xt = torch.from_numpy(x).to("cuda", rt)
    X = p.forward(xt)
    back = p.inverse(X).cpu().numpy().astype(np.float64)
    Xh = X.cpu().numpy().astype(np.complex128)
    xs = x if prec == "f64" else x.astype(np.float32).astype(np.float64)
    bar = 1e-14 * 24 if prec == "f64" else 1e-6
    # the whole spectrum against the oracle's RealPlan.RFFT (realplan.go:77-116)
    want = oracle.rfft(xs)
    assert rel_l2(Xh, want) <= bar, (prec, rel_l2(Xh, want))
    # the inverse of the oracle's spectrum against the oracle's IRFFT
    Wd = torch.from_numpy(want).to("cuda", ct)
    got_i = p.inverse(Wd).cpu().numpy().astype(np.float64)
    want_i = oracle.irfft(want if prec == "f64" else want.astype(np.complex64).astype(np.complex128), n)
    assert rel_l2(got_i, want_i) <= bar, (prec, rel_l2(got_i, want_i))
    assert rel_l2(back, xs) <= (1e-14 * 24 if prec == "f64" else 1e-6)
    scale = math.sqrt(float(np.sum(xs * xs)))
    for k in [0, 3, 4999999, n // 2]:
        want = dft_bin_1d(xs, k)
        assert abs(Xh[k] - want) <= bin_tol(prec) * (scale + abs(want)), (prec, k)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_fft2_4096(dev, oracle, prec):
    D, torch = dev
    from fft_amd import _lib as L
    rows = cols = 4096
    rng = np.random.default_rng(4)
    x = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    ct = torch.complex128 if prec == "c128" else torch.complex64
    xs = x if prec == "c128" else x.astype(np.complex64).astype(np.complex128)
    p = D.NDPlan((rows, cols), L.ND_C2C, 1, ct)
    xt = torch.from_numpy(x.ravel()).to("cuda", ct)
    y = p.exec(xt)
    back = p.exec(y, inverse=True).cpu().numpy().astype(np.complex128)
    Y = y.cpu().numpy().astype(np.complex128).reshape(rows, cols)
    assert rel_l2(back, xs.ravel()) <= (1e-14 * 24 if prec == "c128" else 1e-6)
    # the whole transform against the oracle's FFT2 (ndfft.go:22-45)
    bar = 1e-14 * 24 if prec == "c128" else 1e-6
    want = oracle.fftn(xs.ravel(), (rows, cols))
    assert rel_l2(Y, want) <= bar, (prec, rel_l2(Y, want))
    del want
    wanti = oracle.fftn(xs.ravel(), (rows, cols), inverse=True)
    goti = p.exec(xt, inverse=True).cpu().numpy().astype(np.complex128)
    assert rel_l2(goti, wanti) <= bar, (prec, rel_l2(goti, wanti))
    scale = math.sqrt(float(np.sum(np.abs(xs) ** 2)))
    for u, v in [(0, 0), (1, 4095), (2049, 77)]:
        want = dft_bin_2d(xs, u, v)
        assert abs(Y[u, v] - want) <= bin_tol(prec) * (scale + abs(want)), (prec, u, v)


def test_fftn_256_cubed(dev, oracle):
    D, torch = dev
    from fft_amd import _lib as L
    shape = (256, 256, 256)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    p = D.NDPlan(shape, L.ND_C2C, 1, torch.complex64)
    xt = torch.from_numpy(x.ravel()).to("cuda")
    y = p.exec(xt)
    back = p.exec(y, inverse=True).cpu().numpy()
    assert rel_l2(back, x.ravel()) <= 1e-6
    Y = y.cpu().numpy().astype(np.complex128).reshape(shape)
    want = oracle.fftn(x.astype(np.complex128).ravel(), shape).reshape(shape)
    assert rel_l2(Y, want) <= 1e-6
    assert rel_l2(want, np.fft.fftn(x.astype(np.complex128))) <= 1e-13  # the oracle itself


def test_rfft2_8192(dev, oracle):
    D, torch = dev
    from fft_amd import _lib as L
    rows = cols = 8192
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (rows, cols)).astype(np.float32)
    p = D.NDPlan((rows, cols), L.ND_R2C, 1, torch.complex64)
    xt = torch.from_numpy(x.ravel()).to("cuda")
    X = p.rfft2(xt)
    back = p.irfft2(X).cpu().numpy()
    assert rel_l2(back, x.ravel()) <= 1e-6
    Xh = X.cpu().numpy().astype(np.complex128).reshape(rows, cols // 2 + 1)
    xs = x.astype(np.float64)
    # the whole half spectrum against the oracle's RFFT2 (rfft2.go:22-74), and
    # the inverse of the oracle's spectrum against its IRFFT2 (rfft2.go:87-138)
    want = oracle.rfft2(xs.ravel(), (rows, cols))
    assert rel_l2(Xh, want) <= 1e-6, rel_l2(Xh, want)
    Wd = torch.from_numpy(want.astype(np.complex64)).to("cuda")
    got_i = p.irfft2(Wd).cpu().numpy().astype(np.float64)
    want_i = oracle.irfft2(want.astype(np.complex64).astype(np.complex128), (rows, cols))
    del want, Wd
    assert rel_l2(got_i, want_i) <= 1e-6, rel_l2(got_i, want_i)
    scale = math.sqrt(float(np.sum(xs * xs)))
    for u, v in [(0, 0), (4096, 4096), (123, 4000)]:
        want = dft_bin_2d(xs, u, v)
        assert abs(Xh[u, v] - want) <= bin_tol("f32") * (scale + abs(want)), (u, v, Xh[u, v], want)
