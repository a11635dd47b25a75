"""Pins the oracle's floating-point transforms: long-double naive DFT,
numpy/pocketfft FP64, the reference's closed-form tests, and properties
(round trips, linearity, Parseval, parallel == serial). CPU only."""
import math

import numpy as np
import pytest

from conftest import cgauss, cmplx_signal, real_signal, rel_l2, rgauss

# plan_test.go:14-19, rader_test.go:13-15, radix7_test.go:45
MIXED = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 18, 20, 21, 24, 25, 27, 30, 32, 33, 35, 36,
         45, 48, 49, 50, 60, 64, 72, 81, 100, 105, 121, 125, 128, 169, 1000, 1296, 17, 19, 23, 29, 31,
         34, 46, 289, 101, 97, 257, 521, 769, 1009, 14, 28, 63, 70, 98, 147, 245, 343]


def bound(n):
    return 1e-14 * max(1.0, math.log2(n))


@pytest.mark.parametrize("n", MIXED)
def test_against_long_double_dft(oracle, n):
    x = cmplx_signal(n)
    for inv in (False, True):
        want = oracle.naive_dft(x, inverse=inv)
        got = oracle.fft(x, inverse=inv) * (n if inv else 1)
        assert rel_l2(got, want) <= bound(n), (n, inv, oracle.route(n))


@pytest.mark.parametrize("n", [1024, 4096, 65536, 6720, 1009, 10007, 2017, 9973, 5003, 1 << 20])
def test_against_numpy(oracle, n):
    x = cgauss(n, seed=n)
    assert rel_l2(oracle.fft(x), np.fft.fft(x)) <= bound(n)
    assert rel_l2(oracle.ifft(x), np.fft.ifft(x)) <= bound(n)


def test_bluestein_recurrence_error_is_restated(oracle):
    # The reference convolves with recurrence twiddles (kernels.go:38-59); at
    # N = 2*10007 (m = 65536) that leaves ~1e-12 relative error against an exact
    # DFT, while exact twiddles give ~1e-15. A restatement with direct twiddles
    # would fail the lower bound.
    n = 20014
    x = cgauss(n, seed=0)
    e = rel_l2(oracle.fft(x), np.fft.fft(x))
    assert 3e-13 < e < 5e-12, e
    # Small Bluestein lengths stay accurate.
    for n in [17, 23, 641]:
        x = cgauss(n, seed=n)
        assert rel_l2(oracle.fft(x), np.fft.fft(x)) < 1e-13


def test_go_sincos(oracle):
    # Go's math.Sin(math.Pi) and math.Cos(math.Pi/2) print these (Cephes
    # reduction), unlike glibc's 1.2246467991473532e-16 / 6.123233995736766e-17.
    assert oracle.go_sincos(math.pi)[0] == 1.2246467991473515e-16
    assert oracle.go_sincos(math.pi / 2)[1] == 6.123233995736757e-17
    # The per-stage angles of kernels.Radix2 (+-2*pi/L, L = 8 .. 2^20): within
    # one ulp of the correctly rounded value (long double). Cephes is not
    # correctly rounded: at L = 64 its cosine is one ulp low, which is why the
    # oracle restates Go's sin/cos instead of calling the C library.
    off = []
    for k in range(3, 21):
        a = 2 * math.pi / (1 << k)
        s, c = oracle.go_sincos(-a)
        sl, cl = float(-np.sin(np.longdouble(a))), float(np.cos(np.longdouble(a)))
        assert abs(s - sl) <= np.spacing(abs(sl)) and abs(c - cl) <= np.spacing(cl), k
        if (s, c) != (sl, cl):
            off.append(1 << k)
    assert 64 in off
    # accuracy over a sweep of arguments
    xs = np.linspace(-math.pi, math.pi, 2001)
    for x in xs:
        s, c = oracle.go_sincos(x)
        assert abs(s - math.sin(x)) <= 2.3e-16 and abs(c - math.cos(x)) <= 2.3e-16


def test_radix2_recurrence_twiddles(oracle):
    for L in [8, 1024, 65536]:
        w = oracle.radix2_stage_twiddles(L)
        exact = np.exp(-2j * np.pi * np.arange(L // 2) / L)
        err = np.abs(w - exact).max()
        assert err < 1e-11
        if L == 65536:
            assert err > 1e-14  # the recurrence drifts; direct twiddles would not
        wi = oracle.radix2_stage_twiddles(L, inverse=True)
        assert np.array_equal(wi, np.conj(w))  # inverse recurrence is the exact conjugate
    x = cgauss(64, seed=1)
    y = oracle.radix2_rec(oracle.bit_reverse(x))
    assert rel_l2(y, np.fft.fft(x)) < 5e-15


# --- closed forms of the reference tests -----------------------------------


def test_impulse_constant_sinusoid(oracle):
    for n in [4, 6, 8, 13, 9, 15, 25, 23, 49]:  # fft_test.go:63-75, plan_test.go:141-152
        x = np.zeros(n, complex)
        x[0] = 1
        assert np.abs(oracle.fft(x) - 1).max() < 1e-9
    for n in [4, 7, 8]:  # fft_test.go:77-89
        X = oracle.fft(np.ones(n, complex))
        want = np.zeros(n, complex)
        want[0] = n
        assert np.abs(X - want).max() < 1e-9
    n, f = 16, 3  # fft_test.go:91-111
    x = np.exp(2j * np.pi * f * np.arange(n) / n)
    mag = np.abs(oracle.fft(x))
    assert abs(mag[f] - n) < 1e-9 and np.delete(mag, f).max() < 1e-9


def test_linearity_and_edges(oracle):
    n = 10  # fft_test.go:113-134
    a, b = 2 - 1j, 0.5 + 3j
    i = np.arange(n)
    x = i - 0.2j * i
    y = np.cos(i) + 1j * (i % 4)
    assert np.abs(oracle.fft(a * x + b * y) - (a * oracle.fft(x) + b * oracle.fft(y))).max() < 1e-9
    assert len(oracle.fft(np.zeros(0, complex))) == 0
    assert oracle.fft(np.array([7 - 3j]))[0] == 7 - 3j
    assert oracle.ifft(np.array([7 - 3j]))[0] == 7 - 3j


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 16, 25, 32, 63, 100, 4096, 6720, 10007, 20014])
def test_round_trip(oracle, n):
    x = cgauss(n, seed=3)
    assert rel_l2(oracle.ifft(oracle.fft(x)), x) < 1e-11


def test_rfft_closed_forms(oracle):
    for n in [2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 17, 31, 32, 1000, 1024, 1 << 16]:
        x = real_signal(n)
        X = oracle.rfft(x)
        assert len(X) == n // 2 + 1
        assert rel_l2(X, np.fft.rfft(x)) <= bound(n)
        assert rel_l2(X, oracle.fft(x.astype(complex))[: n // 2 + 1]) <= bound(n)
        assert rel_l2(oracle.irfft(X, n), x) <= bound(n)
    n, f = 16, 3  # rfft_test.go:82-100
    X = oracle.rfft(np.cos(2 * np.pi * f * np.arange(n) / n))
    mag = np.abs(X)
    assert abs(mag[f] - n / 2) < 1e-9 and np.delete(mag, f).max() < 1e-9
    # rfft_test.go:105-124: Hermitian mirror for n = 6
    half = np.array([3, 1 - 2j, -0.5 + 0.4j, 2])
    full = np.concatenate([half, np.conj(half[1:3][::-1])])
    assert np.abs(oracle.irfft(half, 6) - np.fft.ifft(full).real).max() < 1e-12
    # short spectrum zero-pads (rfft_test.go:206-212), imaginary DC/Nyquist ignored
    assert np.allclose(oracle.irfft(np.array([2 + 0j]), 4), oracle.irfft(np.array([2, 0, 0]), 4))
    s = oracle.rfft(real_signal(8))
    s2 = s.copy()
    s2[0] += 5j
    s2[4] += 3j
    assert np.array_equal(oracle.irfft(s, 8), oracle.irfft(s2, 8))
    assert len(oracle.rfft(np.zeros(0))) == 0 and oracle.rfft(np.array([5.0]))[0] == 5
    assert len(oracle.irfft(np.array([1 + 0j]), 0)) == 0
    assert abs(oracle.irfft(np.array([4 + 0j]), 1)[0] - 4) < 1e-12


def test_psd_parseval(oracle):
    # freq_test.go:78-130
    for n, d in [(32, 0.1), (15, 1.0)]:
        i = np.arange(n)
        if n == 32:
            x = np.sin(2 * np.pi * 3 * i / n) + 0.5 * np.cos(2 * np.pi * 7 * i / n)
        else:
            x = np.cos(2 * np.pi * 2 * i / n) + (i % 3)
        p = oracle.psd(x, d)
        f = oracle.rfftfreq(n, d)
        assert len(p) == n // 2 + 1
        assert abs(p.sum() * (f[1] - f[0]) - np.mean(x * x)) < 1e-9
    x = rgauss(1000, seed=2)
    X = np.fft.rfft(x)
    want = np.abs(X) ** 2 * 0.5 / 1000
    want[1:-1] *= 2
    assert rel_l2(oracle.psd(x, 0.5), want) < 1e-14


def test_spectrogram_matches_framed_psd(oracle):
    x = rgauss(5000, seed=4)
    w = oracle.window("hann", 256)
    S = oracle.spectrogram(x, 256, 100, w, 2.0)
    step = 156
    assert S.shape == ((5000 - 256) // step + 1, 129)
    for t in [0, 7, S.shape[0] - 1]:
        seg = x[t * step: t * step + 256] * w
        assert np.array_equal(S[t], oracle.psd(seg, 2.0))
    Sp = oracle.spectrogram(x, 256, 100, w, 2.0, threads=4)
    assert np.array_equal(S, Sp)


def test_windows_and_freqs(oracle):
    # window_test.go:40-99, freq_test.go:10-72
    assert np.allclose(oracle.window("hann", 5), [0, 0.5, 1, 0.5, 0], atol=1e-15)
    k = np.arange(5)
    assert np.allclose(oracle.window("hamming", 5), 0.54 - 0.46 * np.cos(2 * np.pi * k / 4), atol=1e-15)
    assert np.allclose(oracle.window("bartlett", 5), [0, 0.5, 1, 0.5, 0])
    assert np.allclose(oracle.window("bartlett", 4), [0, 2 / 3, 2 / 3, 0])
    assert np.allclose(oracle.window("hann", 1024), np.hanning(1024), atol=1e-15)
    assert np.allclose(oracle.window("blackman", 33), np.blackman(33), atol=1e-15)
    assert np.allclose(oracle.fftfreq(8, 1), [0, 0.125, 0.25, 0.375, -0.5, -0.375, -0.25, -0.125])
    assert np.allclose(oracle.fftfreq(5, 1), [0, 0.2, 0.4, -0.4, -0.2])
    assert np.allclose(oracle.fftfreq(4, 0.5), [0, 0.5, -1.0, -0.5])
    assert np.allclose(oracle.rfftfreq(8, 1), [0, 0.125, 0.25, 0.375, 0.5])
    assert np.allclose(oracle.rfftfreq(4, 0.25), [0, 1, 2])


# --- N-D ----------------------------------------------------------------------


@pytest.mark.parametrize("shape", [[2, 2], [3, 4], [4, 3], [1, 5], [5, 1], [2, 3, 4], [3, 1, 2],
                                   [2, 2, 2, 2], [6, 5], [17, 8], [64, 64]])
def test_fftn(oracle, shape):
    total = int(np.prod(shape))
    x = cgauss(total, seed=total)
    want = np.fft.fftn(x.reshape(shape)).ravel()
    assert rel_l2(oracle.fftn(x, shape), want) <= bound(total)
    assert rel_l2(oracle.fftn(oracle.fftn(x, shape), shape, inverse=True), x) <= bound(total)


def test_fftn_shape_panics(oracle):
    with pytest.raises(ValueError, match="shape product does not match"):
        oracle.fftn(np.zeros(10, complex), [3, 4])
    with pytest.raises(ValueError, match="shape lengths must be positive"):
        oracle.fftn(np.zeros(0, complex), [0, 4])
    with pytest.raises(ValueError, match="shape lengths must be positive"):
        oracle.fftn(np.zeros(4, complex), [-1, 4])
    x = np.array([3 - 2j])
    assert oracle.fftn(x, [])[0] == x[0]


@pytest.mark.parametrize("shape", [(4, 6), (3, 4), (5, 8), (2, 2), (4, 5), (1, 6), (6, 1), (7, 3), (64, 128)])
def test_rfft2(oracle, shape):
    r, c = shape
    x = rgauss(r * c, seed=r * c)
    got = oracle.rfft2(x, shape)
    assert rel_l2(got, np.fft.rfft2(x.reshape(shape)).ravel()) <= bound(r * c)
    assert rel_l2(oracle.irfft2(got, shape), x) <= bound(r * c)


def test_parallel_matches_serial(oracle):
    # parallel_test.go:64-106: line fan-out must not change a single bit
    x = cgauss(256 * 256, seed=9)
    rx = ((np.arange(256 * 256) * 7 + 1) % 13) * 0.1
    oracle.set_workers(4)
    par = oracle.fftn(x, [256, 256]), oracle.rfft2(rx, (256, 256))
    oracle.set_workers(1)
    ser = oracle.fftn(x, [256, 256]), oracle.rfft2(rx, (256, 256))
    oracle.set_workers(4)
    assert np.array_equal(par[0], ser[0]) and np.array_equal(par[1], ser[1])
