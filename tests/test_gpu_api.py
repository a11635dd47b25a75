"""The reference's own test semantics (fft_test.go, plan_test.go, rfft_test.go,
realplan_test.go, ndfft_test.go, rfft2_test.go, freq_test.go, window_test.go)
run against the GPU-backed mirror of its API, fft_amd.api."""
import math

import numpy as np
import pytest

from conftest import cmplx_signal, real_signal, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-9  # fft_test.go:9


@pytest.fixture(scope="module")
def api():
    import torch
    assert torch.cuda.is_available()
    from fft_amd import api
    return api


def close(got, want, tol=TOL):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape
    assert np.abs(got - want).max(initial=0) <= tol


def test_fft_against_naive(api, oracle):
    for n in [2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 17, 31, 64] + [1000, 1296, 769, 1009, 289, 101]:
        x = cmplx_signal(n)
        close(api.FFT(x), oracle.naive_dft(x), tol=1e-13 * n * n)


def test_round_trip_and_plan_reuse(api):
    for n in [1, 2, 3, 5, 8, 11, 16, 25, 32, 63, 100]:
        x = np.arange(n) * 0.5 - 1 + 1j * np.sin(np.arange(n))
        close(api.IFFT(api.FFT(x)), x)
    for n in [8, 9, 12, 1000, 1296, 23]:  # plan_test.go:46-67
        p = api.NewPlan(n)
        assert p.Len() == n
        x = cmplx_signal(n)
        dst = np.zeros(n, complex)
        p.FFT(dst, x)
        back = np.zeros(n, complex)
        p.IFFT(back, dst)
        close(back, x)
        dst2 = np.zeros(n, complex)
        p.FFT(dst2, x)
        close(dst2, dst, tol=0)


def test_alias_and_trivial(api, oracle):
    for n in [8, 12, 1296, 23]:
        x = cmplx_signal(n)
        want = oracle.fft(x)
        p = api.NewPlan(n)
        p.FFT(x, x)
        close(x, want)
    p0 = api.NewPlan(0)
    assert len(p0.FFT(np.zeros(0, complex), np.zeros(0, complex))) == 0
    one = np.array([7 - 3j])
    assert api.NewPlan(1).FFT(np.zeros(1, complex), one)[0] == one[0]
    assert api.NewPlan(1).IFFT(np.zeros(1, complex), one)[0] == one[0]
    assert len(api.FFT(np.zeros(0, complex))) == 0
    assert api.FFT(one)[0] == one[0]


def test_no_mutation(api):
    for n in [8, 10]:
        x = np.arange(n) - 1j * np.arange(n)
        orig = x.copy()
        api.FFT(x)
        api.IFFT(x)
        assert np.array_equal(x, orig)
    x = real_signal(8)
    orig = x.copy()
    spec = api.RFFT(x)
    so = spec.copy()
    api.IRFFT(spec, 8)
    assert np.array_equal(x, orig) and np.array_equal(spec, so)


def test_rfft_edges(api):
    for n in [2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 17, 31, 32]:
        x = real_signal(n)
        close(api.RFFT(x), api.FFTReal(x)[: n // 2 + 1])
    for n in [1, 2, 3, 4, 5, 8, 9, 12, 15, 16, 17, 32]:
        x = real_signal(n)
        close(api.IRFFT(api.RFFT(x), n), x)
    assert len(api.RFFT(np.zeros(0))) == 0
    assert api.RFFT(np.array([5.0]))[0] == 5
    assert len(api.IRFFT(np.array([1 + 0j]), 0)) == 0
    assert len(api.IRFFT(np.array([1 + 0j]), -3)) == 0
    assert abs(api.IRFFT(np.array([4 + 0j]), 1)[0] - 4) <= TOL
    close(api.IRFFT(np.array([2 + 0j]), 4), api.IRFFT(np.array([2, 0, 0], dtype=complex), 4))
    half = np.array([3, 1 - 2j, -0.5 + 0.4j, 2])  # rfft_test.go:105-124
    full = np.concatenate([half, np.conj(half[1:3][::-1])])
    close(api.IRFFT(half, 6), np.fft.ifft(full).real)
    for n in [2, 4, 6, 8, 10, 12, 16, 20, 24, 32, 36, 64, 100, 128]:  # packed vs full inverse
        x = real_signal(n)
        spec = api.RFFT(x)
        close(api.IRFFT(spec, n), x)
        assert abs(spec[0].imag) <= TOL and abs(spec[n // 2].imag) <= TOL
    rp = api.NewRealPlan(1296)
    assert rp.Len() == 1296
    x = real_signal(1296)
    got = rp.RFFT(np.zeros(649, complex), x)
    close(got, api.FFTReal(x)[:649])


def test_nd(api, oracle):
    for shape in [[2, 2], [3, 4], [4, 3], [1, 5], [5, 1], [2, 3, 4], [3, 1, 2], [2, 2, 2, 2], [6, 5]]:
        total = int(np.prod(shape))
        i = np.arange(total)
        x = np.sin(i * 0.7) + (i % 3) + 1j * (np.cos(i * 0.3) - (i % 2))
        close(api.FFTN(x, shape), oracle.fftn(x, shape))
        close(api.IFFTN(api.FFTN(x, shape), shape), x)
    x = np.array([3 - 2j])
    assert api.FFTN(x, [])[0] == x[0] and api.IFFTN(x, [])[0] == x[0]
    for shape, total in [([3, 4], 10), ([0, 4], 0), ([-1, 4], 4)]:
        with pytest.raises(api.Panic):
            api.FFTN(np.zeros(total, complex), shape)
    with pytest.raises(api.Panic, match="fft: shape product does not match len\\(data\\)"):
        api.FFTN(np.zeros(10, complex), [3, 4])
    with pytest.raises(api.Panic, match="fft: shape lengths must be positive"):
        api.FFTN(np.zeros(0, complex), [0, 4])


def test_rfft2(api, oracle):
    for shape in [(4, 6), (3, 4), (5, 8), (2, 2), (4, 5), (1, 6), (6, 1)]:
        r, c = shape
        i = np.arange(r * c)
        x = np.sin(i * 0.7) + 0.3 * np.cos(i * 1.9) + (i % 4)
        got = api.RFFT2(x, shape)
        full = api.FFT2(x.astype(complex), shape).reshape(r, c)
        close(got.reshape(r, c // 2 + 1), full[:, : c // 2 + 1])
        close(api.IRFFT2(got, shape), x)
    x = np.ones(12)
    X = api.RFFT2(x, (3, 4))
    want = np.zeros(len(X), complex)
    want[0] = 12
    close(X, want)
    full = api.RFFT2(real_signal(8), (2, 4))
    short = full[:-1]
    close(api.IRFFT2(short, (2, 4)), api.IRFFT2(np.concatenate([short, [0]]), (2, 4)))
    for fn in [lambda: api.RFFT2(np.zeros(4), (0, 4)), lambda: api.RFFT2(np.zeros(4), (2, -4)),
               lambda: api.RFFT2(np.zeros(5), (2, 4)), lambda: api.IRFFT2(np.zeros(4, complex), (0, 4)),
               lambda: api.IRFFT2(np.zeros(4, complex), (2, -4))]:
        with pytest.raises(api.Panic):
            fn()


def test_psd_spectrogram(api):
    n, d = 32, 0.1  # freq_test.go:78-104
    i = np.arange(n)
    x = np.sin(2 * np.pi * 3 * i / n) + 0.5 * np.cos(2 * np.pi * 7 * i / n)
    p = api.PSD(x, d)
    f = api.RFFTFreq(n, d)
    assert abs(p.sum() * (f[1] - f[0]) - np.mean(x * x)) <= TOL
    n = 15
    i = np.arange(n)
    x = np.cos(2 * np.pi * 2 * i / n) + (i % 3)
    p = api.PSD(x, 1.0)
    assert abs(p.sum() * (1.0 / n) - np.mean(x * x)) <= TOL
    assert len(api.PSD(np.zeros(0), 1)) == 0
    x = np.arange(20, dtype=float)  # freq_test.go:150-177
    frames = api.Spectrogram(x, 8, 4, api.Hann(8), 1.0)
    assert len(frames) == 4 and all(len(f) == 5 for f in frames)
    for t, fr in enumerate(frames):
        close(fr, api.PSD(x[t * 4: t * 4 + 8] * api.Hann(8), 1.0))
    assert api.Spectrogram(np.zeros(5), 8, 4, api.Hann(8), 1.0) is None
    with pytest.raises(api.Panic, match="fft: segment length must be positive"):
        api.Spectrogram(x, 0, 0, [], 1.0)
    with pytest.raises(api.Panic, match="fft: window length must equal segment length"):
        api.Spectrogram(x, 8, 0, api.Hann(4), 1.0)
    with pytest.raises(api.Panic, match="fft: overlap must satisfy 0 <= overlap < segment"):
        api.Spectrogram(x, 8, 8, api.Hann(8), 1.0)


def test_windows_and_freqs(api):
    close(api.Hann(5), [0, 0.5, 1, 0.5, 0])
    close(api.Bartlett(5), [0, 0.5, 1, 0.5, 0])
    close(api.Bartlett(4), [0, 2 / 3, 2 / 3, 0])
    for fn in [api.Hann, api.Hamming, api.Blackman, api.BlackmanHarris, api.Bartlett]:
        assert len(fn(0)) == 0 and len(fn(-3)) == 0 and fn(1)[0] == 1
        w = fn(9)
        close(w, w[::-1])
    close(api.FFTFreq(8, 1), [0, 0.125, 0.25, 0.375, -0.5, -0.375, -0.25, -0.125])
    close(api.FFTFreq(5, 1), [0, 0.2, 0.4, -0.4, -0.2])
    close(api.RFFTFreq(4, 0.25), [0, 1, 2])
    assert len(api.FFTFreq(0, 1)) == 0 and len(api.RFFTFreq(-5, 1)) == 0


def test_vjp_api_adjoint_identities(api):
    # For each transform T, its VJP G satisfies Re<T(x), g> = Re<x, G(g)>.
    rng = np.random.default_rng(3)

    def c(n):
        return rng.standard_normal(n) + 1j * rng.standard_normal(n)

    def re_dot(a, b):
        return float(np.real(np.vdot(b, a)))

    for n in [8, 17, 1000, 1009]:
        x, g = c(n), c(n)
        assert math.isclose(re_dot(api.FFT(x), g), re_dot(x, api.FFTVJP(g)), rel_tol=1e-11)
        assert math.isclose(re_dot(api.IFFT(x), g), re_dot(x, api.IFFTVJP(g)), rel_tol=1e-11)
        xr, gr = rng.standard_normal(n), rng.standard_normal(n)
        gb = c(n // 2 + 1)
        assert math.isclose(re_dot(api.RFFT(xr), gb), re_dot(xr, api.RFFTVJP(gb, n)), rel_tol=1e-11)
        assert math.isclose(re_dot(api.IRFFT(gb, n), gr), re_dot(gb, api.IRFFTVJP(gr)), rel_tol=1e-11)
    for shape in [(4, 6), (3, 5, 7)]:
        m = int(np.prod(shape))
        x, g = c(m), c(m)
        assert math.isclose(re_dot(api.FFTN(x, shape), g), re_dot(x, api.FFTNVJP(g, shape)), rel_tol=1e-11)
        assert math.isclose(re_dot(api.IFFTN(x, shape), g), re_dot(x, api.IFFTNVJP(g, shape)), rel_tol=1e-11)
    rows, cols = 6, 10
    xr, gr = rng.standard_normal(rows * cols), rng.standard_normal(rows * cols)
    gb = c(rows * (cols // 2 + 1))
    assert math.isclose(re_dot(api.RFFT2(xr, (rows, cols)), gb), re_dot(xr, api.RFFT2VJP(gb, (rows, cols))),
                        rel_tol=1e-11)
    assert math.isclose(re_dot(api.IRFFT2(gb, (rows, cols)), gr), re_dot(gb, api.IRFFT2VJP(gr, (rows, cols))),
                        rel_tol=1e-11)
    with pytest.raises(api.Panic, match="shape lengths must be positive"):
        api.RFFT2VJP(gb, (0, cols))
    with pytest.raises(api.Panic, match="gradient length"):
        api.RFFTVJP(gb, 7)
