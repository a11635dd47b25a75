"""GPU parity: complex transforms through the C ABI against the oracle.

Bars: the plain complex128 path within rel-L2 1e-14*log2(n) of the oracle
(which restates the reference); the complex64 path within rel-L2 1e-6 of the
FP64 oracle.
"""
import math

import numpy as np
import pytest

from conftest import cgauss, cmplx_signal, rel_l2

pytestmark = pytest.mark.gpu


def bound64(n):
    return 1e-14 * max(1.0, math.log2(n))


BOUND32 = 1e-6


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


def _run(D, torch, n, batch, x, dtype, inverse=False, flags=0):
    p = D.C2CPlan(n, batch, dtype, flags)
    xt = torch.from_numpy(x).to("cuda")
    y = p.exec(xt, inverse=inverse)
    torch.cuda.synchronize()
    return y.cpu().numpy()


POW2 = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 65536, 1 << 18, 1 << 20]
SMOOTH = [3, 5, 6, 7, 9, 10, 11, 12, 13, 15, 21, 49, 60, 100, 121, 169, 1000, 1296, 6720, 1008, 20580, 9408, 20160,
          3 * 5 * 7 * 11 * 13, 729, 2 * 3 * 5 * 7 * 64 * 3]
RADER = [769, 1009, 2017, 5003, 9973, 10007, 20011]
BLUE = [17, 19, 23, 34, 46, 101, 289, 641]


@pytest.mark.parametrize("n", POW2 + SMOOTH + RADER + BLUE)
def test_c2c_fp64_vs_oracle(dev, oracle, n):
    D, torch = dev
    batch = 3
    x = cgauss(n * batch, seed=n).reshape(batch, n)
    want = oracle.fft_batch(x)
    got = _run(D, torch, n, batch, x.ravel(), torch.complex128).reshape(batch, n)
    assert rel_l2(got, want) <= bound64(n), (n, rel_l2(got, want))
    wanti = oracle.fft_batch(x, inverse=True)
    goti = _run(D, torch, n, batch, x.ravel(), torch.complex128, inverse=True).reshape(batch, n)
    assert rel_l2(goti, wanti) <= bound64(n), (n, rel_l2(goti, wanti))


@pytest.mark.parametrize("n", [4096, 65536, 6720, 1009, 10007, 20014, 1 << 20, 641, 1296, 10006, 9973, 5003,
                               20011, 39239])
def test_c2c_fp32_vs_fp64_oracle(dev, oracle, n):
    D, torch = dev
    batch = 2
    x = cgauss(n * batch, seed=n + 1).reshape(batch, n)
    want = oracle.fft_batch(x)
    got = _run(D, torch, n, batch, x.astype(np.complex64).ravel(), torch.complex64).reshape(batch, n)
    e = rel_l2(got, want)
    assert e <= BOUND32, (n, e)
    goti = _run(D, torch, n, batch, x.astype(np.complex64).ravel(), torch.complex64, inverse=True)
    assert rel_l2(goti.reshape(batch, n), oracle.fft_batch(x, inverse=True)) <= BOUND32


@pytest.mark.parametrize("n", [17, 641, 2000, 20014, 2 * 9973, 40000])
def test_bluestein_reference_parity(dev, oracle, n):
    # Plain complex128 Bluestein reproduces the reference's recurrence-twiddle
    # rounding (~1e-12 away from the exact DFT at n = 20014), so it matches
    # the oracle to the 1e-14*log2(n) bar in both directions.
    D, torch = dev
    x = cgauss(2 * n, seed=n).reshape(2, n)
    for inv in (False, True):
        got = _run(D, torch, n, 2, x.ravel(), torch.complex128, inverse=inv).reshape(2, n)
        want = oracle.fft_batch(x, inverse=inv)
        assert rel_l2(got, want) <= bound64(n), (n, inv, rel_l2(got, want))


def test_bluestein_accurate_flag(dev, oracle):
    # With accurate twiddles the same length reaches the exact DFT to ~1e-15.
    from fft_amd import _lib as L
    D, torch = dev
    n = 20014
    x = cgauss(n, seed=5)
    got = _run(D, torch, n, 1, x, torch.complex128, flags=L.FLAG_ACCURATE_TWIDDLES)
    assert rel_l2(got, np.fft.fft(x)) <= bound64(n) / 10
    assert rel_l2(oracle.fft(x), np.fft.fft(x)) > 3e-13  # the reference's own error


def test_config1_roundtrip(dev, oracle):
    # N = 4096, batch 256, re/im N(0,1) seed 0: FFT then IFFT.
    D, torch = dev
    n, batch = 4096, 256
    x = cgauss(n * batch, seed=0).reshape(batch, n)
    p = D.C2CPlan(n, batch, torch.complex128)
    xt = torch.from_numpy(x.ravel()).to("cuda")
    y = p.exec(xt)
    z = p.exec(y, inverse=True)
    torch.cuda.synchronize()
    y, z = y.cpu().numpy().reshape(batch, n), z.cpu().numpy().reshape(batch, n)
    assert rel_l2(y, oracle.fft_batch(x)) <= bound64(n)
    assert rel_l2(z, x) <= bound64(n)


def test_aliasing_and_trivial(dev, oracle):
    D, torch = dev
    for n in [8, 12, 1296, 23, 1009, 65536]:  # plan_test.go:70-78, rader_test.go:118-126
        x = cmplx_signal(n)
        p = D.C2CPlan(n, 1, torch.complex128)
        t = torch.from_numpy(x.copy()).to("cuda")
        p.exec(t, out=t)
        torch.cuda.synchronize()
        assert rel_l2(t.cpu().numpy(), oracle.fft(x)) <= bound64(n), n
    for n in [0, 1]:  # plan_test.go:81-98
        x = np.array([7 - 3j])[:n]
        p = D.C2CPlan(n, 1, torch.complex128)
        if n:
            t = torch.from_numpy(x).to("cuda")
            assert p.exec(t).cpu().numpy()[0] == x[0]
            assert p.exec(t, inverse=True).cpu().numpy()[0] == x[0]


def test_mixed_length_batch(dev, oracle):
    # Many uneven lengths in one plan: the mixed-length workload's length set.
    D, torch = dev
    rng = np.random.default_rng(7)
    choices = [1024, 4096, 65536, 6720, 1009, 10007, 20014]
    lengths = list(rng.choice(choices, size=40)) + choices
    total = int(sum(lengths))
    x = cgauss(total, seed=11)
    for dtype, bar in [(torch.complex128, None), (torch.complex64, BOUND32)]:
        p = D.ManyPlan(lengths, dtype)
        xt = torch.from_numpy(x.astype(np.complex128 if dtype == torch.complex128 else np.complex64)).to("cuda")
        y = p.exec(xt)
        torch.cuda.synchronize()
        y = y.cpu().numpy()
        off = 0
        worst = {}
        for n in lengths:
            seg = x[off: off + n]
            want = oracle.fft(seg)
            e = rel_l2(y[off: off + n], want)
            worst[n] = max(worst.get(n, 0), e)
            off += n
        for n, e in worst.items():
            assert e <= (bar if bar else bound64(n)), (dtype, n, e)


def test_vjp_c2c_adjoint_identity(dev):
    # <F x, y> = <x, F^H y>: the VJP of y = FFT(x) is gx = F^H gy.
    D, torch = dev
    for n in [1024, 6720, 1009, 20014]:
        x = cgauss(n, seed=1)
        gy = cgauss(n, seed=2)
        p = D.C2CPlan(n, 1, torch.complex128)
        fx = p.exec(torch.from_numpy(x).to("cuda")).cpu().numpy()
        gx = p.vjp(torch.from_numpy(gy).to("cuda")).cpu().numpy()
        lhs = np.vdot(gy, fx)
        rhs = np.vdot(gx, x)
        # 20014 reproduces the reference's ~1e-12 Bluestein rounding
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), n
