"""GPU parity: vector-Jacobian products of the real and N-D transforms.

Convention: for a real loss L and a complex output y, the upstream gradient is
gy = dL/dRe(y) + i dL/dIm(y), so dL = Re sum conj(gy) dy. Each VJP is checked
against a reference computed without the weighting formula the library uses:
the adjoint sum written out with exact integer phases, a zero-padded numpy
inverse, or central finite differences (eps = 1e-3) of the oracle's FP64
transforms. The blur loss of the differentiation workload is checked against
2*A(Ax - y), which needs forward transforms only because the Gaussian blur A is
symmetric.
"""
import math

import numpy as np
import pytest

from conftest import cgauss, rel_l2, rgauss

pytestmark = pytest.mark.gpu

EPS = 1e-3


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


def bound64(n, oracle):
    # Bluestein-routed lengths reproduce the reference's own rounding.
    return 5e-12 if oracle.route(n) == 4 else 1e-13 * max(1.0, math.log2(n))


def adj_rfft_exact(gy, n):
    """gx_j = Re sum_{k<=n/2} gy_k exp(+2 pi i j k / n), exact integer phases."""
    nb = n // 2 + 1
    if n <= 2048:
        j = np.arange(n)[:, None]
        k = np.arange(nb)[None, :]
        ph = (j * k) % n
        E = np.exp(2j * np.pi * ph.astype(np.float64) / n)
        return np.real(E @ gy)
    z = np.zeros(n, dtype=np.complex128)
    z[:nb] = gy
    return np.real(np.fft.ifft(z) * n)


RLEN = [2, 7, 8, 17, 64, 1000, 1024, 2046, 20014, 1 << 17]


@pytest.mark.parametrize("n", RLEN)
def test_vjp_rfft(dev, oracle, n):
    D, torch = dev
    nb = n // 2 + 1
    gy = cgauss(2 * nb, seed=n).reshape(2, nb)
    want = np.stack([adj_rfft_exact(g, n) for g in gy])
    p = D.R2CPlan(n, 2, torch.float64)
    got = p.vjp_forward(torch.from_numpy(gy.ravel()).to("cuda")).cpu().numpy().reshape(2, n)
    assert rel_l2(got, want) <= bound64(n, oracle), (n, rel_l2(got, want))
    p32 = D.R2CPlan(n, 2, torch.float32)
    got32 = p32.vjp_forward(torch.from_numpy(gy.astype(np.complex64).ravel()).to("cuda")).cpu().numpy()
    assert rel_l2(got32.reshape(2, n), want) <= 1e-6, n
    # finite differences of L(x) = Re <gy, RFFT(x)> through the oracle; L is
    # linear, so its gradient is the same at every x, and at x = 0 the
    # differences carry no cancellation
    x = np.zeros(n)
    rng = np.random.default_rng(n)
    for j in rng.choice(n, size=min(n, 8), replace=False):
        e = np.zeros(n)
        e[j] = EPS
        lp = np.real(np.vdot(gy[0], oracle.rfft(x + e)))
        lm = np.real(np.vdot(gy[0], oracle.rfft(x - e)))
        fd = (lp - lm) / (2 * EPS)
        assert abs(fd - got[0, j]) <= 1e-9 * max(1.0, abs(fd)), (n, j, fd, got[0, j])


@pytest.mark.parametrize("n", [2, 7, 8, 17, 64, 1000, 1024, 2046, 20014, 1 << 17])
def test_vjp_irfft(dev, oracle, n):
    D, torch = dev
    nb = n // 2 + 1
    gx = rgauss(2 * n, seed=n + 3).reshape(2, n)
    p = D.R2CPlan(n, 2, torch.float64)
    got = p.vjp_inverse(torch.from_numpy(gx.ravel()).to("cuda")).cpu().numpy().reshape(2, nb)
    # central differences of L(X) = <gx, IRFFT(X)> in Re and Im of each bin
    X = cgauss(nb, seed=2)
    rng = np.random.default_rng(n)
    ks = sorted(set([0, nb - 1] + list(rng.choice(nb, size=min(nb, 10), replace=False))))
    for k in ks:
        for part in (1.0, 1j):
            e = np.zeros(nb, dtype=np.complex128)
            e[k] = EPS * part
            fd = (np.dot(gx[0], oracle.irfft(X + e, n)) - np.dot(gx[0], oracle.irfft(X - e, n))) / (2 * EPS)
            g = got[0, k].real if part == 1.0 else got[0, k].imag
            assert abs(fd - g) <= 1e-9 * max(1.0, np.abs(got[0]).max()), (n, k, part, fd, g)
    # FP32 against the FP64 result
    p32 = D.R2CPlan(n, 2, torch.float32)
    got32 = p32.vjp_inverse(torch.from_numpy(gx.astype(np.float32).ravel()).to("cuda")).cpu().numpy()
    assert rel_l2(got32.reshape(2, nb), got) <= 1e-6, n
    # DC and (even n) Nyquist gradients are real: the IRFFT ignores their imaginary parts
    assert np.all(got[:, 0].imag == 0)
    if n % 2 == 0:
        assert np.all(got[:, -1].imag == 0)


def adj_rfft2_exact(gy, rows, cols):
    full = np.zeros((rows, cols), dtype=np.complex128)
    full[:, : cols // 2 + 1] = gy
    return np.real(np.fft.ifft2(full) * rows * cols)


@pytest.mark.parametrize("shape", [(8, 16), (7, 10), (64, 64), (16, 1000), (256, 256), (5, 9)])
def test_vjp_rfft2_irfft2(dev, oracle, shape):
    D, torch = dev
    from fft_amd import _lib as L
    rows, cols = shape
    rc = cols // 2 + 1
    batch = 2
    gy = cgauss(batch * rows * rc, seed=rows + cols).reshape(batch, rows, rc)
    p = D.NDPlan(shape, kind=L.ND_R2C, batch=batch, dtype=torch.complex128)
    out = torch.empty(batch * rows * cols, dtype=torch.float64, device="cuda")
    got = p.vjp(torch.from_numpy(gy.ravel()).to("cuda"), out, direction=0).cpu().numpy()
    got = got.reshape(batch, rows, cols)
    want = np.stack([adj_rfft2_exact(g, rows, cols) for g in gy])
    assert rel_l2(got, want) <= 1e-13 * math.log2(rows * cols) + 5e-12 * (oracle.route(cols) == 4), shape
    # IRFFT2 VJP by central differences of L(X) = <gx, IRFFT2(X)> (oracle, FP64)
    gx = rgauss(batch * rows * cols, seed=rows * cols).reshape(batch, rows, cols)
    outX = torch.empty(batch * rows * rc, dtype=torch.complex128, device="cuda")
    gX = p.vjp(torch.from_numpy(gx.ravel()).to("cuda"), outX, direction=1).cpu().numpy()
    gX = gX.reshape(batch, rows, rc)
    X = cgauss(rows * rc, seed=9)
    rng = np.random.default_rng(rows * 31 + cols)
    for idx in rng.choice(rows * rc, size=min(rows * rc, 12), replace=False):
        for part in (1.0, 1j):
            e = np.zeros(rows * rc, dtype=np.complex128)
            e[idx] = EPS * part
            lp = np.dot(gx[1].ravel(), oracle.irfft2(X + e, shape))
            lm = np.dot(gx[1].ravel(), oracle.irfft2(X - e, shape))
            fd = (lp - lm) / (2 * EPS)
            g = gX[1].ravel()[idx]
            g = g.real if part == 1.0 else g.imag
            assert abs(fd - g) <= 1e-9 * max(1.0, np.abs(gX[1]).max()), (shape, idx, part, fd, g)


def gaussian_h(rows, cols, sigma):
    fr = np.fft.fftfreq(rows)[:, None]
    fc = np.fft.rfftfreq(cols)[None, :]
    return np.exp(-2.0 * (np.pi * sigma) ** 2 * (fr ** 2 + fc ** 2))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_blur_loss_gradient(dev, dtype):
    # Differentiation workload at its full size, 64 images of 1024 x 1024: loss
    # ||IRFFT2(RFFT2(x) H) - y||^2 with a Gaussian (sigma 3 px) H; the gradient
    # from the adjoint kernels against 2 A (A x - y) (A symmetric) in FP64 and
    # against central differences of the FP64 loss on 32 pixels.
    D, torch = dev
    from fft_amd import _lib as L
    rows = cols = 1024
    batch = 64
    rc = cols // 2 + 1
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, size=(batch, rows, cols))
    H = gaussian_h(rows, cols, 3.0)

    def A(v):
        return np.fft.irfft2(np.fft.rfft2(v) * H, s=(rows, cols))

    y = A(rng.uniform(0, 1, size=(batch, rows, cols))) + rng.normal(0, 0.01, size=(batch, rows, cols))
    want = 2.0 * A(A(x) - y)

    rt, ct = (torch.float64, torch.complex128) if dtype == "f64" else (torch.float32, torch.complex64)
    p = D.NDPlan((rows, cols), kind=L.ND_R2C, batch=batch, dtype=ct)
    xt = torch.from_numpy(x.ravel()).to("cuda", rt)
    yt = torch.from_numpy(y.ravel()).to("cuda", rt)
    Ht = torch.from_numpy(np.tile(H.ravel(), batch)).to("cuda", rt)
    S = p.rfft2(xt)
    r = p.irfft2(S * Ht) - yt
    gS = p.vjp(2.0 * r, torch.empty_like(S), direction=1)
    gx = p.vjp(gS * Ht, torch.empty_like(xt), direction=0)
    torch.cuda.synchronize()
    got = gx.cpu().numpy().astype(np.float64).reshape(batch, rows, cols)
    tol = 1e-12 if dtype == "f64" else 2e-6
    assert rel_l2(got, want) <= tol, (dtype, rel_l2(got, want))

    def loss_b(v, b):
        # the loss terms of image b, the only ones a pixel of image b moves
        return float(np.sum((A(v) - y[b]) ** 2))

    for flat in rng.choice(batch * rows * cols, size=32, replace=False):
        b, rem = divmod(int(flat), rows * cols)
        i, j = divmod(rem, cols)
        e = np.zeros((rows, cols))
        e[i, j] = EPS
        fd = (loss_b(x[b] + e, b) - loss_b(x[b] - e, b)) / (2 * EPS)
        assert abs(fd - want[b, i, j]) <= 1e-6 * max(1.0, np.abs(want).max()), (b, i, j, fd, want[b, i, j])
        assert abs(fd - got[b, i, j]) <= (1e-6 if dtype == "f64" else 1e-4) * max(1.0, np.abs(want).max())


def test_vjp_fftn_adjoint_identity(dev):
    # <F x, y> = <x, F^H y> for FFTN (direction 0) and IFFTN (direction 1).
    D, torch = dev
    for shape in [(8, 12), (4, 5, 6), (64, 64), (16, 1009)]:
        n = int(np.prod(shape))
        x = cgauss(n, seed=n)
        gy = cgauss(n, seed=n + 1)
        p = D.NDPlan(shape, dtype=torch.complex128)
        for inv in (False, True):
            fx = p.exec(torch.from_numpy(x).to("cuda"), inverse=inv).cpu().numpy()
            gx = p.vjp_c2c(torch.from_numpy(gy).to("cuda"), inverse=inv).cpu().numpy()
            lhs = np.vdot(gy, fx)
            rhs = np.vdot(gx, x)
            assert abs(lhs - rhs) <= 1e-12 * abs(lhs) * math.log2(n), (shape, inv)


def test_host_staged_vjps(dev):
    # The host-staged VJP entry points agree with the device-resident ones.
    import ctypes
    from fft_amd import _lib as L
    D, torch = dev
    n, batch = 1000, 3
    nb = n // 2 + 1
    gy = cgauss(batch * nb, seed=4)
    gx = rgauss(batch * n, seed=5)
    gc = cgauss(batch * n, seed=6)
    p = D.R2CPlan(n, batch, torch.float64)
    out = np.zeros(batch * n)
    L.check(L.lib().gofft_vjp_r2c_host(p.handle, gy.ctypes.data_as(ctypes.c_void_p),
                                       out.ctypes.data_as(ctypes.c_void_p)))
    assert np.array_equal(out, p.vjp_forward(torch.from_numpy(gy).to("cuda")).cpu().numpy())
    outc = np.zeros(batch * nb, dtype=np.complex128)
    L.check(L.lib().gofft_vjp_c2r_host(p.handle, gx.ctypes.data_as(ctypes.c_void_p),
                                       outc.ctypes.data_as(ctypes.c_void_p)))
    assert np.array_equal(outc, p.vjp_inverse(torch.from_numpy(gx).to("cuda")).cpu().numpy())
    pc = D.C2CPlan(n, batch, torch.complex128)
    outg = np.zeros(batch * n, dtype=np.complex128)
    L.check(L.lib().gofft_vjp_c2c_host(pc.handle, gc.ctypes.data_as(ctypes.c_void_p),
                                       outg.ctypes.data_as(ctypes.c_void_p), L.FORWARD))
    assert np.array_equal(outg, pc.vjp(torch.from_numpy(gc).to("cuda")).cpu().numpy())
