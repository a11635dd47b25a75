"""GPU: gofft_filter2_loss_grad, the spectral filter loss
||IRFFT2(RFFT2(x) H) - y||^2 and its gradient with the forward recomputed per
chunk of arrays. Every chunk size gives the same bits as holding the whole
batch; the whole-batch result equals the same composition of the separate
RFFT2 / IRFFT2 / VJP entry points; and both precisions match the FP64 loss and
gradient 2 A (A x - y) (A = the filter, real and symmetric here) and central
differences."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    from fft_amd import _lib as L
    return D, L, torch


def gaussian_h(rows, cols, sigma):
    fr = np.fft.fftfreq(rows)[:, None]
    fc = np.fft.rfftfreq(cols)[None, :]
    return np.exp(-2.0 * (np.pi * sigma) ** 2 * (fr ** 2 + fc ** 2))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("shape,batch", [((256, 320), 12), ((1024, 1024), 64)])
def test_filter_loss_grad(dev, prec, shape, batch):
    D, L, torch = dev
    rows, cols = shape
    rt, ct = (torch.float32, torch.complex64) if prec == "f32" else (torch.float64, torch.complex128)
    rng = np.random.default_rng(rows + batch)
    x = rng.uniform(0, 1, (batch, rows, cols))
    Hf = gaussian_h(rows, cols, 3.0)

    def A(v):
        return np.fft.irfft2(np.fft.rfft2(v) * Hf, s=(rows, cols))

    y = A(rng.uniform(0, 1, (batch, rows, cols))) + rng.normal(0, 0.01, (batch, rows, cols))
    p = D.NDPlan(shape, L.ND_R2C, batch, ct)
    xt = torch.from_numpy(x.ravel()).to("cuda", rt)
    yt = torch.from_numpy(y.ravel()).to("cuda", rt)
    Ht = torch.from_numpy(Hf.ravel().astype(np.complex128)).to("cuda", ct)
    loss_all, g_all = p.filter_loss_grad(xt, Ht, yt, chunk=batch)
    for ch in (1, 5, 0):
        loss_c, g_c = p.filter_loss_grad(xt, Ht, yt, chunk=ch)
        assert torch.equal(loss_c, loss_all) and torch.equal(g_c, g_all), ch
    # the same composition through the separate entry points
    S = p.rfft2(xt)
    r = p.irfft2(S * Ht.repeat(batch)) - yt
    gS = p.vjp(2.0 * r, torch.empty_like(S), direction=1)
    g_ref = p.vjp(gS * Ht.conj().repeat(batch), torch.empty_like(xt), direction=0)
    tol_same = 1e-6 if prec == "f32" else 1e-13
    assert rel_l2(g_all.cpu().numpy(), g_ref.cpu().numpy()) <= tol_same
    loss_ref = (r.reshape(batch, -1).double() ** 2).sum(1)
    assert rel_l2(loss_all.double().cpu().numpy(), loss_ref.cpu().numpy()) <= (1e-5 if prec == "f32" else 1e-12)
    # FP64 ground truth
    xs = x if prec == "f64" else x.astype(np.float32).astype(np.float64)
    ys = y if prec == "f64" else y.astype(np.float32).astype(np.float64)
    res = A(xs) - ys
    want_g = 2.0 * A(res)
    want_l = (res.reshape(batch, -1) ** 2).sum(1)
    got_g = g_all.cpu().numpy().astype(np.float64).reshape(batch, rows, cols)
    assert rel_l2(got_g, want_g) <= (3e-5 if prec == "f32" else 1e-12), rel_l2(got_g, want_g)
    assert rel_l2(loss_all.cpu().numpy().astype(np.float64), want_l) <= (1e-5 if prec == "f32" else 1e-12)
    # central differences of the FP64 loss on a few pixels
    eps = 1e-3
    for flat in rng.choice(batch * rows * cols, size=8, replace=False):
        b, rem = divmod(int(flat), rows * cols)
        i, j = divmod(rem, cols)
        e = np.zeros((rows, cols))
        e[i, j] = eps
        lp = float(np.sum((A(xs[b] + e) - ys[b]) ** 2))
        lm = float(np.sum((A(xs[b] - e) - ys[b]) ** 2))
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - got_g[b, i, j]) <= (1e-4 if prec == "f32" else 1e-6) * max(1.0, np.abs(want_g).max())
