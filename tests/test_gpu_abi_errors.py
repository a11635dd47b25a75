"""GPU: argument checks of the C ABI's transform calls, driven through the raw
entry points. A NULL data pointer of a plan with data, a direction other than
GOFFT_FORWARD / GOFFT_INVERSE and a negative bin count return
GOFFT_ERR_INVALID_ARG with a message, before any work is enqueued; the same
plans then still transform correctly. Empty plans accept NULL pointers, and
an inverse given zero bins reads nothing from its input."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import _lib as L
    from fft_amd import device as D
    return D, L, L.lib(), torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _rejects(L, rc, what):
    assert rc == L.ERR_INVALID_ARG, (rc, L.last_error())
    assert what in L.last_error(), L.last_error()


@pytest.mark.parametrize("dt", ["c64", "c128"])
def test_c2c_calls_reject_bad_arguments(env, dt):
    D, L, lib, torch = env
    ct = torch.complex64 if dt == "c64" else torch.complex128
    n, batch = 1024, 2
    p = D.C2CPlan(n, batch, ct)
    x = torch.randn(n * batch, dtype=ct, device="cuda")
    y = torch.empty_like(x)
    h, s = p.handle, None
    _rejects(L, lib.gofft_exec_c2c(h, None, _ptr(y), L.FORWARD, s), "null data pointer")
    _rejects(L, lib.gofft_exec_c2c(h, _ptr(x), None, L.INVERSE, s), "null data pointer")
    _rejects(L, lib.gofft_exec_c2c(h, _ptr(x), _ptr(y), 7, s), "direction")
    _rejects(L, lib.gofft_vjp_c2c(h, _ptr(x), _ptr(y), -1, s), "direction")
    _rejects(L, lib.gofft_vjp_c2c(h, _ptr(x), None, L.FORWARD, s), "null data pointer")
    xh = np.zeros(n * batch, dtype=np.complex64 if dt == "c64" else np.complex128)
    _rejects(L, lib.gofft_exec_c2c_host(h, None, xh.ctypes.data_as(ctypes.c_void_p), L.FORWARD), "null data")
    _rejects(L, lib.gofft_exec_c2c_host(h, xh.ctypes.data_as(ctypes.c_void_p),
                                        xh.ctypes.data_as(ctypes.c_void_p), 3), "direction")
    _rejects(L, lib.gofft_vjp_c2c_host(h, xh.ctypes.data_as(ctypes.c_void_p), None, L.FORWARD), "null data")
    g = ctypes.c_void_p()
    st = torch.cuda.Stream()
    _rejects(L, lib.gofft_graph_c2c(ctypes.byref(g), h, None, _ptr(y), L.FORWARD,
                                    ctypes.c_void_p(st.cuda_stream)), "null data pointer")
    assert not g.value
    # the plan still works, on the same stream as before
    torch.cuda.synchronize()
    got = p.exec(x)
    want = torch.fft.fft(x.reshape(batch, n).to(torch.complex128)).reshape(-1)
    e = ((got.to(torch.complex128) - want).norm() / want.norm()).item()
    assert e <= (1e-6 if dt == "c64" else 1e-13), e


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_real_calls_reject_bad_arguments(env, dt):
    D, L, lib, torch = env
    rt = torch.float32 if dt == "f32" else torch.float64
    n = 4096
    p = D.R2CPlan(n, 1, rt)
    x = torch.randn(n, dtype=rt, device="cuda")
    X = p.forward(x)
    y = torch.empty_like(x)
    h = p.handle
    _rejects(L, lib.gofft_exec_r2c(h, None, _ptr(X), None), "null data pointer")
    _rejects(L, lib.gofft_exec_r2c(h, _ptr(x), None, None), "null data pointer")
    _rejects(L, lib.gofft_exec_c2r(h, _ptr(X), n // 2 + 1, None, None), "null data pointer")
    _rejects(L, lib.gofft_exec_c2r(h, _ptr(X), -1, _ptr(y), None), "negative bin count")
    _rejects(L, lib.gofft_exec_c2r_host(h, None, -5, None), "negative bin count")
    _rejects(L, lib.gofft_vjp_r2c(h, None, _ptr(y), None), "null data pointer")
    _rejects(L, lib.gofft_vjp_c2r(h, _ptr(x), None, None), "null data pointer")
    # zero bins: nothing is read, the output is the inverse of an all-zero spectrum
    y.fill_(1.0)
    assert lib.gofft_exec_c2r(h, None, 0, _ptr(y), None) == L.OK, L.last_error()
    torch.cuda.synchronize()
    assert torch.count_nonzero(y).item() == 0
    back = p.inverse(X)
    assert ((back - x).norm() / x.norm()).item() <= (1e-6 if dt == "f32" else 1e-13)


def test_nd_psd_and_spectrogram_calls_reject_bad_arguments(env):
    D, L, lib, torch = env
    nd = D.NDPlan((64, 96), L.ND_R2C, 1, torch.complex64)
    x = torch.rand(64 * 96, dtype=torch.float32, device="cuda")
    X = nd.rfft2(x)
    y = torch.empty_like(x)
    _rejects(L, lib.gofft_exec_nd_r2c(nd.handle, _ptr(x), 0, _ptr(X), 5, None), "direction")
    _rejects(L, lib.gofft_exec_nd_r2c(nd.handle, _ptr(X), -1, _ptr(y), L.INVERSE, None), "negative bin count")
    _rejects(L, lib.gofft_exec_nd_r2c(nd.handle, None, 0, _ptr(X), L.FORWARD, None), "null data pointer")
    _rejects(L, lib.gofft_vjp_nd_r2c(nd.handle, _ptr(X), None, L.FORWARD, None), "null data pointer")
    ps = D.PSDPlan(1024, 1, 1.0, torch.float32)
    z = torch.randn(1024, dtype=torch.float32, device="cuda")
    _rejects(L, lib.gofft_exec_psd(ps.handle, None, _ptr(z), None), "null data pointer")
    win = [1.0] * 64
    sp = D.SpectrogramPlan(1024, 64, 32, win, 1.0, torch.float32)
    _rejects(L, lib.gofft_exec_spectrogram(sp.handle, _ptr(z), None, None), "null data pointer")
    back = nd.irfft2(X)
    assert ((back - x).norm() / x.norm()).item() <= 1e-6


def test_empty_plans_accept_null_pointers(env):
    D, L, lib, torch = env
    for n, batch in ((0, 4), (16, 0)):
        p = D.C2CPlan(n, batch, torch.complex64)
        assert lib.gofft_exec_c2c(p.handle, None, None, L.FORWARD, None) == L.OK, L.last_error()
    r = D.R2CPlan(0, 1, torch.float64)
    assert lib.gofft_exec_r2c(r.handle, None, None, None) == L.OK, L.last_error()


def test_real_calls_reject_overlapping_buffers(env):
    # Real transforms (and their VJPs) do not run in place: a batched call
    # writes line l's output where line l+1's input lies. Overlapping input
    # and output ranges return GOFFT_ERR_INVALID_ARG; the complex transforms
    # keep accepting dst == src.
    D, L, lib, torch = env
    n, batch = 1024, 4
    p = D.R2CPlan(n, batch, torch.float64)
    buf = torch.zeros(batch * (n + 2), dtype=torch.float64, device="cuda")
    h = p.handle
    msg = "must not overlap"
    _rejects(L, lib.gofft_exec_r2c(h, _ptr(buf), _ptr(buf), None), msg)
    _rejects(L, lib.gofft_exec_c2r(h, _ptr(buf), n // 2 + 1, _ptr(buf), None), msg)
    _rejects(L, lib.gofft_vjp_r2c(h, _ptr(buf), _ptr(buf), None), msg)
    _rejects(L, lib.gofft_vjp_c2r(h, _ptr(buf), _ptr(buf), None), msg)
    # a partial overlap: output starting in the middle of the input
    _rejects(L, lib.gofft_exec_r2c(h, _ptr(buf), ctypes.c_void_p(buf.data_ptr() + 8 * n), None), msg)
    ps = D.PSDPlan(n, batch, 1.0, torch.float64)
    _rejects(L, lib.gofft_exec_psd(ps.handle, _ptr(buf), _ptr(buf), None), msg)
    sp = D.SpectrogramPlan(batch * n, 256, 128, [1.0] * 256, 1.0, torch.float64)
    _rejects(L, lib.gofft_exec_spectrogram(sp.handle, _ptr(buf), _ptr(buf), None), msg)
    nd = D.NDPlan((32, 64), L.ND_R2C, 1, torch.complex128)
    _rejects(L, lib.gofft_exec_nd_r2c(nd.handle, _ptr(buf), 0, _ptr(buf), L.FORWARD, None), msg)
    _rejects(L, lib.gofft_exec_nd_r2c(nd.handle, _ptr(buf), 32 * 33, _ptr(buf), L.INVERSE, None), msg)
    _rejects(L, lib.gofft_vjp_nd_r2c(nd.handle, _ptr(buf), _ptr(buf), L.FORWARD, None), msg)
    # adjacent, non-overlapping ranges are fine
    x = torch.randn(batch * n + batch * (n + 2), dtype=torch.float64, device="cuda")
    X = x[batch * n:].view(torch.complex128)
    assert lib.gofft_exec_r2c(h, _ptr(x), _ptr(X), None) == L.OK, L.last_error()
    torch.cuda.synchronize()
    want = torch.fft.rfft(x[: batch * n].reshape(batch, n)).reshape(-1)
    assert ((X - want).norm() / want.norm()).item() <= 1e-13
    c = D.C2CPlan(n, batch, torch.complex128)
    z = torch.randn(batch * n, dtype=torch.complex128, device="cuda")
    want = torch.fft.fft(z.reshape(batch, n)).reshape(-1)
    assert lib.gofft_exec_c2c(c.handle, _ptr(z), _ptr(z), L.FORWARD, None) == L.OK
    torch.cuda.synchronize()
    assert ((z - want).norm() / want.norm()).item() <= 1e-13
