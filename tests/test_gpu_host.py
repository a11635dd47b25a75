"""GPU: the host-staged calls (host slices in and out, the path a Go binding
uses for Plan.FFT, RealPlan.RFFT/IRFFT, FFT2/RFFT2, PSD and Spectrogram).
Each call stages through the plan's pinned buffers on a stream of its own and
must give the same bits as the device-resident call on the same input; one
plan is driven from four threads at once (plan.go:12-14, realplan.go:10-11:
plans are safe for concurrent use)."""
import ctypes
import threading

import numpy as np
import pytest

from conftest import cgauss, rgauss

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import _lib as L
    from fft_amd import device as D
    return D, L, L.lib(), torch


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def dev_of(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_c2c_host_matches_device(env, prec):
    # 4096 x 256 complex128 (config 1's shape): 16 MiB, several staging chunks
    D, L, lib, torch = env
    n, batch = 4096, 256
    cdt = np.complex128 if prec == "c128" else np.complex64
    x = cgauss(n * batch, seed=1).astype(cdt)
    p = D.C2CPlan(n, batch, torch.complex128 if prec == "c128" else torch.complex64)
    for direction in (L.FORWARD, L.INVERSE):
        out = np.zeros_like(x)
        L.check(lib.gofft_exec_c2c_host(p.handle, vp(x), vp(out), direction))
        want = p.exec(dev_of(torch, x), inverse=direction == L.INVERSE).cpu().numpy()
        assert np.array_equal(out, want), direction
    # dst may alias src on the host path
    y = x.copy()
    L.check(lib.gofft_exec_c2c_host(p.handle, vp(y), vp(y), L.FORWARD))
    assert np.array_equal(y, p.exec(dev_of(torch, x)).cpu().numpy())


def test_real_host_forms_match_device(env):
    D, L, lib, torch = env
    # RFFT / IRFFT (with a short spectrum)
    n, batch = 6000, 5
    x = rgauss(n * batch, seed=2)
    p = D.R2CPlan(n, batch, torch.float64)
    X = np.zeros(batch * (n // 2 + 1), dtype=np.complex128)
    L.check(lib.gofft_exec_r2c_host(p.handle, vp(x), vp(X)))
    assert np.array_equal(X, p.forward(dev_of(torch, x)).cpu().numpy())
    bins = 1000
    Xs = np.ascontiguousarray(X.reshape(batch, -1)[:, :bins]).ravel()
    back = np.zeros(batch * n)
    L.check(lib.gofft_exec_c2r_host(p.handle, vp(Xs), bins, vp(back)))
    assert np.array_equal(back, p.inverse(dev_of(torch, Xs), bins=bins).cpu().numpy())
    # PSD of a long line (several chunks)
    m = 1 << 22
    z = rgauss(m, seed=3).astype(np.float32)
    ps = D.PSDPlan(m, 1, 0.25, torch.float32)
    out = np.zeros(m // 2 + 1, dtype=np.float32)
    L.check(lib.gofft_exec_psd_host(ps.handle, vp(z), vp(out)))
    assert np.array_equal(out, ps.exec(dev_of(torch, z)).cpu().numpy())
    # Spectrogram
    from oracle import oracle as O
    sig = rgauss(1 << 20, seed=4)
    win = O.window("hann", 512)
    sp = D.SpectrogramPlan(len(sig), 512, 384, list(win), 1.0, torch.float64)
    so = np.zeros(sp.frames * sp.bins)
    L.check(lib.gofft_exec_spectrogram_host(sp.handle, vp(sig), vp(so)))
    assert np.array_equal(so, sp.exec(dev_of(torch, sig)).cpu().numpy())
    # RFFT2 / IRFFT2 and the RFFT2 VJP
    rows, cols = 384, 1000
    img = np.random.default_rng(5).uniform(0, 1, rows * cols).astype(np.float32)
    nd = D.NDPlan((rows, cols), L.ND_R2C, 1, torch.complex64)
    S = np.zeros(rows * (cols // 2 + 1), dtype=np.complex64)
    L.check(lib.gofft_exec_nd_r2c_host(nd.handle, vp(img), 0, vp(S), L.FORWARD))
    Sd = nd.rfft2(dev_of(torch, img))
    assert np.array_equal(S, Sd.cpu().numpy())
    im2 = np.zeros_like(img)
    L.check(lib.gofft_exec_nd_r2c_host(nd.handle, vp(S), len(S), vp(im2), L.INVERSE))
    assert np.array_equal(im2, nd.irfft2(Sd).cpu().numpy())
    g = np.zeros_like(img)
    L.check(lib.gofft_vjp_nd_r2c_host(nd.handle, vp(S), vp(g), L.FORWARD))
    assert np.array_equal(g, nd.vjp(Sd, torch.empty(rows * cols, dtype=torch.float32, device="cuda"),
                                    direction=0).cpu().numpy())


def test_fft2_host(env):
    D, L, lib, torch = env
    shape = (512, 768)
    x = cgauss(512 * 768, seed=6)
    nd = D.NDPlan(shape, L.ND_C2C, 1, torch.complex128)
    out = np.zeros_like(x)
    L.check(lib.gofft_exec_c2c_host(nd.handle, vp(x), vp(out), L.FORWARD))
    assert np.array_equal(out, nd.exec(dev_of(torch, x)).cpu().numpy())


@pytest.mark.parametrize("kind", ["c2c", "r2c", "many"])
def test_host_calls_from_four_threads(env, kind):
    D, L, lib, torch = env
    T, ROUNDS = 4, 6
    if kind == "c2c":
        n, batch = 4096, 64
        p = D.C2CPlan(n, batch, torch.complex128)
        xs = [cgauss(n * batch, seed=20 + t) for t in range(T)]
        outs = [np.zeros_like(v) for v in xs]
        call = lambda t: lib.gofft_exec_c2c_host(p.handle, vp(xs[t]), vp(outs[t]), L.FORWARD)
        want = [p.exec(dev_of(torch, v)).cpu().numpy() for v in xs]
    elif kind == "r2c":
        n, batch = 20014, 16  # Bluestein half length: multi-kernel real path with workspace
        p = D.R2CPlan(n, batch, torch.float32)
        xs = [rgauss(n * batch, seed=30 + t).astype(np.float32) for t in range(T)]
        outs = [np.zeros(batch * (n // 2 + 1), dtype=np.complex64) for _ in range(T)]
        call = lambda t: lib.gofft_exec_r2c_host(p.handle, vp(xs[t]), vp(outs[t]))
        want = [p.forward(dev_of(torch, v)).cpu().numpy() for v in xs]
    else:
        lengths = [1024, 10007, 20014, 65536, 1009, 6720] * 4
        p = D.ManyPlan(lengths, torch.complex64)
        xs = [cgauss(sum(lengths), seed=40 + t).astype(np.complex64) for t in range(T)]
        outs = [np.zeros_like(v) for v in xs]
        call = lambda t: lib.gofft_exec_c2c_host(p.handle, vp(xs[t]), vp(outs[t]), L.FORWARD)
        want = [p.exec(dev_of(torch, v)).cpu().numpy() for v in xs]
    torch.cuda.synchronize()
    errors = []

    def worker(t):
        for r in range(ROUNDS):
            outs[t][...] = 0
            rc = call(t)
            if rc != L.OK:
                errors.append((t, r, rc, L.last_error()))
                return
            if not np.array_equal(outs[t], want[t]):
                errors.append((t, r, "mismatch"))
                return

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(T)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors, errors
