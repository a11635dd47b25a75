"""GPU: one plan executed from several host threads at once, each on its own
stream (the C ABI's contract: plans are immutable, with one workspace per
stream). Every thread checks each of its results bit for bit against the same
plan's single-threaded result on the same input."""
import threading

import numpy as np
import pytest

from conftest import cgauss

pytestmark = pytest.mark.gpu

THREADS = 4
ROUNDS = 12


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


def _hammer(torch, run, make_input):
    """run(x, stream) -> output tensor. Thread t owns input t and its own
    stream; the expected outputs come from one thread first."""
    xs = [make_input(t) for t in range(THREADS)]
    want = [run(x, torch.cuda.current_stream()).clone() for x in xs]
    torch.cuda.synchronize()
    errors = []

    def worker(t):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for r in range(ROUNDS):
                    got = run(xs[t], s)
                    s.synchronize()
                    if not torch.equal(got, want[t]):
                        errors.append((t, r))
                        return
        except Exception as e:  # surfaced in the main thread
            errors.append((t, repr(e)))

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(THREADS)]
    for th in ths:
        th.start()
    for th in ths:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in ths), "a worker did not finish"
    torch.cuda.synchronize()
    assert not errors, errors


@pytest.mark.parametrize("n,batch", [(20014, 6), (65536, 4), (10007, 5), (4096, 64)])
@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_c2c_plan_shared_by_threads(dev, n, batch, prec):
    D, torch = dev
    dt = torch.complex64 if prec == "c64" else torch.complex128
    p = D.C2CPlan(n, batch, dt)

    def make(t):
        return torch.from_numpy(cgauss(n * batch, seed=100 + t)).to("cuda", dt)
    _hammer(torch, lambda x, s: p.exec(x, stream=s), make)


def test_mixed_plan_shared_by_threads(dev):
    # a mixed-length plan also forks its groups onto its own lane streams
    D, torch = dev
    lengths = [1024, 4096, 65536, 6720, 1009, 10007, 20014] * 2
    total = sum(lengths)
    for dt in (torch.complex64, torch.complex128):
        p = D.ManyPlan(lengths, dt)

        def make(t):
            return torch.from_numpy(cgauss(total, seed=200 + t)).to("cuda", dt)
        _hammer(torch, lambda x, s: p.exec(x, stream=s), make)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_real_and_nd_plans_shared_by_threads(dev, prec):
    D, torch = dev
    rt = torch.float32 if prec == "f32" else torch.float64
    ct = torch.complex64 if prec == "f32" else torch.complex128
    n = 1 << 18  # multi-pass half-length transform with fused untangle / pack
    rp = D.R2CPlan(n, 2, rt)

    def make_real(t):
        return torch.from_numpy(np.random.default_rng(300 + t).standard_normal(2 * n)).to("cuda", rt)
    _hammer(torch, lambda x, s: rp.forward(x, stream=s), make_real)
    X = rp.forward(make_real(0))
    torch.cuda.synchronize()
    _hammer(torch, lambda x, s: rp.inverse(x, stream=s), lambda t: X * (t + 1))
    nd = D.NDPlan((512, 768), D.L.ND_C2C, 1, ct)

    def make_nd(t):
        return torch.from_numpy(cgauss(512 * 768, seed=400 + t)).to("cuda", ct)
    _hammer(torch, lambda x, s: nd.exec(x, stream=s), make_nd)
