"""GPU: complex transforms with the output aliasing the input (plan.go:88-90:
dst may alias src), for every Rader form: the fused single-kernel convolution
(whose input loads all end before its first barrier), and the three-launch
chains, whose last launch must not read the caller's input again (x0 is kept
from the first launch). N-D axes after the first run in place on dst, so
FFTN/FFT2 with a Rader last axis take the same path. Each result is compared
with the oracle."""
import math

import numpy as np
import pytest

from conftest import cgauss, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    from fft_amd import _lib as L
    return D, L, torch


def bar(n, prec):
    return 1e-14 * math.log2(n) if prec == "c128" else 1e-6


@pytest.mark.parametrize("n,batch,prec", [
    (10007, 585, "c128"),   # three-launch chain
    (10007, 585, "c64"),    # fused wide convolution (24 points per thread)
    (5003, 64, "c128"),
    (1009, 64, "c64"),      # fused, several lines per workgroup
    (600011, 2, "c128"),    # chain of multi-pass transforms
    (600011, 2, "c64"),
    (20014, 16, "c64"),     # Bluestein chain
    (65536, 4, "c128"),     # two-pass Stockham
])
def test_c2c_in_place(dev, oracle, n, batch, prec):
    D, L, torch = dev
    dt = torch.complex128 if prec == "c128" else torch.complex64
    x = cgauss(n * batch, seed=n + batch).reshape(batch, n)
    xs = x if prec == "c128" else x.astype(np.complex64).astype(np.complex128)
    p = D.C2CPlan(n, batch, dt)
    for inverse in (False, True):
        t = torch.from_numpy(x.ravel()).to("cuda", dt)
        out = p.exec(t, out=t, inverse=inverse)
        assert out.data_ptr() == t.data_ptr()
        got = t.cpu().numpy().astype(np.complex128).reshape(batch, n)
        want = oracle.fft_batch(xs, inverse=inverse, threads=16)
        assert rel_l2(got, want) <= bar(n, prec), (n, prec, inverse, rel_l2(got, want))


@pytest.mark.parametrize("shape,prec", [((256, 10007), "c128"), ((8, 600011), "c128"), ((8, 600011), "c64"),
                                        ((64, 5003), "c128"), ((32, 10007), "c64")])
def test_fftn_rader_last_axis(dev, oracle, shape, prec):
    D, L, torch = dev
    dt = torch.complex128 if prec == "c128" else torch.complex64
    n = int(np.prod(shape))
    x = cgauss(n, seed=7)
    xs = x if prec == "c128" else x.astype(np.complex64).astype(np.complex128)
    p = D.NDPlan(shape, L.ND_C2C, 1, dt)
    oracle.set_workers(16)
    for inverse in (False, True):
        t = torch.from_numpy(x).to("cuda", dt)
        got = p.exec(t, inverse=inverse).cpu().numpy().astype(np.complex128)  # out of place: axis 1 in place on dst
        want = oracle.fftn(xs, shape, inverse=inverse)
        assert rel_l2(got, want) <= bar(n, prec), (shape, prec, inverse, rel_l2(got, want))
        t2 = torch.from_numpy(x).to("cuda", dt)
        p.exec(t2, out=t2, inverse=inverse)  # fully in place
        assert rel_l2(t2.cpu().numpy().astype(np.complex128), want) <= bar(n, prec)


def test_mixed_plan_in_place(dev, oracle):
    D, L, torch = dev
    lengths = [10007, 1024, 1009, 20014, 600011, 10007, 6720, 5003] * 2
    x = cgauss(sum(lengths), seed=11)
    for dt, prec in ((torch.complex128, "c128"), (torch.complex64, "c64")):
        p = D.ManyPlan(lengths, dt)
        t = torch.from_numpy(x).to("cuda", dt)
        p.exec(t, out=t)
        got = t.cpu().numpy().astype(np.complex128)
        off = 0
        for n in lengths:
            seg = x[off: off + n] if prec == "c128" else x[off: off + n].astype(np.complex64).astype(np.complex128)
            assert rel_l2(got[off: off + n], oracle.fft(seg)) <= bar(n, prec), (prec, n)
            off += n
