"""GPU parity across batch sizes.

The planner packs 1 to 16 lines into a workgroup depending on the batch (a
small batch still spreads over every CU), and the pass kernel maps columns to
threads in several forms (plain, hooked, the 16 x 16 middle launch, the
two-stage register form). These cases put single lines, partial last
workgroups and large batches through every kernel family and check them
against the oracle: complex128 within rel-L2 1e-14*log2(n), complex64 within
1e-6.
"""
import math

import numpy as np
import pytest

from conftest import cgauss, rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    (1009, 1), (1009, 7), (1009, 585), (1009, 1501),  # fused Rader convolution, 1..8 lines per workgroup
    (1024, 3), (1024, 2049),  # single pass, several lines per workgroup, partial last one
    (6720, 1), (6720, 700),  # single pass with composite radices
    (10007, 1), (10007, 33),  # Rader: complex64 wide fused form, complex128 convolution chain
    (20000, 1), (20000, 3),  # two passes with partial column groups
    (20014, 1), (20014, 5),  # Bluestein: complex64 chain, complex128 reference twiddles
    (40960, 2),  # two passes 160 x 256 (two-stage register form)
    (65536, 3),
]


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


@pytest.mark.parametrize("n,batch", CASES)
def test_batch_shapes(dev, oracle, n, batch):
    D, torch = dev
    x = cgauss(n * batch, seed=n * 7 + batch).reshape(batch, n)
    want = oracle.fft_batch(x)
    p64 = D.C2CPlan(n, batch, torch.complex128)
    got = p64.exec(torch.from_numpy(x.ravel()).to("cuda")).cpu().numpy().reshape(batch, n)
    assert rel_l2(got, want) <= 1e-14 * max(1.0, math.log2(n)), (n, batch, rel_l2(got, want))
    p32 = D.C2CPlan(n, batch, torch.complex64)
    g32 = p32.exec(torch.from_numpy(x.astype(np.complex64).ravel()).to("cuda")).cpu().numpy()
    e = rel_l2(g32.reshape(batch, n).astype(np.complex128), want)
    assert e <= 1e-6, (n, batch, e)
    # every line individually too, so a wrong tail line cannot hide in the norm
    worst = max(rel_l2(g32.reshape(batch, n)[i].astype(np.complex128), want[i]) for i in range(batch))
    assert worst <= 3e-6, (n, batch, worst)


@pytest.mark.parametrize("n,batch", [(1009, 1501), (20014, 5), (40960, 2), (1024, 2049)])
def test_batch_shapes_inverse(dev, oracle, n, batch):
    D, torch = dev
    x = cgauss(n * batch, seed=n * 11 + batch).reshape(batch, n)
    want = oracle.fft_batch(x, inverse=True)
    p32 = D.C2CPlan(n, batch, torch.complex64)
    g32 = p32.exec(torch.from_numpy(x.astype(np.complex64).ravel()).to("cuda"), inverse=True).cpu().numpy()
    assert rel_l2(g32.reshape(batch, n).astype(np.complex128), want) <= 1e-6
