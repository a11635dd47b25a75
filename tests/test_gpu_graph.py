"""GPU: captured execs (gofft_graph_c2c) replay to the same bits as eager
execs. Each graph is replayed several times, with eager work, changed inputs
and synchronisations between replays, and every replay is checked."""
import numpy as np
import pytest

from conftest import cgauss

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    return D, torch


def _replays(D, torch, plan, n_total, dtype, inverse=False):
    s = torch.cuda.Stream()
    x = torch.empty(n_total, dtype=dtype, device="cuda")
    y = torch.empty_like(x)
    with torch.cuda.stream(s):
        x.copy_(torch.from_numpy(cgauss(n_total, seed=1).astype(np.complex128)).to(dtype))
        plan.exec(x, out=y, inverse=inverse, stream=s)  # warm-up: workspace for stream s
        g = D.Graph(plan, x, y, inverse=inverse, stream=s)
    for rep in range(4):
        fresh = torch.from_numpy(cgauss(n_total, seed=10 + rep)).to("cuda", dtype)
        with torch.cuda.stream(s):
            x.copy_(fresh)
            g.launch(stream=s)
        s.synchronize()
        got = y.clone()
        want = plan.exec(fresh.clone(), inverse=inverse)  # eager, on the current stream
        torch.cuda.synchronize()
        assert torch.equal(got, want), rep
        # unrelated eager work between replays
        plan.exec(torch.zeros_like(x))
        torch.cuda.synchronize()


def test_graph_mixed_lengths(dev):
    D, torch = dev
    lengths = [1024, 4096, 65536, 6720, 1009, 10007, 20014] * 3
    for dt in (torch.complex64, torch.complex128):
        p = D.ManyPlan(lengths, dt)
        _replays(D, torch, p, sum(lengths), dt)
        _replays(D, torch, p, sum(lengths), dt, inverse=True)


@pytest.mark.parametrize("n,batch", [(65536, 8), (20014, 4), (4096, 64), (10007, 3)])
def test_graph_c2c(dev, n, batch):
    D, torch = dev
    for dt in (torch.complex64, torch.complex128):
        p = D.C2CPlan(n, batch, dt)
        _replays(D, torch, p, n * batch, dt)


def test_graph_nd(dev):
    D, torch = dev
    p = D.NDPlan((64, 48, 20), dtype=torch.complex128)
    _replays(D, torch, p, 64 * 48 * 20, torch.complex128)
    _replays(D, torch, p, 64 * 48 * 20, torch.complex128, inverse=True)
