"""GPU: mixed-length plans run as one persistent work-queue kernel
(GOFFT_FLAG_QUEUE). The queue re-plans each launch's columns per workgroup
but keeps every transform's arithmetic, so its results must equal the
stream path's bit for bit: forward, inverse, in place, both VJPs, a replayed
graph and the host-staged call, at the 4096-transform mixed-length batch's
full size in both precisions. The stream path itself is checked against the
oracle elsewhere; a few lines here are checked against the oracle directly.
The bounded wait's time-out is forced from the host (every wait asks for one
item more than its stage has and may last 1 ms): the kernel ends, the exec
reports the error, and the next plan runs normally."""
import ctypes
import os
import sys

import numpy as np
import pytest

from conftest import cgauss, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import _lib as L
    from fft_amd import device as D
    return D, L, torch


def headline_lengths():
    import bench
    return bench.workload_lengths(0)


def plans(D, L, torch, lengths, dt, extra=0):
    q = D.ManyPlan(lengths, dt, flags=L.FLAG_QUEUE | extra)
    s = D.ManyPlan(lengths, dt, flags=extra)
    return q, s


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_queue_matches_streams_full_size(env, oracle, prec):
    D, L, torch = env
    dt = torch.complex64 if prec == "c64" else torch.complex128
    lengths = headline_lengths()
    q, s = plans(D, L, torch, lengths, dt)
    text = q.describe()
    print(text.splitlines()[0], flush=True)
    assert "queue threads=" in text
    x = torch.from_numpy(cgauss(sum(lengths), seed=3)).to("cuda", dt)
    for inverse in (False, True):
        a = q.exec(x, inverse=inverse)
        b = s.exec(x, inverse=inverse)
        torch.cuda.synchronize()
        assert torch.equal(a, b), ("exec", prec, inverse)
        ga = q.vjp(x, inverse=inverse)
        gb = s.vjp(x, inverse=inverse)
        assert torch.equal(ga, gb), ("vjp", prec, inverse)
    y = x.clone()
    q.exec(y, out=y)
    assert torch.equal(y, s.exec(x)), ("in place", prec)
    # a few lines of every length against the oracle
    xh = x.cpu().numpy().astype(np.complex128)
    got = s.exec(x).cpu().numpy().astype(np.complex128)
    offs = np.concatenate([[0], np.cumsum(lengths)])
    seen = {}
    for i, n in enumerate(lengths):
        if seen.get(n, 0) >= 2:
            continue
        seen[n] = seen.get(n, 0) + 1
        seg = slice(offs[i], offs[i + 1])
        want = oracle.fft(xh[seg])
        bar = 1e-14 * np.log2(n) if prec == "c128" else 1e-6
        assert rel_l2(got[seg], want) <= bar, (n, prec)


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_queue_mixed_with_lane_groups(env, prec):
    # lengths beyond the 4096-transform batch: a multi-pass Rader chain,
    # small Bluestein and trivial groups (the trivial one stays on a lane)
    D, L, torch = env
    dt = torch.complex64 if prec == "c64" else torch.complex128
    lengths = [600011, 17, 1, 0, 1024, 10007, 20014, 6720, 65536, 1009, 3, 5003] * 3
    q, s = plans(D, L, torch, lengths, dt)
    print(q.describe().splitlines()[0], flush=True)
    x = torch.from_numpy(cgauss(sum(lengths), seed=5)).to("cuda", dt)
    for inverse in (False, True):
        assert torch.equal(q.exec(x, inverse=inverse), s.exec(x, inverse=inverse)), (prec, inverse)
    if prec == "c64":
        # FP32 twiddles: the multi-pass groups stay on the lanes beside the queue
        q2, s2 = plans(D, L, torch, lengths, dt, L.FLAG_TWIDDLES_FP32)
        assert "queued:" in q2.describe()
        assert torch.equal(q2.exec(x), s2.exec(x))


def test_queue_graph_and_host(env):
    D, L, torch = env
    lengths = [1024, 4096, 65536, 6720, 1009, 10007, 20014] * 40
    dt = torch.complex64
    q, s = plans(D, L, torch, lengths, dt)
    x = torch.from_numpy(cgauss(sum(lengths), seed=7)).to("cuda", dt)
    want = s.exec(x)
    out = torch.empty_like(x)
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(st):
        g = D.Graph(q, x, out, stream=st)
    for r in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.launch(stream=st)
        torch.cuda.synchronize()
        assert torch.equal(out, want), r
        q.exec(x)  # eager work between replays
    xh = x.cpu().numpy()
    oh = np.zeros_like(xh)
    lib = L.lib()
    L.check(lib.gofft_exec_c2c_host(q.handle, xh.ctypes.data_as(ctypes.c_void_p),
                                    oh.ctypes.data_as(ctypes.c_void_p), L.FORWARD))
    assert np.array_equal(oh, want.cpu().numpy())


def test_queue_wait_time_limit(env, monkeypatch):
    D, L, torch = env
    lengths = [65536, 20014, 1024] * 20
    monkeypatch.setenv("GOFFT_QUEUE_TEST_STALL", "1")
    monkeypatch.setenv("GOFFT_QUEUE_SPIN_TICKS", "100000")  # 1 ms
    monkeypatch.setenv("GOFFT_QUEUE_CHECK", "1")
    q = D.ManyPlan(lengths, torch.complex64, flags=L.FLAG_QUEUE)
    x = torch.from_numpy(cgauss(sum(lengths), seed=9)).to("cuda", torch.complex64)
    out = torch.empty_like(x)
    rc = L.lib().gofft_exec_c2c(q.handle, x.data_ptr(), out.data_ptr(), L.FORWARD, None)
    assert rc == L.ERR_HIP and "time limit" in L.last_error(), (rc, L.last_error())
    torch.cuda.synchronize()
    monkeypatch.delenv("GOFFT_QUEUE_TEST_STALL")
    q2 = D.ManyPlan(lengths, torch.complex64, flags=L.FLAG_QUEUE)
    s2 = D.ManyPlan(lengths, torch.complex64)
    assert torch.equal(q2.exec(x), s2.exec(x))
