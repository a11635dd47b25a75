"""GPU: gofft_exec_c2c_noalloc (the complex transform with its workspace in
the library's static device scratch, no allocation) gives the same bits as
gofft_exec_c2c on the same plan and input, over two calls in a row, the
second with a smaller input, and rejects a plan whose workspace exceeds the
scratch with a clear error."""
import ctypes

import numpy as np
import pytest

from conftest import cgauss

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


def _both(D, L, lib, torch, plan, x, inverse):
    """(noalloc output, exec_c2c output) of one call each on the current stream."""
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    d = L.INVERSE if inverse else L.FORWARD
    y0 = torch.full_like(x, complex(float("nan"), float("nan")))
    y1 = torch.empty_like(x)
    L.check(lib.gofft_exec_c2c_noalloc(plan.handle, _ptr(x), _ptr(y0), d, s))
    L.check(lib.gofft_exec_c2c(plan.handle, _ptr(x), _ptr(y1), d, s))
    torch.cuda.synchronize()
    return y0, y1


def test_noalloc_matches_exec_two_calls(env):
    D, L, lib, torch = env
    # first call: a mixed-length plan with multi-pass, chain and fused groups
    # (workspace on every lane); second: a smaller two-pass plan
    lengths = [65536] * 3 + [20014] * 4 + [10007] * 3 + [1024] * 8 + [6720] * 2
    big = D.ManyPlan(lengths, torch.complex64)
    assert big.workspace_bytes() > 0
    xa = torch.from_numpy(cgauss(sum(lengths), seed=1).astype(np.complex64)).to("cuda")
    small = D.C2CPlan(65536, 2, torch.complex128)
    assert 0 < small.workspace_bytes() < big.workspace_bytes()
    xb = torch.from_numpy(cgauss(65536 * 2, seed=2)).to("cuda")
    assert xb.numel() < xa.numel()
    for inverse in (False, True):
        y0, y1 = _both(D, L, lib, torch, big, xa, inverse)
        assert torch.equal(y0, y1), ("mixed", inverse)
        y0, y1 = _both(D, L, lib, torch, small, xb, inverse)
        assert torch.equal(y0, y1), ("small", inverse)
    # in place, as gofft_exec_c2c allows
    t = xb.clone()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(lib.gofft_exec_c2c_noalloc(small.handle, _ptr(t), _ptr(t), L.FORWARD, s))
    torch.cuda.synchronize()
    assert torch.equal(t, small.exec(xb))


def test_noalloc_rejects_plans_beyond_the_scratch(env):
    D, L, lib, torch = env
    p = D.C2CPlan(65536, 1024, torch.complex128)  # 1 GiB of workspace
    assert p.workspace_bytes() > (64 << 20)
    x = torch.zeros(65536 * 1024, dtype=torch.complex128, device="cuda")
    rc = lib.gofft_exec_c2c_noalloc(p.handle, _ptr(x), _ptr(x), L.FORWARD, None)
    assert rc == L.ERR_UNSUPPORTED, rc
    assert "bytes of scratch" in L.last_error(), L.last_error()
    # a plan without workspace runs (nothing to place in the scratch)
    q = D.C2CPlan(4096, 256, torch.complex128)
    assert q.workspace_bytes() == 0
    z = torch.from_numpy(cgauss(4096 * 256, seed=3)).to("cuda")
    y0, y1 = _both(D, L, lib, torch, q, z, False)
    assert torch.equal(y0, y1)
