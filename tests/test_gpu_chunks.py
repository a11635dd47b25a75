"""GPU parity with tiny line chunks: multi-kernel transforms split their lines
into many chunks (GOFFT_CHUNK_BYTES), exercising the chunk/offset logic of the
Stockham, Rader and Bluestein paths. Runs in a child process because the chunk
budget is read once per process."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

SCRIPT = r"""
import math, sys
import numpy as np, torch
sys.path.insert(0, %r)
sys.path.insert(0, %r)
from conftest import cgauss, rel_l2
from fft_amd import device as D
from oracle import oracle as O
rng = np.random.default_rng(3)
choices = [65536, 10007, 20014, 6720, 1009]
lengths = [int(v) for v in rng.choice(choices, size=23)]
total = sum(lengths)
x = cgauss(total, seed=21)
for dt, nd in [(torch.complex128, np.complex128), (torch.complex64, np.complex64)]:
    for inv in (False, True):
        p = D.ManyPlan(lengths, dt)
        y = p.exec(torch.from_numpy(x.astype(nd)).to('cuda'), inverse=inv).cpu().numpy()
        off = 0
        for n in lengths:
            want = O.fft(x[off:off + n], inverse=inv)
            e = rel_l2(y[off:off + n], want)
            bar = 1e-14 * math.log2(n) if nd == np.complex128 else 1e-6
            assert e <= bar, (n, dt, inv, e)
            off += n
    # strided axis passes of a long 2-D axis
    shape = [65536, 4]
    xs = cgauss(65536 * 4, seed=5)
    q = D.NDPlan(shape, 0, 1, dt)
    got = q.exec(torch.from_numpy(xs.astype(nd)).to('cuda')).cpu().numpy()
    e = rel_l2(got, O.fftn(xs, shape))
    assert e <= (1e-14 * 18 if nd == np.complex128 else 1e-6), e
    # real transforms whose pack / untangle run in the half-length transform's
    # first / last pass, over several chunks of lines
    rt = torch.float64 if nd == np.complex128 else torch.float32
    n, batch = 1 << 17, 7
    xr = np.random.default_rng(9).standard_normal((batch, n))
    rp = D.R2CPlan(n, batch, rt)
    X = rp.forward(torch.from_numpy(xr.ravel()).to('cuda', rt)).cpu().numpy().reshape(batch, n // 2 + 1)
    xs32 = xr if rt == torch.float64 else xr.astype(np.float32).astype(np.float64)
    want = np.stack([O.rfft(r) for r in xs32])
    bar = 1e-14 * 17 if rt == torch.float64 else 1e-6
    assert rel_l2(X, want) <= bar, ('rfft', rt, rel_l2(X, want))
    back = rp.inverse(torch.from_numpy(want.ravel().astype(nd)).to('cuda')).cpu().numpy().reshape(batch, n)
    wantb = np.stack([O.irfft(w, n) for w in want])
    assert rel_l2(back, wantb) <= bar, ('irfft', rt, rel_l2(back, wantb))
print('chunks ok')
"""


def test_tiny_chunks():
    env = dict(os.environ, GOFFT_CHUNK_BYTES=str(3 << 20))
    code = SCRIPT % (ROOT, os.path.join(ROOT, "tests"))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "chunks ok" in r.stdout
