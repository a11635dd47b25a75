"""GPU parity at the mixed-length workload's full size: 4096 complex
transforms, lengths drawn uniformly (numpy default_rng(0)) from {1024, 4096,
65536, 6720, 1009, 10007, 20014}, re/im N(0,1) from seed 0 (the inputs bench.py
times), one plan, both precisions. Every transform is checked against the
oracle's Plan.FFT of the same input: complex128 within rel-L2 1e-14*log2(n),
complex64 within 1e-6; the inverse likewise.
"""
import math
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.fixture(scope="module")
def workload(oracle):
    lengths = bench.workload_lengths(0)
    assert len(lengths) == 4096
    x = bench.gaussian_complex(sum(lengths), 0)
    # the oracle per length group, several lines per call (one thread each)
    want = np.zeros_like(x)
    wanti = np.zeros_like(x)
    offs = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    for n in sorted(set(lengths)):
        idx = [i for i, v in enumerate(lengths) if v == n]
        rows = np.stack([x[offs[i]: offs[i] + n] for i in idx])
        y = oracle.fft_batch(rows, threads=16)
        yi = oracle.fft_batch(rows, inverse=True, threads=16)
        for k, i in enumerate(idx):
            want[offs[i]: offs[i] + n] = y[k]
            wanti[offs[i]: offs[i] + n] = yi[k]
    return lengths, offs, x, want, wanti


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_mixed_4096_transforms_vs_oracle(workload, prec):
    import torch
    from fft_amd import device as D
    lengths, offs, x, want, wanti = workload
    dt = torch.complex128 if prec == "c128" else torch.complex64
    plan = D.ManyPlan(lengths, dt)
    xt = torch.from_numpy(x.astype(np.complex128 if prec == "c128" else np.complex64)).to("cuda")
    y = plan.exec(xt).cpu().numpy().astype(np.complex128)
    yi = plan.exec(xt, inverse=True).cpu().numpy().astype(np.complex128)
    worst = {}
    for i, n in enumerate(lengths):
        s = slice(offs[i], offs[i] + n)
        for got, ref in ((y, want), (yi, wanti)):
            e = float(np.linalg.norm(got[s] - ref[s]) / np.linalg.norm(ref[s]))
            worst[n] = max(worst.get(n, 0.0), e)
    for n, e in sorted(worst.items()):
        bar = 1e-14 * math.log2(n) if prec == "c128" else 1e-6
        assert e <= bar, (prec, n, e, bar)
    print(prec, {n: "%.3g" % e for n, e in sorted(worst.items())})
